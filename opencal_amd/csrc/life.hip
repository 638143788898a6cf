// Conway Life device program.
//
// Reference: OpenCAL-OMP/examples/calomp_life-glut/source/life.c:16-27 — the
// sum of the 8 Moore neighbours (calGetX2D, n = 1..8), alive' = sum == 3 ||
// (sum == 2 && alive), written to the next layer of every cell; then the full
// next->current copy (cal2D.c:738-776).  Toroidal.  Integer work: results are
// bit-identical by construction.
//
// Kernels:
//   life_step   one step, one thread per 4 consecutive cells of a row.
//   life_multi  many steps of a grid that fits in one workgroup's LDS as a
//               bit-packed board (two boards of up to 512 Ki cells = 128 KiB
//               of the 160 KiB LDS; columns a multiple of 64): the board is
//               loaded once, stepped with 64-cell bit-sliced adders, and
//               written back once.  Used by calRun2D when no host callback
//               runs between steps; the first step runs per-cell so the board
//               holds only 0/1 values.
#include "engine.hpp"
#include "device_common.hpp"

namespace calhip {

namespace {

template <typename T>
__global__ void life_step(const T* __restrict__ cur, T* __restrict__ nxt, int rows, int cols)
{
	size_t n = (size_t)rows * cols;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		int i = (int)(c / cols), j = (int)(c % cols);
		int in = dev::wrap1(i - 1, rows), is = dev::wrap1(i + 1, rows);
		int jw = dev::wrap1(j - 1, cols), je = dev::wrap1(j + 1, cols);
		const T* rn = cur + (size_t)in * cols;
		const T* rc = cur + (size_t)i * cols;
		const T* rs = cur + (size_t)is * cols;
		// neighbour order of the reference: N, W, E, S, NW, SW, SE, NE
		int sum = (int)rn[j] + (int)rc[jw] + (int)rc[je] + (int)rs[j] + (int)rn[jw] + (int)rs[jw] + (int)rs[je] +
		          (int)rn[je];
		nxt[c] = (sum == 3 || (sum == 2 && rc[j] == 1)) ? (T)1 : (T)0;
	}
}

// ---- LDS-resident bit-packed multi-step ----
constexpr int LM_THREADS = 1024;
constexpr int LM_MAX_WORDS = 16384;   // 64-bit words: 1 Mi cells, 128 KiB

// bit-sliced full adder helpers on 64-cell words
__device__ __forceinline__ void add2(uint64_t a, uint64_t b, uint64_t& s, uint64_t& c)
{
	s = a ^ b;
	c = a & b;
}

__device__ __forceinline__ void add3(uint64_t a, uint64_t b, uint64_t d, uint64_t& s, uint64_t& c)
{
	uint64_t t = a ^ b;
	s = t ^ d;
	c = (a & b) | (t & d);
}

// next state of 64 cells given the 8 neighbour words and the centre word.
// Only the sums 2 and 3 matter: bit0/bit1 of the count and "count >= 4".
__device__ __forceinline__ uint64_t life_word(uint64_t nw, uint64_t n, uint64_t ne, uint64_t w, uint64_t c,
                                              uint64_t e, uint64_t sw, uint64_t s, uint64_t se)
{
	uint64_t s0, c0, s1, c1, s2, c2;
	add3(nw, n, ne, s0, c0);
	add3(w, e, sw, s1, c1);
	add2(s, se, s2, c2);
	// ones: s0 + s1 + s2
	uint64_t o, oc;
	add3(s0, s1, s2, o, oc);
	// twos: c0 + c1 + c2 + oc (each worth 2)
	uint64_t t1, tc1, t2, tc2;
	add3(c0, c1, c2, t1, tc1);
	add2(t1, oc, t2, tc2);
	// count = o + 2*t2 + 4*(tc1 + tc2 + ...) ; >= 4 iff tc1 | tc2
	uint64_t ge4 = tc1 | tc2;
	uint64_t is3 = o & t2 & ~ge4;
	uint64_t is2 = ~o & t2 & ~ge4;
	return is3 | (is2 & c);
}

// Board layout in LDS: rows x wpr words, bit b of word (i, w) = cell (i, 64*w + b).
// Toroidal wrap at the word level needs cols % 64 == 0 (checked on the host).
template <typename T>
__global__ __launch_bounds__(LM_THREADS) void life_multi(T* __restrict__ cur, T* __restrict__ nxt, int rows,
                                                         int cols, int steps)
{
	extern __shared__ uint64_t lds[];
	const int wpr = cols / 64;
	const int words = rows * wpr;
	uint64_t* A = lds;
	uint64_t* B = lds + words;
	for (int w = threadIdx.x; w < words; w += LM_THREADS) {
		int i = w / wpr, q = w % wpr;
		const T* src = cur + (size_t)i * cols + (size_t)q * 64;
		uint64_t v = 0;
		for (int b = 0; b < 64; ++b) v |= (uint64_t)(src[b] == 1) << b;
		A[w] = v;
	}
	__syncthreads();
	for (int s = 0; s < steps; ++s) {
		for (int w = threadIdx.x; w < words; w += LM_THREADS) {
			int i = w / wpr, q = w % wpr;
			int in = i == 0 ? rows - 1 : i - 1, is = i == rows - 1 ? 0 : i + 1;
			int ql = q == 0 ? wpr - 1 : q - 1, qr = q == wpr - 1 ? 0 : q + 1;
			uint64_t cn = A[in * wpr + q], cc = A[i * wpr + q], cs = A[is * wpr + q];
			uint64_t ln = A[in * wpr + ql], lc = A[i * wpr + ql], ls = A[is * wpr + ql];
			uint64_t rn = A[in * wpr + qr], rc = A[i * wpr + qr], rs = A[is * wpr + qr];
			// west neighbour of bit b is bit b-1 (bit 63 of the left word for b = 0)
			uint64_t wn = (cn << 1) | (ln >> 63), wc = (cc << 1) | (lc >> 63), ws = (cs << 1) | (ls >> 63);
			uint64_t en = (cn >> 1) | (rn << 63), ec = (cc >> 1) | (rc << 63), es = (cs >> 1) | (rs << 63);
			B[w] = life_word(wn, cn, en, wc, cc, ec, ws, cs, es);
		}
		__syncthreads();
		uint64_t* t = A;
		A = B;
		B = t;
	}
	for (int w = threadIdx.x; w < words; w += LM_THREADS) {
		int i = w / wpr, q = w % wpr;
		T* dst = cur + (size_t)i * cols + (size_t)q * 64;
		uint64_t v = A[w];
		for (int b = 0; b < 64; ++b) dst[b] = (T)((v >> b) & 1u);
	}
	(void)nxt;
}

}  // namespace

struct LifeProgram : Program {
	SubX* q = nullptr;
	const char* name() const override { return "Life"; }

	int launch_step(ModelX* m)
	{
		size_t n = m->cells();
		int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
		hipEvent_t ev;
		timer_begin(m, &ev);
		if (q->type == 'b')
			hipLaunchKernelGGL(life_step<int8_t>, dim3(grid), dim3(256), 0, m->stream, (const int8_t*)q->d_cur,
			                   (int8_t*)q->d_nxt, m->pub.rows, m->pub.columns);
		else
			hipLaunchKernelGGL(life_step<int32_t>, dim3(grid), dim3(256), 0, m->stream, (const int32_t*)q->d_cur,
			                   (int32_t*)q->d_nxt, m->pub.rows, m->pub.columns);
		timer_end(m, ev);
		m->stats.launches++;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		flip(q);
		return 0;
	}

	int step(ModelX* m) override
	{
		if (q->nxt_dirty) flip(q);
		return launch_step(m);
	}

	bool lds_fits(ModelX* m) const
	{
		return m->pub.columns % 64 == 0 && (size_t)m->pub.rows * (m->pub.columns / 64) * 2 <= (size_t)LM_MAX_WORDS;
	}

	int run_steps(ModelX* m, int n) override
	{
		if (q->nxt_dirty) flip(q);
		if (n <= 1 || !lds_fits(m) || getenv("CALHIP_LIFE_NO_LDS")) {
			for (int k = 0; k < n; ++k)
				if (launch_step(m)) return -1;
			return 0;
		}
		if (launch_step(m)) return -1;   // afterwards every cell is 0 or 1
		n -= 1;
		size_t words = (size_t)m->pub.rows * (m->pub.columns / 64);
		size_t lds = words * 2 * sizeof(uint64_t);
		hipEvent_t ev;
		timer_begin(m, &ev);
		if (q->type == 'b') {
			hipFuncSetAttribute((const void*)life_multi<int8_t>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
			hipLaunchKernelGGL(life_multi<int8_t>, dim3(1), dim3(LM_THREADS), lds, m->stream, (int8_t*)q->d_cur,
			                   (int8_t*)q->d_nxt, m->pub.rows, m->pub.columns, n);
		} else {
			hipFuncSetAttribute((const void*)life_multi<int32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
			hipLaunchKernelGGL(life_multi<int32_t>, dim3(1), dim3(LM_THREADS), lds, m->stream, (int32_t*)q->d_cur,
			                   (int32_t*)q->d_nxt, m->pub.rows, m->pub.columns, n);
		}
		timer_end(m, ev);
		m->stats.launches++;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		q->nxt_stale = true;   // the result is in place in the current layer
		q->nxt_dirty = false;
		return 0;
	}
};

}  // namespace calhip

using namespace calhip;

extern "C" int calhipBindLife2D(struct CALModel2D* ca2D, CALCallbackFunc2D transition, struct CALSubstate2Db* Qb,
                                struct CALSubstate2Di* Qi)
{
	ModelX* m = X(ca2D);
	if (ca2D->X_id != CAL_MOORE_NEIGHBORHOOD_2D || ca2D->T != CAL_SPACE_TOROIDAL ||
	    ca2D->OPTIMIZATION != CAL_NO_OPT) {
		set_error(30, "calhipBindLife2D: the Life program needs a Moore, toroidal, full-grid model");
		return -1;
	}
	if (!transition || (!Qb == !Qi)) {
		set_error(31, "calhipBindLife2D: give the transition and exactly one substate");
		return -1;
	}
	if (ca2D->rows < 3 || ca2D->columns < 3) {
		set_error(32, "calhipBindLife2D: the grid must be at least 3 x 3");
		return -1;
	}
	auto* p = new LifeProgram();
	p->q = Qb ? S(Qb) : S(Qi);
	delete m->prog;
	m->prog = p;
	m->prog_eps.assign(1, transition);
	m->prog_steering = nullptr;
	return 0;
}
