// Conway Life device program.
//
// Reference: OpenCAL-OMP/examples/calomp_life-glut/source/life.c:16-27 — the
// sum of the 8 Moore neighbours (calGetX2D, n = 1..8), alive' = sum == 3 ||
// (sum == 2 && alive), written to the next layer of every cell; then the full
// next->current copy (cal2D.c:738-776).  Toroidal.  Integer work: results are
// bit-identical by construction.
//
// Kernels:
//   life_step   one step, one thread per cell (grid-stride).
//   life_tiles  k steps per launch by temporal blocking (columns a multiple of
//               64): a block holds a bit-packed region of (tw + 2) 64-cell
//               words x (tr + 2k) rows in LDS, advances it k steps with
//               64-cell bit-sliced adders, and writes back its tile of tw
//               words x tr rows.  A cell whose Moore neighbourhood leaves the
//               region is wrong after one step and the error moves one cell
//               per step, so after k <= 64 steps the tile (k rows and one word
//               from every side of the region) is exact.  Used by calRun2D
//               when no host callback runs between steps; the first step
//               runs per cell so the board holds only 0/1 values.
#include <stdio.h>
#include <stdlib.h>

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

// ---- temporal blocking in LDS, bit-packed ----
constexpr int LT_THREADS = 256;
constexpr int LT_TW = 1;     // tile width in 64-cell words
constexpr int LT_TR = 16;    // tile height in rows
constexpr int LT_K = 32;     // steps per launch (<= 64: the side halo is one word)

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
	// count = o + 2*t2 + 4*(tc1 + tc2 + ...); >= 4 iff tc1 | tc2
	uint64_t ge4 = tc1 | tc2;
	uint64_t is3 = o & t2 & ~ge4;
	uint64_t is2 = ~o & t2 & ~ge4;
	return is3 | (is2 & c);
}

__device__ __forceinline__ int wrapm(int x, int n)
{
	x %= n;
	return x < 0 ? x + n : x;
}

// 64 cells of row i starting at column 64*q as a bit word (bit b = cell b).
template <typename T>
__device__ __forceinline__ uint64_t pack64(const T* __restrict__ row)
{
	uint64_t v = 0;
	if (sizeof(T) == 1) {
		const uint4* p = reinterpret_cast<const uint4*>(row);
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			const uint4 x = p[k];
			const unsigned w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
			for (int d = 0; d < 4; ++d)
#pragma unroll
				for (int b = 0; b < 4; ++b)
					v |= (uint64_t)(((w[d] >> (8 * b)) & 0xffu) == 1u) << (16 * k + 4 * d + b);
		}
	} else {
		for (int b = 0; b < 64; ++b) v |= (uint64_t)(row[b] == 1) << b;
	}
	return v;
}

template <typename T>
__device__ __forceinline__ void unpack64(uint64_t v, T* __restrict__ row)
{
	if (sizeof(T) == 1) {
		uint4* p = reinterpret_cast<uint4*>(row);
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			unsigned w[4];
#pragma unroll
			for (int d = 0; d < 4; ++d) {
				unsigned x = 0;
#pragma unroll
				for (int b = 0; b < 4; ++b) x |= (unsigned)((v >> (16 * k + 4 * d + b)) & 1u) << (8 * b);
				w[d] = x;
			}
			p[k] = make_uint4(w[0], w[1], w[2], w[3]);
		}
	} else {
		for (int b = 0; b < 64; ++b) row[b] = (T)((v >> b) & 1u);
	}
}

// k steps of a rows x cols toroidal board (cols % 64 == 0), cur -> out.
// Region word (ri, rq) holds board row i0 + ri and word column q0 + rq, both
// wrapped (the region may be larger than the board: each region cell still
// sees its true toroidal neighbours inside the region).
template <typename T>
__global__ __launch_bounds__(LT_THREADS) void life_tiles(const T* __restrict__ cur, T* __restrict__ out, int rows,
                                                         int cols, int k, int tw, int tr)
{
	extern __shared__ uint64_t lds[];
	const int wpr = cols / 64;
	const int tiles_x = (wpr + tw - 1) / tw;
	const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
	const int RW = tw + 2, RH = tr + 2 * k;
	const int q0 = tx * tw - 1, i0 = ty * tr - k;
	const int words = RW * RH;
	uint64_t* A = lds;
	uint64_t* B = lds + words;
	for (int w = threadIdx.x; w < words; w += LT_THREADS) {
		const int ri = w / RW, rq = w % RW;
		const int i = wrapm(i0 + ri, rows), q = wrapm(q0 + rq, wpr);
		A[w] = pack64(cur + (size_t)i * cols + (size_t)q * 64);
	}
	__syncthreads();
	for (int s = 0; s < k; ++s) {
		for (int w = threadIdx.x; w < words; w += LT_THREADS) {
			const int ri = w / RW, rq = w % RW;
			// neighbours outside the region read as dead
			const bool hn = ri > 0, hs = ri < RH - 1, hl = rq > 0, hr = rq < RW - 1;
			const uint64_t cn = hn ? A[w - RW] : 0, cc = A[w], cs = hs ? A[w + RW] : 0;
			const uint64_t ln = (hn && hl) ? A[w - RW - 1] : 0, lc = hl ? A[w - 1] : 0;
			const uint64_t ls = (hs && hl) ? A[w + RW - 1] : 0;
			const uint64_t rn = (hn && hr) ? A[w - RW + 1] : 0, rc = hr ? A[w + 1] : 0;
			const uint64_t rs = (hs && hr) ? A[w + RW + 1] : 0;
			// west neighbour of bit b is bit b-1 (bit 63 of the left word for b = 0)
			const uint64_t wn = (cn << 1) | (ln >> 63), wc = (cc << 1) | (lc >> 63), ws = (cs << 1) | (ls >> 63);
			const uint64_t en = (cn >> 1) | (rn << 63), ec = (cc >> 1) | (rc << 63), es = (cs >> 1) | (rs << 63);
			B[w] = life_word(wn, cn, en, wc, cc, ec, ws, cs, es);
		}
		__syncthreads();
		uint64_t* t = A;
		A = B;
		B = t;
	}
	for (int w = threadIdx.x; w < tw * tr; w += LT_THREADS) {
		const int ti = w / tw, tq = w % tw;
		const int i = ty * tr + ti, q = tx * tw + tq;
		if (i >= rows || q >= wpr) continue;
		unpack64(A[(k + ti) * RW + 1 + tq], out + (size_t)i * cols + (size_t)q * 64);
	}
}

}  // namespace

struct LifeProgram : Program {
	SubX* q = nullptr;
	const char* name() const override { return "Life"; }

	// one sweep current -> next (no flip)
	int launch_sweep_only(ModelX* m)
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
		return CALHIP_CHECK(hipGetLastError());
	}

	int launch_step(ModelX* m)
	{
		if (launch_sweep_only(m)) return -1;
		flip(q);
		return 0;
	}

	// calApplyElementaryProcess2D (cal2D.c:717-734) with the Life transition:
	// the sweep writes every cell of the next layer; the caller's
	// calUpdate2D (or calUpdateSubstate2D*) flips it in.
	int apply(ModelX* m, const void* ep) override
	{
		if (m->prog_eps.empty() || (const void*)m->prog_eps[0] != ep) {
			set_error(33, "Life device program: calApplyElementaryProcess2D with a process that is not the bound "
			              "transition");
			return -1;
		}
		if (launch_sweep_only(m)) return -1;
		q->nxt_stale = false;
		q->nxt_dirty = true;
		return 0;
	}

	int step(ModelX* m) override
	{
		if (q->nxt_dirty) flip(q);
		return launch_step(m);
	}

	bool tiles_fit(ModelX* m) const { return m->pub.columns % 64 == 0; }

	int run_steps(ModelX* m, int n, bool with_steering) override
	{
		if (with_steering) return Program::run_steps(m, n, true);   // Life has no steering twin
		if (q->nxt_dirty) flip(q);
		if (n <= 1 || !tiles_fit(m) || getenv("CALHIP_LIFE_NO_TILES")) {
			for (int k = 0; k < n; ++k)
				if (launch_step(m)) return -1;
			return 0;
		}
		if (launch_step(m)) return -1;   // afterwards every cell is 0 or 1
		n -= 1;
		const int rows = m->pub.rows, cols = m->pub.columns;
		const int wpr = cols / 64;
		int tw = LT_TW, tr = LT_TR, kmax = LT_K;
		if (const char* e = getenv("CALHIP_LIFE_TILE"))   // "words,rows,steps"
			if (sscanf(e, "%d,%d,%d", &tw, &tr, &kmax) != 3 || tw < 1 || tr < 1 || kmax < 1 || kmax > 64) {
				tw = LT_TW;
				tr = LT_TR;
				kmax = LT_K;
			}
		const int blocks = ((wpr + tw - 1) / tw) * ((rows + tr - 1) / tr);
		while (n > 0) {
			const int k = n < kmax ? n : kmax;
			const size_t lds = (size_t)2 * (tw + 2) * (tr + 2 * k) * sizeof(uint64_t);
			hipEvent_t ev;
			timer_begin(m, &ev);
			if (q->type == 'b')
				hipLaunchKernelGGL(life_tiles<int8_t>, dim3(blocks), dim3(LT_THREADS), lds, m->stream,
				                   (const int8_t*)q->d_cur, (int8_t*)q->d_nxt, rows, cols, k, tw, tr);
			else
				hipLaunchKernelGGL(life_tiles<int32_t>, dim3(blocks), dim3(LT_THREADS), lds, m->stream,
				                   (const int32_t*)q->d_cur, (int32_t*)q->d_nxt, rows, cols, k, tw, tr);
			timer_end(m, ev);
			m->stats.launches++;
			if (CALHIP_CHECK(hipGetLastError())) return -1;
			flip(q);
			n -= k;
		}
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
