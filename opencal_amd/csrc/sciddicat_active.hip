// SciddicaT active cells, K steps per launch (the mode is described in
// sciddicat_active.hpp).
//
// Kernels:
//   act_check   entry check over the grid: counts the set's cells and the
//               cells outside it with h > eps (the mode needs none), and
//               marks the ghosts (cells outside the set with stored flows
//               that are not +0.0);
//   act_setup   makes the second copy of every state buffer equal to the
//               first and lists the tiles within reach of the set;
//   act_fused<K> K steps of the listed tiles (tile + 2K halo in LDS, one
//               cell per thread), lists the tiles of the next launch;
//   act_leave   writes the flows' logical content back into the current
//               layer (zero outside the set except at ghosts; zero
//               everywhere once the set emptied) and finishes the 0.0
//               additions of a run that emptied the set;
//   act_normalise  the latter alone (before a read of h).
// The minimisation is flows_wet_body (sciddicat_device.hpp), operation for
// operation the reference's (sciddicaT.c:36-88); the width sum follows
// sciddicaT.c:91-101 term by term.
#include "sciddicat_active.hpp"

#include "device_common.hpp"
#include "engine.hpp"
#include "sciddicat_device.hpp"

namespace calhip {
namespace actf {

namespace {

using namespace sct;

__device__ __forceinline__ int wrapm(int x, int n)
{
	const int r = x % n;
	return r < 0 ? r + n : r;
}

__device__ __forceinline__ unsigned long long bits_of(double v) { return (unsigned long long)__double_as_longlong(v); }

// Tiles a set cell at tile-local (la, lb) of a tile of tre x tce cells must
// list: bit (dr + 1) * 3 + (dc + 1) for the tile at offset (dr, dc); the
// tile itself and those within km cells.
__device__ __forceinline__ unsigned reach_marks(int la, int lb, int tre, int tce, int km)
{
	const bool n = la < km, s = la >= tre - km, w = lb < km, e = lb >= tce - km;
	unsigned mk = 1u << 4;
	if (n) mk |= 1u << 1;
	if (s) mk |= 1u << 7;
	if (w) mk |= 1u << 3;
	if (e) mk |= 1u << 5;
	if (n && w) mk |= 1u << 0;
	if (n && e) mk |= 1u << 2;
	if (s && w) mk |= 1u << 6;
	if (s && e) mk |= 1u << 8;
	return mk;
}

// Thread q < 9 of a block lists the tile at offset (q / 3 - 1, q % 3 - 1)
// from tile (ti, tj) when bit q of marks is set and the tile is not listed
// yet for this epoch.
__device__ __forceinline__ void list_marked(unsigned marks, int q, int ti, int tj, int ty, int tx, unsigned* tag,
                                            unsigned e1, int* list, int* ctr)
{
	if (q >= 9 || !((marks >> q) & 1u)) return;
	const int ui = wrapm(ti + q / 3 - 1, ty), uj = wrapm(tj + q % 3 - 1, tx);
	const int u = ui + uj * ty;
	if (atomicMax(tag + u, e1) < e1) list[atomicAdd(ctr, 1)] = u;
}

__global__ __launch_bounds__(256) void act_check(const double* __restrict__ h, const uint8_t* __restrict__ fl,
                                                 const double* __restrict__ F0, const double* __restrict__ F1,
                                                 const double* __restrict__ F2, const double* __restrict__ F3,
                                                 int f_zero, uint8_t* __restrict__ gh, int* chk, size_t n,
                                                 double eps)
{
	__shared__ int sh[2];
	if (threadIdx.x == 0) sh[0] = sh[1] = 0;
	__syncthreads();
	int viol = 0, cnt = 0;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		const bool in = fl[c] != 0;
		bool g = false;
		if (!in && !f_zero)
			g = (bits_of(F0[c]) | bits_of(F1[c]) | bits_of(F2[c]) | bits_of(F3[c])) != 0ull;
		gh[c] = g ? 1 : 0;
		viol += (!in && !(h[c] <= eps)) ? 1 : 0;
		cnt += in ? 1 : 0;
	}
	if (viol) atomicAdd(sh, viol);
	if (cnt) atomicAdd(sh + 1, cnt);
	__syncthreads();
	if (threadIdx.x == 0) {
		if (sh[0]) atomicAdd(chk, sh[0]);
		if (sh[1]) atomicAdd(chk + 1, sh[1]);
	}
}

__global__ __launch_bounds__(TILE * TILE) void act_setup(Launch p)
{
	__shared__ unsigned sh_mark;
	const int tid = threadIdx.x, la = tid / TILE, lb = tid % TILE;
	for (int t = blockIdx.x; t < p.n_tiles; t += gridDim.x) {
		if (tid == 0) sh_mark = 0u;
		__syncthreads();
		const int ti = t % p.ty, tj = t / p.ty;
		const int tre = min(TILE, p.rows - ti * TILE), tce = min(TILE, p.cols - tj * TILE);
		if (la < tre && lb < tce) {
			const size_t c = (size_t)(ti * TILE + la) * p.cols + tj * TILE + lb;
			p.h_out[c] = p.h_in[c];
			const unsigned f = p.fl_in[c] != 0 ? 1u : 0u;
			p.fl_out[c] = (uint8_t)f;
			const uint8_t g = p.gh_in[c];
			p.gh_out[c] = g;
			if (g)
				for (int q = 0; q < 4; ++q) p.F_out[q][c] = p.F_in[q][c];
			if (f) atomicOr(&sh_mark, reach_marks(la, lb, tre, tce, p.km));
		}
		__syncthreads();
		const unsigned marks = sh_mark;
		__syncthreads();
		list_marked(marks, tid, ti, tj, p.ty, p.tx, p.L.tag, p.e1, p.L.list[p.par], p.L.ctr + p.rot);
	}
}

template <int K>
__global__ __launch_bounds__((TILE + 4 * K) * (TILE + 4 * K)) __attribute__((amdgpu_waves_per_eu(K == 4 ? 8 : 1)))
void act_fused(Launch p)
{
	constexpr int H = 2 * K, RW = TILE + 2 * H, NT = RW * RW;
	__shared__ double su[NT];       // levels z + h
	__shared__ double so[4][NT];    // outflows towards N, W, E, S
	__shared__ uint8_t sk[NT];      // minimisation keep bits (1..4: N, W, E, S)
	__shared__ unsigned sh_mark, sh_bits;
	const int tid = threadIdx.x;
	const int a = tid / RW, b = tid - (tid / RW) * RW;
	const int rot = p.rot;
	const int n_in = p.L.ctr[rot];
	const unsigned pm = p.L.mask[(rot + 2) % 3];
	// this block's first list entry beside the count (gridDim.x <= n_tiles)
	const int t_first = p.L.list[p.par][blockIdx.x];
	if (blockIdx.x == 0 && tid == 0) {
		p.L.ctr[(rot + 2) % 3] = 0;       // appended to by the next launch
		p.L.mask[(rot + 1) % 3] = 0u;     // written by the next launch
	}
	if (!(pm & MASK_SET_AFTER_LAST)) {
		// The set is empty: every step covers the grid with zero flows.
		for (int t = blockIdx.x; t < p.n_tiles; t += gridDim.x) {
			const int ti = t % p.ty, tj = t / p.ty;
			for (int q = tid; q < TILE * TILE; q += NT) {
				const int i = ti * TILE + q / TILE, j = tj * TILE + q % TILE;
				if (i < p.rows && j < p.cols) {
					const size_t c = (size_t)i * p.cols + j;
					double v = p.h_in[c];
					v += 0.0;   // h + (0.0 - 0.0), sciddicaT.c:98-99 with every flow zero
					p.h_out[c] = v;
					p.fl_out[c] = 0;
					p.gh_out[c] = 0;
				}
			}
		}
		if (blockIdx.x == 0 && tid == 0) atomicOr(p.L.mask + rot, MASK_WHOLE_GRID);
		return;
	}
	const double r = p.r, eps = p.eps;
	for (int k = blockIdx.x; k < n_in; k += gridDim.x) {
		const int t = k == (int)blockIdx.x ? t_first : p.L.list[p.par][k];
		const int ti = t % p.ty, tj = t / p.ty;
		if (tid == 0) {
			sh_mark = 0u;
			sh_bits = 0u;
		}
		const int gi = wrapm(ti * TILE - H + a, p.rows), gj = wrapm(tj * TILE - H + b, p.cols);
		const size_t c = (size_t)gi * p.cols + gj;
		const double h0 = p.h_in[c];
		const double z = p.z[c];
		const unsigned fl0 = p.fl_in[c] != 0 ? 1u : 0u;
		const unsigned gh0 = p.gh_in[c];
		// A ghost's stored flows live in its own outflow slots: each step
		// rewrites them with what the cell sends (its stored flows, or the
		// flows it computes in the step it leaves the set).
		if (gh0)
			for (int q = 0; q < 4; ++q) so[q][tid] = p.F_in[q][c];
		double h = h0;
		unsigned fl = fl0, gh = gh0;
		const int la = a - H, lb = b - H;
		const int tre = min(TILE, p.rows - ti * TILE), tce = min(TILE, p.cols - tj * TILE);
		const bool out = la >= 0 && la < tre && lb >= 0 && lb < tce;
		unsigned bits = 0u;   // bit s: this output cell is in the set after step s + 1
#pragma unroll 1
		for (int st = 0; st < K; ++st) {
			// The cells whose h is still valid are [lo, hi)^2; this step needs
			// the flows of [lo + 1, hi - 1)^2 and updates [lo + 2, hi - 2)^2,
			// which holds the tile after the last step.
			const int lo = 2 * st, hi = RW - 2 * st;
			const bool fb = a >= lo + 1 && a < hi - 1 && b >= lo + 1 && b < hi - 1;
			const bool wb = a >= lo + 2 && a < hi - 2 && b >= lo + 2 && b < hi - 2;
			if (a >= lo && a < hi && b >= lo && b < hi) su[tid] = z + h;
			__syncthreads();
			double o[4] = {0.0, 0.0, 0.0, 0.0};
			if (fb) {
				unsigned keep = 0u;
				if (!fl && gh) {
#pragma unroll
					for (int q = 0; q < 4; ++q) o[q] = so[q][tid];
				}
				if (fl && !(h <= eps)) {
					const FlowOut fo =
					    flows_wet_body(h, z, su[tid - RW], su[tid - 1], su[tid + 1], su[tid + RW], r, eps);
					o[0] = fo.f0;
					o[1] = fo.f1;
					o[2] = fo.f2;
					o[3] = fo.f3;
					keep = fo.keep;
				}
#pragma unroll
				for (int q = 0; q < 4; ++q) so[q][tid] = o[q];
				sk[tid] = (uint8_t)keep;
			}
			__syncthreads();
			if (wb) {
				// receivers: the north neighbour kept its south level, and so on
				const unsigned recv =
				    ((sk[tid - RW] >> 4) | (sk[tid - 1] >> 3) | (sk[tid + 1] >> 2) | (sk[tid + RW] >> 1)) & 1u;
				if (fl || recv) {
					double v = h;
					v += so[3][tid - RW] - o[0];
					v += so[2][tid - 1] - o[1];
					v += so[1][tid + 1] - o[2];
					v += so[0][tid + RW] - o[3];
					const unsigned nf = !(v <= eps) ? 1u : 0u;
					if (nf) {
						gh = 0u;   // the steering resets the flows of the set's cells
					} else if (fl) {
						// left the set: keeps the flows it wrote (zero for a dry
						// cell), which are in its outflow slots
						gh = (bits_of(o[0]) | bits_of(o[1]) | bits_of(o[2]) | bits_of(o[3])) != 0ull ? 1u : 0u;
					}
					fl = nf;
					h = v;
				}
			}
			if (out && fl) bits |= 1u << st;
		}
		unsigned mk = 0u;
		if (out) {
			p.h_out[c] = h;
			p.fl_out[c] = (uint8_t)fl;
			p.gh_out[c] = (uint8_t)gh;
			bool changed = bits_of(h) != bits_of(h0) || fl != fl0 || gh != gh0;
			if (gh) {
#pragma unroll
				for (int q = 0; q < 4; ++q) {
					const double v = so[q][tid];
					p.F_out[q][c] = v;
					// a ghost at the start too: compare with what it held
					changed = changed || !gh0 || bits_of(v) != bits_of(p.F_in[q][c]);
				}
			}
			if (fl) mk = reach_marks(la, lb, tre, tce, p.km);
			if (changed) mk |= 1u << 4;
		}
		if (mk) atomicOr(&sh_mark, mk);
		if (bits) atomicOr(&sh_bits, bits);
		__syncthreads();
		const unsigned marks = sh_mark, sb = sh_bits;
		__syncthreads();   // read before the next tile clears them
		list_marked(marks, tid, ti, tj, p.ty, p.tx, p.L.tag, p.e1, p.L.list[1 - p.par], p.L.ctr + (rot + 1) % 3);
		if (tid == 9) {
			unsigned m = ((sb >> (K - 1)) & 1u) ? MASK_SET_AFTER_LAST : 0u;
			if (K >= 2 && ((sb >> (K >= 2 ? K - 2 : 0)) & 1u)) m |= MASK_SET_BEFORE_LAST;
			if (m) atomicOr(p.L.mask + rot, m);
		}
	}
	// K = 1: the launch started from a non-empty set
	if (K == 1 && blockIdx.x == 0 && tid == 10) atomicOr(p.L.mask + rot, MASK_SET_BEFORE_LAST);
}

__global__ __launch_bounds__(256) void act_leave(double* __restrict__ h, const uint8_t* __restrict__ fl,
                                                 const uint8_t* __restrict__ gh, double* __restrict__ F0,
                                                 double* __restrict__ F1, double* __restrict__ F2,
                                                 double* __restrict__ F3, const unsigned* mask_last, size_t n)
{
	const unsigned m = *mask_last;
	const bool empty = !(m & MASK_SET_AFTER_LAST);
	const bool norm = empty && !(m & (MASK_SET_BEFORE_LAST | MASK_WHOLE_GRID));
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		if (empty || (!fl[c] && !gh[c])) {
			F0[c] = 0.0;
			F1[c] = 0.0;
			F2[c] = 0.0;
			F3[c] = 0.0;
		}
		if (norm) {
			double v = h[c];
			v += 0.0;
			h[c] = v;
		}
	}
}

__global__ __launch_bounds__(256) void act_normalise(double* __restrict__ h, const unsigned* mask_last, size_t n)
{
	const unsigned m = *mask_last;
	if ((m & MASK_SET_AFTER_LAST) || (m & (MASK_SET_BEFORE_LAST | MASK_WHOLE_GRID))) return;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		double v = h[c];
		v += 0.0;
		h[c] = v;
	}
}

// ---- full grid ----

// The K-step full-grid kernel computes a listed tile in NSUB parts of SUB x
// SUB cells, one block each, so that a tile's steps spread over several
// CUs; a part's region with its halo is (SUB + 4K)^2 cells, over at most
// 1024 threads.
template <int K, int SUB_>
struct PlainShape {
	static constexpr int SUB = SUB_, PER_ROW = TILE / SUB, NSUB = PER_ROW * PER_ROW;
	static constexpr int RW = SUB + 4 * K, NC = RW * RW;
	static constexpr int CPT = (NC + 1023) / 1024;
	static constexpr int NT = (NC + CPT - 1) / CPT;
};

// Lists every tile within p.km of a wet cell (h > eps or NaN) of the grid in
// list_out with tag e; cnt_out is zero on entry.  Every tile starts with no
// copy behind it (same = 0).
__global__ __launch_bounds__(TILE * TILE) void plain_setup(PlainLaunch p)
{
	__shared__ unsigned sh_mark;
	const int tid = threadIdx.x, la = tid / TILE, lb = tid % TILE;
	for (int t = blockIdx.x; t < p.n_tiles; t += gridDim.x) {
		if (tid == 0) {
			sh_mark = 0u;
			p.same[t] = 0u;
		}
		__syncthreads();
		const int ti = t % p.ty, tj = t / p.ty;
		const int tre = min(TILE, p.rows - ti * TILE), tce = min(TILE, p.cols - tj * TILE);
		if (la < tre && lb < tce) {
			const size_t c = (size_t)(ti * TILE + la) * p.cols + tj * TILE + lb;
			if (!(p.h_in[c] <= p.eps)) atomicOr(&sh_mark, reach_marks(la, lb, tre, tce, p.km));
		}
		__syncthreads();
		const unsigned marks = sh_mark;
		__syncthreads();
		list_marked(marks, tid, ti, tj, p.ty, p.tx, p.tag_out, p.e, p.list_out, p.cnt_out);
	}
}

// Blocks [0, grid_comp): K steps of the listed tiles, a block per part of
// SUB x SUB cells (with its own halo; parts of one tile overlap only in
// what they read).  The others: the
// copy h + 0.0 of every unlisted tile, each block a band of 16 rows (one
// tile row) and 2 * NT columns, two cells per thread and row.
//
// An unlisted tile stays as it is (every flow around its cells is zero): h'
// = h + 0.0, which only turns -0.0 into +0.0.  Once it has been copied by
// two consecutive launches, both layers hold h + 0.0 of the same values, and
// every later launch would write what the layer it writes already holds: the
// copy of such a tile is left out until the tile is listed again.  same[t]
// holds the e of the first launch of the tile's run of copies (0: listed
// since, or no copy yet); a launch copies the tile unless 0 < same[t] <=
// e - 2.  The thread that finds 0 stores e: threads of the same launch that
// read it afterwards see e and copy as well.
template <int K, int SUB>
__global__ __launch_bounds__((PlainShape<K, SUB>::NT)) void plain_tiles(PlainLaunch p)
{
	using S = PlainShape<K, SUB>;
	constexpr int H = 2 * K, RW = S::RW, NC = S::NC, CPT = S::CPT, NT = S::NT;
	__shared__ double su[NC];
	__shared__ double so[4][NC];
	// the tile marks live in so[0][0]: outflows are written for rows and
	// columns 1 .. RW - 2 and read for 1 .. RW - 2 only, so the corner entry
	// is never a flow
	unsigned& sh_mark = *reinterpret_cast<unsigned*>(&so[0][0]);
	const int tid = threadIdx.x;
	if ((int)blockIdx.x >= p.grid_comp) {
		// copy role
		const int band = (int)blockIdx.x - p.grid_comp;
		const int per_row = (p.cols + 2 * NT - 1) / (2 * NT);
		const int ti = band / per_row, bc = band - ti * per_row;
		if (ti >= p.ty) return;
		const int j = bc * 2 * NT + 2 * tid;
		if (j >= p.cols) return;
		const int tj0 = j / TILE, tj1 = (j + 1) / TILE;   // j even: one tile
		const bool c0 = p.tag_in[ti + tj0 * p.ty] != p.e;
		const bool c1 = j + 1 < p.cols && p.tag_in[ti + tj1 * p.ty] != p.e;
		if (c0) {
			unsigned* sm = p.same + ti + tj0 * p.ty;
			const unsigned first = *sm;
			if (first != 0u && first + 2u <= p.e) return;
			if (first == 0u) *sm = p.e;
		}
		const int i0 = ti * TILE, nr = min(TILE, p.rows - i0);
		const bool pair = c0 && c1 && (p.cols % 2 == 0);   // j even: 16-byte aligned pair
		if (pair) {
			// every row's load before the first store (h_in and h_out are
			// different layers), so the band's reads are in flight together
			double2 v[TILE];
#pragma unroll
			for (int r = 0; r < TILE; ++r)
				if (r < nr) v[r] = *reinterpret_cast<const double2*>(p.h_in + (size_t)(i0 + r) * p.cols + j);
#pragma unroll
			for (int r = 0; r < TILE; ++r)
				if (r < nr) {
					v[r].x += 0.0;   // h + (0.0 - 0.0) x 4: every flow around the cell is zero
					v[r].y += 0.0;
					*reinterpret_cast<double2*>(p.h_out + (size_t)(i0 + r) * p.cols + j) = v[r];
				}
		} else {
			for (int r = 0; r < nr; ++r) {
				const size_t c = (size_t)(i0 + r) * p.cols + j;
				if (c0) p.h_out[c] = p.h_in[c] + 0.0;
				if (c1) p.h_out[c + 1] = p.h_in[c + 1] + 0.0;
			}
		}
		return;
	}
	if (blockIdx.x == 0 && tid == 0) *p.cnt_clear = 0;
	const int n_in = *p.cnt_in;
	const double r = p.r, eps = p.eps;
	for (int k = blockIdx.x; k < n_in * S::NSUB; k += p.grid_comp) {
		const int t = p.list_in[k / S::NSUB], part = k % S::NSUB;
		const int ti = t % p.ty, tj = t / p.ty;
		const int tre = min(TILE, p.rows - ti * TILE), tce = min(TILE, p.cols - tj * TILE);
		// the part's first row and column in the tile; a part outside the
		// grid (a ragged last tile) has nothing to do (uniform over the block)
		const int oa = (part / S::PER_ROW) * S::SUB, ob = (part % S::PER_ROW) * S::SUB;
		if (oa >= tre || ob >= tce) continue;
		if (tid == 0) sh_mark = 0u;
		double z[CPT], h[CPT];
#pragma unroll
		for (int u = 0; u < CPT; ++u) {
			const int q = tid + u * NT;
			const int a = q / RW, b = q - (q / RW) * RW;
			z[u] = 0.0;
			h[u] = 0.0;
			if (q < NC) {
				const size_t c = (size_t)wrapm(ti * TILE + oa - H + a, p.rows) * p.cols +
				                 wrapm(tj * TILE + ob - H + b, p.cols);
				h[u] = p.h_in[c];
				z[u] = p.z[c];
			}
		}
#pragma unroll 1
		for (int st = 0; st < K; ++st) {
			const int lo = 2 * st, hi = RW - 2 * st;
#pragma unroll
			for (int u = 0; u < CPT; ++u) {
				const int q = tid + u * NT;
				const int a = q / RW, b = q - (q / RW) * RW;
				if (q < NC && a >= lo && a < hi && b >= lo && b < hi) su[q] = z[u] + h[u];
			}
			__syncthreads();
#pragma unroll
			for (int u = 0; u < CPT; ++u) {
				const int q = tid + u * NT;
				const int a = q / RW, b = q - (q / RW) * RW;
				if (q < NC && a >= lo + 1 && a < hi - 1 && b >= lo + 1 && b < hi - 1) {
					double o[4] = {0.0, 0.0, 0.0, 0.0};
					if (!(h[u] <= eps)) {
						const FlowOut fo = flows_wet_body(h[u], z[u], su[q - RW], su[q - 1], su[q + 1], su[q + RW], r, eps);
						o[0] = fo.f0;
						o[1] = fo.f1;
						o[2] = fo.f2;
						o[3] = fo.f3;
					}
#pragma unroll
					for (int d = 0; d < 4; ++d) so[d][q] = o[d];
				}
			}
			__syncthreads();
#pragma unroll
			for (int u = 0; u < CPT; ++u) {
				const int q = tid + u * NT;
				const int a = q / RW, b = q - (q / RW) * RW;
				if (q < NC && a >= lo + 2 && a < hi - 2 && b >= lo + 2 && b < hi - 2) {
					double v = h[u];
					v += so[3][q - RW] - so[0][q];
					v += so[2][q - 1] - so[1][q];
					v += so[1][q + 1] - so[2][q];
					v += so[0][q + RW] - so[3][q];
					h[u] = v;
				}
			}
			// the next step writes su and so only after its first barrier,
			// when every read of this step is done
		}
		unsigned mk = 0u;
#pragma unroll
		for (int u = 0; u < CPT; ++u) {
			const int q = tid + u * NT;
			const int a0 = q / RW - H, b0 = q - (q / RW) * RW - H;
			const int la = oa + a0, lb = ob + b0;
			if (q < NC && a0 >= 0 && a0 < S::SUB && b0 >= 0 && b0 < S::SUB && la < tre && lb < tce) {
				const size_t c = (size_t)(ti * TILE + la) * p.cols + tj * TILE + lb;
				p.h_out[c] = h[u];
				if (!(h[u] <= eps)) mk |= reach_marks(la, lb, tre, tce, p.km);
			}
		}
		if (mk) atomicOr(&sh_mark, mk);
		if (tid == 0) p.same[t] = 0u;   // the layers differ from here
		__syncthreads();
		const unsigned marks = sh_mark;
		__syncthreads();   // read before the next tile clears it
		list_marked(marks, tid, ti, tj, p.ty, p.tx, p.tag_out, p.e + 1u, p.list_out, p.cnt_out);
	}
}

int grid_for(size_t n)
{
	const size_t b = (n + 255) / 256;
	return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

}  // namespace

int launch_check(hipStream_t st, const double* h, const uint8_t* fl, const double* const F[4], bool f_zero,
                 uint8_t* gh, int* chk, size_t n, double eps)
{
	hipLaunchKernelGGL(act_check, dim3(grid_for(n)), dim3(256), 0, st, h, fl, F[0], F[1], F[2], F[3],
	                   f_zero ? 1 : 0, gh, chk, n, eps);
	return CALHIP_CHECK(hipGetLastError());
}

int launch_setup(hipStream_t st, const Launch& p)
{
	const int grid = p.n_tiles < 2048 ? p.n_tiles : 2048;
	hipLaunchKernelGGL(act_setup, dim3(grid), dim3(TILE * TILE), 0, st, p);
	return CALHIP_CHECK(hipGetLastError());
}

int launch_fused(hipStream_t st, int K, const Launch& p, int grid)
{
	switch (K) {
	case 1:
		hipLaunchKernelGGL(act_fused<1>, dim3(grid), dim3((TILE + 4) * (TILE + 4)), 0, st, p);
		break;
	case 2:
		hipLaunchKernelGGL(act_fused<2>, dim3(grid), dim3((TILE + 8) * (TILE + 8)), 0, st, p);
		break;
	case 4:
		hipLaunchKernelGGL(act_fused<4>, dim3(grid), dim3((TILE + 16) * (TILE + 16)), 0, st, p);
		break;
	default:
		set_error(13, "SciddicaT active launch: steps per launch must be 1, 2 or 4");
		return -1;
	}
	return CALHIP_CHECK(hipGetLastError());
}

int launch_leave(hipStream_t st, double* h, const uint8_t* fl, const uint8_t* gh, double* const F[4],
                 const unsigned* mask_last, size_t n)
{
	hipLaunchKernelGGL(act_leave, dim3(grid_for(n)), dim3(256), 0, st, h, fl, gh, F[0], F[1], F[2], F[3],
	                   mask_last, n);
	return CALHIP_CHECK(hipGetLastError());
}

template <int SUB>
static int plain_nt(int K)
{
	switch (K) {
	case 1: return PlainShape<1, SUB>::NT;
	case 2: return PlainShape<2, SUB>::NT;
	case 3: return PlainShape<3, SUB>::NT;
	default: return PlainShape<4, SUB>::NT;
	}
}

int plain_threads(int K, bool parts) { return parts ? plain_nt<PLAIN_PART>(K) : plain_nt<TILE>(K); }

int plain_copy_blocks(int K, bool parts, int rows, int cols)
{
	const int nt = plain_threads(K, parts);
	return ((rows + TILE - 1) / TILE) * ((cols + 2 * nt - 1) / (2 * nt));
}

int launch_plain_setup(hipStream_t st, const PlainLaunch& p)
{
	const int grid = p.n_tiles < 4096 ? p.n_tiles : 4096;
	hipLaunchKernelGGL(plain_setup, dim3(grid), dim3(TILE * TILE), 0, st, p);
	return CALHIP_CHECK(hipGetLastError());
}

template <int K>
static void plain_go(hipStream_t st, int grid, bool parts, const PlainLaunch& p)
{
	if (parts)
		hipLaunchKernelGGL((plain_tiles<K, PLAIN_PART>), dim3(grid), dim3(PlainShape<K, PLAIN_PART>::NT), 0, st, p);
	else
		hipLaunchKernelGGL((plain_tiles<K, TILE>), dim3(grid), dim3(PlainShape<K, TILE>::NT), 0, st, p);
}

int launch_plain(hipStream_t st, int K, bool parts, const PlainLaunch& p)
{
	if (K < 1 || K > PLAIN_KMAX) {
		set_error(13, "SciddicaT full-grid launch: steps per launch must be 1 to 4");
		return -1;
	}
	const int grid = p.grid_comp + plain_copy_blocks(K, parts, p.rows, p.cols);
	switch (K) {
	case 1: plain_go<1>(st, grid, parts, p); break;
	case 2: plain_go<2>(st, grid, parts, p); break;
	case 3: plain_go<3>(st, grid, parts, p); break;
	default: plain_go<4>(st, grid, parts, p); break;
	}
	return CALHIP_CHECK(hipGetLastError());
}

int launch_normalise(hipStream_t st, double* h, const unsigned* mask_last, size_t n)
{
	hipLaunchKernelGGL(act_normalise, dim3(grid_for(n)), dim3(256), 0, st, h, mask_last, n);
	return CALHIP_CHECK(hipGetLastError());
}

}  // namespace actf
}  // namespace calhip
