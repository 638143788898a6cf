// SciddicaT device program.
//
// Reference transition (OpenCAL/examples/cal_sciddicaT/source/sciddicaT.c):
//   flows  (:36-88)  minimisation of differences over the von Neumann
//                    neighbourhood: outflows f[0..3] towards N, W, E, S;
//   width  (:91-101) h' = h + sum_{n=1..4} (f[4-n](neighbour n) - f[n-1](self));
//   steering (:133-140) resets the flows to 0 every step;
//   active-cells variant (cal_sciddicaT-activecells-glut/source/sciddicaT.c):
//   receivers of an outflow join the set (:74-77), cells with h <= eps leave
//   it (:94-100).
//
// Kernels:
//   step_fused_al, step_fused2 (sciddicat_device.hpp)
//                 full grid, both processes in one sweep.  Each wave owns a
//                 strip (128 aligned columns for even column counts, 124 plus
//                 a 2-column halo for odd ones), two cells per lane, and walks
//                 a chunk of rows with a register ring of z and h rows;
//                 horizontal neighbours come from the adjacent lanes.  The
//                 outflows of a cell are computed once, consumed by the width
//                 update of its four neighbours through lane shuffles and the
//                 ring, and never written to memory.  HBM traffic per cell: z
//                 and h read, h written = 24 B.
//   step_fused_2x (sweep_tune.hip, strip sweep in sciddicat_2step.hpp)
//                 two steps per sweep when the steering is declared to be the
//                 reference's flow reset (calhipBindSciddicaTSteering2D).
//   flows_full / width_full
//                 one process per sweep with the flows written out, the exact
//                 reference sequence for models whose flows are observed.
//   tiles_from_flags, flows_tiles, width_tiles
//                 the active-cells step on the flag grid over lists of 4 x 64
//                 cell tiles (the set's tiles and the receivers' tiles).
//   flows_active / width_active
//                 over the row-major active list (or the whole grid when it is
//                 empty, as OpenCAL does when A.cells == NULL), for a step
//                 that starts after the host changed the set.
//
// Arithmetic follows the reference operation by operation (no contraction:
// the library is built with -ffp-contract=off), so results are bit-identical
// to the CPU path; the division of the level sum by the level count is
// computed with FMAs in a form that gives the IEEE quotient's bits
// (div_count, sciddicat_device.hpp).
#include <stdio.h>

#include <algorithm>
#include <cmath>

#include <stdlib.h>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_device.hpp"
#include "sciddicat_active.hpp"

namespace calhip {

namespace {

using namespace sct;

struct FlowPtrs {
	double* f[4];
};

// flows process over the whole grid, writing the flows of the next layer.
// Dry cells keep their current flows (the reference does not write them).
__global__ void flows_full(const double* __restrict__ z, const double* __restrict__ h, FlowPtrs fcur,
                           FlowPtrs fnxt, int rows, int cols, double r, double eps)
{
	size_t n = (size_t)rows * cols;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		int i = (int)(c / cols), j = (int)(c % cols);
		int in = dev::wrap1(i - 1, rows), is = dev::wrap1(i + 1, rows);
		int jw = dev::wrap1(j - 1, cols), je = dev::wrap1(j + 1, cols);
		double h0 = h[c];
		if (h0 <= eps) {
			if (fnxt.f[0] != fcur.f[0])
				for (int q = 0; q < 4; ++q) fnxt.f[q][c] = fcur.f[q][c];
			continue;
		}
		size_t cn = (size_t)in * cols + j, cw = (size_t)i * cols + jw, ce = (size_t)i * cols + je,
		       cs = (size_t)is * cols + j;
		double f[4];
		cell_flows(h0, z[c], z[cn] + h[cn], z[cw] + h[cw], z[ce] + h[ce], z[cs] + h[cs], r, eps, f);
		for (int q = 0; q < 4; ++q) fnxt.f[q][c] = f[q];
	}
}

// width process over the whole grid from the current flows.
__global__ void width_full(const double* __restrict__ h, FlowPtrs f, double* __restrict__ hn, int rows, int cols)
{
	size_t n = (size_t)rows * cols;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
		int i = (int)(c / cols), j = (int)(c % cols);
		size_t cn = (size_t)dev::wrap1(i - 1, rows) * cols + j;
		size_t cw = (size_t)i * cols + dev::wrap1(j - 1, cols);
		size_t ce = (size_t)i * cols + dev::wrap1(j + 1, cols);
		size_t cs = (size_t)dev::wrap1(i + 1, rows) * cols + j;
		double o = h[c];
		o += f.f[3][cn] - f.f[0][c];
		o += f.f[2][cw] - f.f[1][c];
		o += f.f[1][ce] - f.f[2][c];
		o += f.f[0][cs] - f.f[3][c];
		hn[c] = o;
	}
}

// flows over the active list (or the whole grid when the list is empty).
// Receivers of an outflow get their flag set.  zero_dry: the flows of the
// visited cells are logically zero on entry (the previous steering reset
// them), so dry cells write zeros — this is the deferred steering reset.
__global__ void flows_active(const double* __restrict__ z, const double* __restrict__ h, FlowPtrs f,
                             uint8_t* __restrict__ flags, const int* __restrict__ cells,
                             const int* __restrict__ count, int rows, int cols, double r, double eps, int zero_dry)
{
	const int cnt = count[0];
	const size_t total = cnt > 0 ? (size_t)cnt : (size_t)rows * cols;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
		const size_t c = cnt > 0 ? (size_t)cells[k] : k;
		const int i = (int)(c / cols), j = (int)(c % cols);
		const double h0 = h[c];
		if (h0 <= eps) {
			if (zero_dry)
				for (int q = 0; q < 4; ++q) f.f[q][c] = 0.0;
			continue;
		}
		size_t nb[4];
		nb[0] = (size_t)dev::wrap1(i - 1, rows) * cols + j;
		nb[1] = (size_t)i * cols + dev::wrap1(j - 1, cols);
		nb[2] = (size_t)i * cols + dev::wrap1(j + 1, cols);
		nb[3] = (size_t)dev::wrap1(i + 1, rows) * cols + j;
		double fl[4];
		unsigned keep = cell_flows(h0, z[c], z[nb[0]] + h[nb[0]], z[nb[1]] + h[nb[1]], z[nb[2]] + h[nb[2]],
		                           z[nb[3]] + h[nb[3]], r, eps, fl);
		for (int q = 0; q < 4; ++q) {
			f.f[q][c] = fl[q];
			if ((keep >> (q + 1)) & 1u) flags[nb[q]] = 1;
		}
	}
}

// width over the active list from hin into hout (the same array for the
// in-place step: the process reads only its own h), optionally followed by
// the removal of cells with h <= eps.
__global__ void width_active(const double* hin, double* hout, FlowPtrs f, uint8_t* __restrict__ flags,
                             const int* __restrict__ cells, const int* __restrict__ count, int rows, int cols,
                             double eps, int remove)
{
	const int cnt = count[0];
	const size_t total = cnt > 0 ? (size_t)cnt : (size_t)rows * cols;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
		const size_t c = cnt > 0 ? (size_t)cells[k] : k;
		const int i = (int)(c / cols), j = (int)(c % cols);
		size_t cn = (size_t)dev::wrap1(i - 1, rows) * cols + j;
		size_t cw = (size_t)i * cols + dev::wrap1(j - 1, cols);
		size_t ce = (size_t)i * cols + dev::wrap1(j + 1, cols);
		size_t cs = (size_t)dev::wrap1(i + 1, rows) * cols + j;
		double o = hin[c];
		o += f.f[3][cn] - f.f[0][c];
		o += f.f[2][cw] - f.f[1][c];
		o += f.f[1][ce] - f.f[2][c];
		o += f.f[0][cs] - f.f[3][c];
		hout[c] = o;
		if (remove && o <= eps) flags[c] = 0;
	}
}

// the removal process alone (active variant sciddicaT.c:94-100): cells of
// the list (or of the grid when it is empty) with h <= eps leave the set
__global__ void remove_active(const double* __restrict__ h, uint8_t* __restrict__ flags,
                              const int* __restrict__ cells, const int* __restrict__ count, size_t n, double eps)
{
	const int cnt = count[0];
	const size_t total = cnt > 0 ? (size_t)cnt : n;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
		const size_t c = cnt > 0 ? (size_t)cells[k] : k;
		if (h[c] <= eps) flags[c] = 0;
	}
}

// zero the flows at the active cells (or everywhere when the list is empty).
__global__ void zero_at_active(FlowPtrs f, const int* __restrict__ cells, const int* __restrict__ count, size_t n)
{
	const int cnt = count[0];
	const size_t total = cnt > 0 ? (size_t)cnt : n;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
		const size_t c = cnt > 0 ? (size_t)cells[k] : k;
		for (int q = 0; q < 4; ++q) f.f[q][c] = 0.0;
	}
}

// The active step driven by the flag grid instead of the row-major list.
// A flag byte is 1 while its cell is in the set; the receivers of a step's
// outflows are marked in a second byte grid (recv), which the width process
// clears again.  The grid is cut into tiles of 4 rows x 64 columns (one wave
// per row segment) and each launch visits only listed tiles.  For step epoch
// e (a counter the host passes, one per step):
//   * S[p]: the tiles that hold cells of the set, each tagged tag[t] == 2e
//     (built from the flags when the step starts from a host-visible state,
//     then by the width process of the step before);
//   * X[p]: the other tiles that hold receivers, each once: a tile joins
//     when its atomic max with 2e + 1 returns a tag below 2e (an older
//     epoch), so a tile of S[p] or one already listed is not listed again.
// The width process visits S[p] and X[p].  No launch depends on a host
// value and no cell list is rebuilt between the processes; the row-major
// list is rebuilt only when it is observed.  Every cell processed is one the
// list path would process: flows visits the set (every cell while it is
// empty, as OpenCAL does when A.cells == NULL), width and removal visit the
// set plus its receivers (every cell while both are empty).  Counters
// ctr[0..1] = |S[0..1]|, ctr[2..3] = |X[0..1]|.
struct TileLists {
	int* ctr;                // [4]
	int* set_list[2];        // [n_tiles] each
	int* x_list[2];          // [n_tiles] each
	unsigned* tag;           // [n_tiles] 2e: in S of epoch e; 2e + 1: reached in epoch e
	uint8_t* recv;           // [cells]
	int ty, tx, n_tiles;     // tile rows, tile columns, tiles
};

__device__ __forceinline__ int tile_id(const TileLists& tl, int i, int j) { return i / 4 + (j / 64) * tl.ty; }

// The set's tiles into S[p] (ctr[p] is zero on entry).
__global__ __launch_bounds__(256) void tiles_from_flags(const uint8_t* __restrict__ flags, TileLists tl, int p,
                                                        unsigned e, int rows, int cols)
{
	for (int t = blockIdx.x; t < tl.n_tiles; t += gridDim.x) {
		const int i = (t % tl.ty) * 4 + (int)(threadIdx.x >> 6);
		const int j = (t / tl.ty) * 64 + (int)(threadIdx.x & 63);
		const bool in = i < rows && j < cols && flags[(size_t)i * cols + j] != 0;
		if (__syncthreads_or((int)in) && threadIdx.x == 0) {
			tl.set_list[p][atomicAdd(tl.ctr + p, 1)] = t;
			tl.tag[t] = 2u * e;
		}
	}
}

// flows over the set (every cell while it is empty).  A listed tile's cells
// load their flag and all their levels at once.  Receivers get recv = 1 (a
// plain byte store: every writer stores 1).  A block collects the tiles its
// receivers lie in (its own, or the tile beyond one of its four edges) and
// lists those not in the set in X[p].  Clears |S[1-p]|, which the width
// process of this step fills.
__global__ __launch_bounds__(256) void flows_tiles(const double* __restrict__ z, const double* __restrict__ h,
                                                   FlowPtrs f, const uint8_t* __restrict__ flags, TileLists tl,
                                                   int p, unsigned e, int rows, int cols, double r, double eps,
                                                   int zero_dry)
{
	__shared__ unsigned sh_dirs;
	if (blockIdx.x == 0 && threadIdx.x == 0) tl.ctr[1 - p] = 0;
	const int n_set = tl.ctr[p];
	// this block's first list entry, loaded beside the count instead of after
	// it (in bounds: gridDim.x <= n_tiles); used only when the count covers it
	const int t_first = tl.set_list[p][blockIdx.x];
	const bool all = n_set == 0;
	const int n_iter = all ? tl.n_tiles : n_set;
	for (int k = blockIdx.x; k < n_iter; k += gridDim.x) {
		const int t = all ? k : (k == (int)blockIdx.x ? t_first : tl.set_list[p][k]);
		if (threadIdx.x == 0) sh_dirs = 0u;
		__syncthreads();
		const int i = (t % tl.ty) * 4 + (int)(threadIdx.x >> 6);
		const int j = (t / tl.ty) * 64 + (int)(threadIdx.x & 63);
		unsigned dirs = 0u;   // bit q: a receiver beyond edge q (N, W, E, S); bit 4: one in this tile
		if (i < rows && j < cols) {
			const size_t c = (size_t)i * cols + j;
			const int iN = dev::wrap1(i - 1, rows), iS = dev::wrap1(i + 1, rows);
			const int jW = dev::wrap1(j - 1, cols), jE = dev::wrap1(j + 1, cols);
			size_t nb[4];
			nb[0] = (size_t)iN * cols + j;
			nb[1] = (size_t)i * cols + jW;
			nb[2] = (size_t)i * cols + jE;
			nb[3] = (size_t)iS * cols + j;
			const bool in = all || flags[c] != 0;
			const double h0 = h[c], z0 = z[c];
			const double u1 = z[nb[0]] + h[nb[0]], u2 = z[nb[1]] + h[nb[1]];
			const double u3 = z[nb[2]] + h[nb[2]], u4 = z[nb[3]] + h[nb[3]];
			if (in) {
				if (h0 <= eps) {
					if (zero_dry)
						for (int q = 0; q < 4; ++q) f.f[q][c] = 0.0;
				} else {
					double fl[4];
					const unsigned keep = cell_flows(h0, z0, u1, u2, u3, u4, r, eps, fl);
					const int ni[4] = {iN, i, i, iS}, nj[4] = {j, jW, jE, j};
					for (int q = 0; q < 4; ++q) {
						f.f[q][c] = fl[q];
						if ((keep >> (q + 1)) & 1u) {
							tl.recv[nb[q]] = 1;
							dirs |= tile_id(tl, ni[q], nj[q]) == t ? 16u : 1u << q;
						}
					}
				}
			}
		}
		if (dirs) atomicOr(&sh_dirs, dirs);
		__syncthreads();
		if (threadIdx.x < 5 && ((sh_dirs >> threadIdx.x) & 1u)) {
			const int tr = t % tl.ty, tc = t / tl.ty;
			int u = t;
			if (threadIdx.x == 0) u = (tr + tl.ty - 1) % tl.ty + tc * tl.ty;
			if (threadIdx.x == 1) u = tr + ((tc + tl.tx - 1) % tl.tx) * tl.ty;
			if (threadIdx.x == 2) u = tr + ((tc + 1) % tl.tx) * tl.ty;
			if (threadIdx.x == 3) u = (tr + 1) % tl.ty + tc * tl.ty;
			if (atomicMax(tl.tag + u, 2u * e + 1u) < 2u * e) tl.x_list[p][atomicAdd(tl.ctr + 2 + p, 1)] = u;
		}
	}
}

// width over the set and its receivers in the tiles of S[p] and X[p] (every
// cell while both are empty), in place, fused with the removal process: a
// visited cell stays in the set (flag 1) when it was in it or received and
// not h' <= eps; recv is cleared.  The tiles keeping cells of the set go to
// S[1-p] tagged 2(e + 1).  Clears |X[1-p]| for the next flows process.
__global__ __launch_bounds__(256) void width_tiles(double* __restrict__ h, FlowPtrs f, uint8_t* __restrict__ flags,
                                                   TileLists tl, int p, unsigned e, int rows, int cols, double eps)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) tl.ctr[3 - p] = 0;
	const int n_set = tl.ctr[p], n_x = tl.ctr[2 + p];
	// Block b takes the set's entries b, b + G, ... and the receivers' entries
	// G-1-b, 2G-1-b, ... (G = gridDim.x <= n_tiles).  Its first entry of each
	// list is loaded beside the counts instead of after them and used only
	// when the count covers it.
	const int G = (int)gridDim.x, b = (int)blockIdx.x;
	const int s_first = tl.set_list[p][b];
	const int x_first = tl.x_list[p][G - 1 - b];
	const bool all = n_set == 0 && n_x == 0;
	auto visit = [&](int t) {
		const int i = (t % tl.ty) * 4 + (int)(threadIdx.x >> 6);
		const int j = (t / tl.ty) * 64 + (int)(threadIdx.x & 63);
		bool kept = false;
		if (i < rows && j < cols) {
			const size_t c = (size_t)i * cols + j;
			const unsigned fl = flags[c], rv = tl.recv[c];
			if (all || fl || rv) {
				const size_t cn = (size_t)dev::wrap1(i - 1, rows) * cols + j;
				const size_t cw = (size_t)i * cols + dev::wrap1(j - 1, cols);
				const size_t ce = (size_t)i * cols + dev::wrap1(j + 1, cols);
				const size_t cs = (size_t)dev::wrap1(i + 1, rows) * cols + j;
				double o = h[c];
				o += f.f[3][cn] - f.f[0][c];
				o += f.f[2][cw] - f.f[1][c];
				o += f.f[1][ce] - f.f[2][c];
				o += f.f[0][cs] - f.f[3][c];
				h[c] = o;
				const unsigned nf = ((fl || rv) && !(o <= eps)) ? 1u : 0u;
				if (nf != fl) flags[c] = (uint8_t)nf;
				if (rv) tl.recv[c] = 0;
				kept = nf != 0u;
			}
		}
		if (__syncthreads_or((int)kept) && threadIdx.x == 0) {
			tl.set_list[1 - p][atomicAdd(tl.ctr + 1 - p, 1)] = t;
			tl.tag[t] = 2u * (e + 1u);
		}
	};
	if (all) {
		for (int k = b; k < tl.n_tiles; k += G) visit(k);
		return;
	}
	for (int k = b; k < n_set; k += G) visit(k == b ? s_first : tl.set_list[p][k]);
	for (int m = G - 1 - b; m < n_x; m += G) visit(m == G - 1 - b ? x_first : tl.x_list[p][m]);
}

int stride_grid(size_t n, int block)
{
	size_t b = (n + block - 1) / block;
	return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

// Device check of div_count: the refined reciprocal of every count 1..5 is
// within 2^-40 of 1/b (the bound its exactness argument needs), and for n
// seeded doubles (SplitMix64 bit patterns, every other one replaced by a
// value of an altitude's magnitude) and every count 0..5 the bits of
// div_count equal the IEEE quotient's (NaNs compare equal to NaNs); failures
// are counted into *bad.
__global__ __launch_bounds__(256) void check_div_count(long long n, unsigned long long seed,
                                                       unsigned long long* bad)
{
	unsigned long long local = 0;
	if (blockIdx.x == 0 && threadIdx.x < 5) {
		const double b = (double)(threadIdx.x + 1);
		local += fabs(__builtin_fma(b, count_reciprocal(b), -1.0)) <= 0x1p-40 ? 0u : 1u;
	}
	for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
	     k += (long long)gridDim.x * blockDim.x) {
		unsigned long long x = seed + (unsigned long long)k * 0x9e3779b97f4a7c15ULL;
		x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
		x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
		x ^= x >> 31;
		double a = __longlong_as_double((long long)x);
		if (k & 1) a = (double)(x >> 11) * 0x1p-53 * 4000.0 - 100.0;
		for (int c = 0; c <= 5; ++c) {
			const double e = c ? a / (double)c : a;
			const double g = div_count(a, c);
			const bool same = __double_as_longlong(e) == __double_as_longlong(g) || (e != e && g != g);
			local += same ? 0u : 1u;
		}
	}
	if (local) atomicAdd(bad, local);
}

}  // namespace

// ------------------------------------------------------------------ program

struct SciddicaTProgram : Program {
	SubX* z = nullptr;
	SubX* h = nullptr;
	SubX* f[4] = {nullptr, nullptr, nullptr, nullptr};
	const CALreal* p_r = nullptr;
	const CALreal* p_eps = nullptr;
	bool active = false;
	bool h_prev_valid = false;   // h->d_nxt still holds the pre-step thickness

	const char* name() const override { return "SciddicaT"; }

	bool is_flow(SubX* s) const
	{
		for (int q = 0; q < 4; ++q)
			if (f[q] == s) return true;
		return false;
	}

	FlowPtrs cur_flows() const
	{
		FlowPtrs p;
		for (int q = 0; q < 4; ++q) p.f[q] = (double*)f[q]->d_cur;
		return p;
	}

	void set_flow_state(Transient t)
	{
		for (int q = 0; q < 4; ++q) f[q]->tstate = t;
	}

	int after_upload(ModelX* m) override
	{
		size_t n = m->cells();
		bool zero = true;
		for (int q = 0; q < 4 && zero; ++q) {
			const double* a = (const double*)f[q]->h_cur();
			const double* b = (const double*)f[q]->h_nxt();
			for (size_t c = 0; c < n; ++c)
				if (a[c] != 0.0 || b[c] != 0.0 || std::signbit(a[c]) || std::signbit(b[c])) {
					zero = false;
					break;
				}
		}
		set_flow_state(zero ? (active ? Transient::ZeroAtActive : Transient::Zero) : Transient::None);
		h_prev_valid = false;
		act_on = false;       // the upload replaced every buffer the mode kept
		act_blocked = false;
		act_pending = 0;
		pl_valid = false;
		pl_declined = false;
		ctr_valid = false;    // the upload brought the list and its count
		list_stale = false;
		return 0;
	}

	double* d_trash = nullptr;   // sink for the stores of halo lanes
	SweepShape shape = {0, 0};   // chunk length and prefetch depth (0: not chosen yet)
	WetOrder wet_order;          // wet-first tile order of the fused sweep
	Sweep2Shape shape2 = {0, 0}; // tiles of the two-step sweep ({0, 0}: one step per launch is faster)
	bool shape2_known = false;   // shape2 chosen
	WetOrder wet_order2;         // its wet-first order

	// flag-driven active step: tile lists by parity (one allocation from
	// tiles.ctr); the row-major list is rebuilt on demand
	TileLists tiles = {};
	unsigned tile_epoch = 0;     // step epoch of the tile lists (restarts before 2^31)
	int parity = 0;
	bool ctr_valid = false;      // tiles.set_list[parity] lists the set's tiles
	bool list_stale = false;     // m->d_cells / d_count lag behind the flags

	// K steps per launch over tiles with halos (sciddicat_active.hpp): used
	// for steps followed by the steering twin once the entry check passed.
	struct ActBufs {
		uint8_t* flags_b = nullptr;   // second flag copy (trades places with m->d_flags)
		uint8_t* gh[2] = {nullptr, nullptr};
		actf::Lists L = {};
		char* block = nullptr;        // gh, tags, lists, counters
		int ty = 0, tx = 0, n_tiles = 0, km = 0;
		int grid = 0;                 // blocks per launch: two per CU (a launch visits fewer tiles
		                              // than that; blocks past the list's end exit at once)
	} act;
	bool act_on = false;        // the mode's copies hold the state
	bool act_blocked = false;   // the entry check failed; retried after the next upload
	int act_gcur = 0;           // current ghost copy
	int act_par = 0, act_rot = 0;
	int act_last_mask = 2;      // L.mask index of the last launch's mask
	unsigned act_epoch = 0;
	int act_list_reach = 0;     // the current tile list covers launches of up to this many steps
	int act_pending = 0;        // steps accepted by step_and_steer, not launched yet

	// The full grid in launches of up to PL_KMAX steps over the same tiles
	// (sciddicat_active.hpp), for steps followed by the steering twin while
	// few tiles are near wet cells.
	static constexpr int PL_KMAX = actf::PLAIN_KMAX;
	struct PlainBufs {
		unsigned* tag[2] = {nullptr, nullptr};
		int* list[2] = {nullptr, nullptr};
		int* cnt = nullptr;        // [3] list lengths by launch index mod 3
		unsigned* same = nullptr;  // per tile: first launch of its run of copies (0: none)
		char* block = nullptr;
		int* h_cnt = nullptr;      // pinned copy of a list length (read without a stream sync)
		hipEvent_t ev = nullptr;
		int ty = 0, tx = 0, n_tiles = 0, km = 0, cus = 0;
	} pl;
	bool pl_valid = false;       // pl.list[pl_L % 2] lists the tiles the next launch needs
	// CALHIP_SCIDDICAT_TILE_PARTS: 1 always quarter tiles, 0 never, unset: by
	// the list's length
	int pl_parts = getenv("CALHIP_SCIDDICAT_TILE_PARTS") ? (atoi(getenv("CALHIP_SCIDDICAT_TILE_PARTS")) != 0) : -1;
	bool pl_declined = false;    // too many tiles near wet cells for the mode to pay
	int pl_L = 0;                // launch index since the list was built
	unsigned pl_e = 0;           // tag of the current list
	int pl_reach = 0;            // the current list covers launches of up to this many steps
	bool pl_cnt_pending = false; // h_cnt awaits ev
	int act_host_chk[2] = {0, 0};
	unsigned act_host_init[6] = {0, 0, 0, 0, 0, 0};

	~SciddicaTProgram() override
	{
		if (d_trash) hipFree(d_trash);
		if (tiles.ctr) hipFree(tiles.ctr);
		if (act.flags_b) hipFree(act.flags_b);
		if (act.block) hipFree(act.block);
		if (pl.block) hipFree(pl.block);
		if (pl.h_cnt) (void)hipHostFree(pl.h_cnt);
		if (pl.ev) (void)hipEventDestroy(pl.ev);
	}

	// The tiles' marking reach: at most KMAX, and no more than the rows
	// (columns) of a last partial tile row (column) that sits between two
	// others, so that a set cell's reach never skips a tile.
	static int act_reach(int rows, int cols)
	{
		const int T = actf::TILE;
		int km = actf::KMAX;
		if ((rows + T - 1) / T >= 3 && rows % T) km = std::min(km, rows % T);
		if ((cols + T - 1) / T >= 3 && cols % T) km = std::min(km, cols % T);
		return km;
	}

	bool act_eligible(ModelX* m) const
	{
		if (!active || act_blocked || m->flags_dirty || getenv("CALHIP_SCIDDICAT_NO_FUSED_ACTIVE")) return false;
		if (z->nxt_dirty || h->nxt_dirty) return false;
		for (int q = 0; q < 4; ++q)
			if ((f[q]->tstate != Transient::ZeroAtActive && f[q]->tstate != Transient::Zero) || f[q]->nxt_dirty)
				return false;
		return true;
	}

	actf::Launch act_launch_args(ModelX* m) const
	{
		actf::Launch p = {};
		p.z = (const double*)z->d_cur;
		p.h_in = (const double*)h->d_cur;
		p.h_out = (double*)h->d_nxt;
		p.fl_in = m->d_flags;
		p.fl_out = act.flags_b;
		p.gh_in = act.gh[act_gcur];
		p.gh_out = act.gh[1 - act_gcur];
		for (int q = 0; q < 4; ++q) {
			p.F_in[q] = (const double*)f[q]->d_cur;
			p.F_out[q] = (double*)f[q]->d_nxt;
		}
		p.L = act.L;
		p.par = act_par;
		p.rot = act_rot;
		p.rows = m->pub.rows;
		p.cols = m->pub.columns;
		p.ty = act.ty;
		p.tx = act.tx;
		p.n_tiles = act.n_tiles;
		p.km = act.km;
		p.r = *p_r;
		p.eps = *p_eps;
		return p;
	}

	unsigned act_next_epoch(ModelX* m)
	{
		if (act_epoch >= 0xfffffff0u) {
			// the tags would wrap: every tile back to "never listed"
			if (CALHIP_CHECK(hipMemsetAsync(act.L.tag, 0, (size_t)act.n_tiles * sizeof(unsigned), m->stream)))
				return 0;
			act_epoch = 0;
		}
		return ++act_epoch;
	}

	// 1: the mode is on; 0: not eligible (or the entry check failed); -1: error.
	int act_ready(ModelX* m)
	{
		if (act_on) return 1;
		if (!act_eligible(m)) return 0;
		const int rows = m->pub.rows, cols = m->pub.columns;
		const size_t n = m->cells();
		if (!act.block) {
			const int T = actf::TILE;
			act.ty = (rows + T - 1) / T;
			act.tx = (cols + T - 1) / T;
			act.n_tiles = act.ty * act.tx;
			act.km = act_reach(rows, cols);
			int cus = 0;
			if (CALHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device)))
				return -1;
			act.grid = std::max(1, std::min(act.n_tiles, 2 * cus));
			const size_t nt = (size_t)act.n_tiles;
			const size_t gh_bytes = (n + 255) & ~(size_t)255;
			const size_t bytes = 2 * gh_bytes + nt * sizeof(unsigned) + 2 * nt * sizeof(int) + 8 * sizeof(int);
			if (CALHIP_CHECK(hipMalloc(&act.flags_b, (n + 15) & ~(size_t)15))) return -1;
			if (CALHIP_CHECK(hipMalloc(&act.block, bytes))) return -1;
			if (CALHIP_CHECK(hipMemsetAsync(act.block, 0, bytes, m->stream))) return -1;
			act.gh[0] = reinterpret_cast<uint8_t*>(act.block);
			act.gh[1] = act.gh[0] + gh_bytes;
			act.L.tag = reinterpret_cast<unsigned*>(act.gh[1] + gh_bytes);
			act.L.list[0] = reinterpret_cast<int*>(act.L.tag + nt);
			act.L.list[1] = act.L.list[0] + nt;
			act.L.ctr = act.L.list[1] + nt;
			act.L.mask = reinterpret_cast<unsigned*>(act.L.ctr + 3);
			act.L.chk = act.L.ctr + 6;
			act_epoch = 0;
		}
		// entry check: no cell outside the set with h > eps, a non-empty set;
		// ghosts marked in gh[0]
		bool f_zero = true;
		const double* F[4];
		for (int q = 0; q < 4; ++q) {
			f_zero = f_zero && f[q]->tstate == Transient::Zero;
			F[q] = (const double*)f[q]->d_cur;
		}
		if (CALHIP_CHECK(hipMemsetAsync(act.L.chk, 0, 2 * sizeof(int), m->stream))) return -1;
		if (actf::launch_check(m->stream, (const double*)h->d_cur, m->d_flags, F, f_zero, act.gh[0], act.L.chk, n,
		                       *p_eps))
			return -1;
		if (CALHIP_CHECK(hipMemcpyAsync(act_host_chk, act.L.chk, sizeof act_host_chk, hipMemcpyDeviceToHost,
		                                m->stream)) ||
		    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
			return -1;
		m->stats.launches++;
		if (act_host_chk[0] != 0 || act_host_chk[1] == 0) {
			act_blocked = true;
			return 0;
		}
		// counters 0, the "previous launch" of the first (index 2) left a
		// non-empty set
		for (int k = 0; k < 6; ++k) act_host_init[k] = 0;
		act_host_init[3 + 2] = actf::MASK_SET_AFTER_LAST | actf::MASK_SET_BEFORE_LAST;
		if (CALHIP_CHECK(hipMemcpyAsync(act.L.ctr, act_host_init, sizeof act_host_init, hipMemcpyHostToDevice,
		                                m->stream)))
			return -1;
		act_gcur = 0;
		act_par = 0;
		act_rot = 0;
		act_last_mask = 2;
		actf::Launch p = act_launch_args(m);
		if (!(p.e1 = act_next_epoch(m))) return -1;
		if (actf::launch_setup(m->stream, p)) return -1;
		m->stats.launches++;
		act_list_reach = act.km;
		act_on = true;
		ctr_valid = false;
		return 1;
	}

	static int act_pick(int lim) { return lim >= 4 ? 4 : (lim >= 2 ? 2 : 1); }

	// n steps, each followed by the steering twin, in launches of K steps.
	// A launch lists the tiles the next launch needs (reach: that launch's
	// steps); the last one lists for as many steps as the grid allows, since
	// the next call usually runs a group of them.
	int act_steps(ModelX* m, int n)
	{
		const int grid = act.grid;
		while (n > 0) {
			const int K = act_pick(std::min(n, act_list_reach));
			const int Kn = act_pick(n > K ? std::min(n - K, act.km) : act.km);
			actf::Launch p = act_launch_args(m);
			p.km = Kn;
			if (!(p.e1 = act_next_epoch(m))) return -1;
			hipEvent_t ev;
			timer_begin(m, &ev);
			if (actf::launch_fused(m->stream, K, p, grid)) return -1;
			timer_end(m, ev);
			m->stats.launches++;
			// the written copies become the current ones
			flip(h);
			for (int q = 0; q < 4; ++q) flip(f[q]);
			std::swap(m->d_flags, act.flags_b);
			act_gcur = 1 - act_gcur;
			act_last_mask = act_rot;
			act_par = 1 - act_par;
			act_rot = (act_rot + 1) % 3;
			act_list_reach = Kn;
			n -= K;
		}
		set_flow_state(Transient::ZeroAtActive);
		m->flags_dirty = false;
		list_stale = true;
		h_prev_valid = false;
		return 0;
	}

	// Back to the ordinary representation: the flows' buffer holds their
	// logical value outside the set (ZeroAtActive), h its final bits.
	int act_leave(ModelX* m)
	{
		if (!act_on) return 0;
		double* F[4];
		for (int q = 0; q < 4; ++q) F[q] = (double*)f[q]->d_cur;
		if (actf::launch_leave(m->stream, (double*)h->d_cur, m->d_flags, act.gh[act_gcur], F,
		                       act.L.mask + act_last_mask, m->cells()))
			return -1;
		m->stats.launches++;
		act_on = false;
		ctr_valid = false;
		list_stale = true;
		return 0;
	}

	// Steps of calRunCAStep2D in the mode are launched in groups: a call
	// records its step and every fourth launches them together.  Anything
	// that can observe or change the state (every other hook of the
	// program, host reads, reductions, synchronisation, timing and
	// statistics queries) launches the recorded steps first.
	int act_flush(ModelX* m)
	{
		if (!act_pending) return 0;
		const int n = act_pending;
		act_pending = 0;
		if (active) return act_steps(m, n);
		// full grid: the steps the tile mode does not take run one at a time
		const int done = pl_run(m, n);
		if (done < 0) return -1;
		return done < n ? Program::run_steps(m, n - done, true) : 0;
	}

	int flush(ModelX* m) override { return act_flush(m); }

	int before_read(ModelX* m) override
	{
		if (act_flush(m)) return -1;
		if (!act_on) return 0;
		m->stats.launches++;
		return actf::launch_normalise(m->stream, (double*)h->d_cur, act.L.mask + act_last_mask, m->cells());
	}

	int step_and_steer(ModelX* m) override
	{
		if (active) {
			const int r = act_ready(m);
			if (r < 0) return -1;
			if (r) {
				list_stale = true;
				if (++act_pending >= actf::KMAX) return act_flush(m);
				return 0;
			}
		} else if (!act_pending) {
			const int r = pl_ready(m);
			if (r < 0) return -1;
			if (r) {
				++act_pending;
				return 0;
			}
		} else {
			if (++act_pending >= PL_KMAX) return act_flush(m);
			return 0;
		}
		return Program::step_and_steer(m);
	}

	// ---- full grid in launches of K steps ----

	bool pl_eligible(ModelX* m) const
	{
		if (active || pl_declined || getenv("CALHIP_SCIDDICAT_NO_TILES")) return false;
		if (z->nxt_dirty || h->nxt_dirty) return false;
		for (int q = 0; q < 4; ++q)
			if (f[q]->tstate != Transient::Zero || f[q]->nxt_dirty) return false;
		(void)m;
		return true;
	}

	// 1: the tile list is built and the mode pays; 0: not used; -1: error.
	int pl_ready(ModelX* m)
	{
		if (pl_valid) return 1;
		if (!pl_eligible(m)) return 0;
		const int rows = m->pub.rows, cols = m->pub.columns;
		if (!pl.block) {
			const int T = actf::TILE;
			pl.ty = (rows + T - 1) / T;
			pl.tx = (cols + T - 1) / T;
			pl.n_tiles = pl.ty * pl.tx;
			int km = PL_KMAX;
			if (pl.ty >= 3 && rows % T) km = std::min(km, rows % T);
			if (pl.tx >= 3 && cols % T) km = std::min(km, cols % T);
			pl.km = km;
			if (CALHIP_CHECK(hipDeviceGetAttribute(&pl.cus, hipDeviceAttributeMultiprocessorCount, m->device)))
				return -1;
			const size_t nt = (size_t)pl.n_tiles;
			const size_t bytes = 3 * nt * sizeof(unsigned) + 2 * nt * sizeof(int) + 4 * sizeof(int);
			if (CALHIP_CHECK(hipMalloc(&pl.block, bytes))) return -1;
			if (CALHIP_CHECK(hipMemsetAsync(pl.block, 0, bytes, m->stream))) return -1;
			pl.tag[0] = reinterpret_cast<unsigned*>(pl.block);
			pl.tag[1] = pl.tag[0] + nt;
			pl.list[0] = reinterpret_cast<int*>(pl.tag[1] + nt);
			pl.list[1] = pl.list[0] + nt;
			pl.cnt = pl.list[1] + nt;
			pl.same = reinterpret_cast<unsigned*>(pl.cnt + 4);
			if (CALHIP_CHECK(hipHostMalloc((void**)&pl.h_cnt, sizeof(int), hipHostMallocDefault))) return -1;
			if (CALHIP_CHECK(hipEventCreateWithFlags(&pl.ev, hipEventDisableTiming))) return -1;
			pl_e = 0;
		}
		if (pl_e >= 0xfffffff0u) {
			// the tags would wrap: every tile back to "never listed" (and, by
			// the setup below, to "no copy yet")
			if (CALHIP_CHECK(hipMemsetAsync(pl.tag[0], 0, 2 * (size_t)pl.n_tiles * sizeof(unsigned), m->stream)))
				return -1;
			pl_e = 0;
		}
		if (CALHIP_CHECK(hipMemsetAsync(pl.cnt, 0, 3 * sizeof(int), m->stream))) return -1;
		pl_L = 0;
		actf::PlainLaunch p = pl_args(m);
		p.e = ++pl_e;
		p.tag_out = pl.tag[0];
		p.list_out = pl.list[0];
		p.cnt_out = pl.cnt;
		p.km = pl.km;
		if (actf::launch_plain_setup(m->stream, p)) return -1;
		m->stats.launches++;
		if (CALHIP_CHECK(hipMemcpyAsync(pl.h_cnt, pl.cnt, sizeof(int), hipMemcpyDeviceToHost, m->stream)) ||
		    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
			return -1;
		pl_cnt_pending = false;
		// CALHIP_SCIDDICAT_TILES_ALWAYS keeps the mode on whatever the wet area
		const bool always = getenv("CALHIP_SCIDDICAT_TILES_ALWAYS") != nullptr;
		if (!always && *pl.h_cnt > pl.n_tiles / 8) {
			pl_declined = true;
			return 0;
		}
		pl_valid = true;
		pl_reach = pl.km;
		return 1;
	}

	actf::PlainLaunch pl_args(ModelX* m) const
	{
		actf::PlainLaunch p = {};
		p.z = (const double*)z->d_cur;
		p.h_in = (const double*)h->d_cur;
		p.h_out = (double*)h->d_nxt;
		p.tag_in = pl.tag[pl_L % 2];
		p.tag_out = pl.tag[(pl_L + 1) % 2];
		p.list_in = pl.list[pl_L % 2];
		p.list_out = pl.list[(pl_L + 1) % 2];
		p.cnt_in = pl.cnt + pl_L % 3;
		p.cnt_out = pl.cnt + (pl_L + 1) % 3;
		p.cnt_clear = pl.cnt + (pl_L + 2) % 3;
		p.same = pl.same;
		p.e = pl_e;
		p.rows = m->pub.rows;
		p.cols = m->pub.columns;
		p.ty = pl.ty;
		p.tx = pl.tx;
		p.n_tiles = pl.n_tiles;
		p.r = *p_r;
		p.eps = *p_eps;
		return p;
	}

	static int pl_pick(int lim) { return std::max(1, std::min(lim, PL_KMAX)); }

	// Up to n steps of the full grid (each followed by the steering twin);
	// returns the steps taken: fewer than n when the mode stops paying (too
	// many tiles near wet cells), -1 on error.
	int pl_run(ModelX* m, int n)
	{
		const int r0 = pl_ready(m);
		if (r0 <= 0) return r0 < 0 ? -1 : 0;
		int done = 0;
		while (done < n) {
			if (pl_cnt_pending && hipEventQuery(pl.ev) == hipSuccess) {
				pl_cnt_pending = false;
				if (*pl.h_cnt > pl.n_tiles / 4 && !getenv("CALHIP_SCIDDICAT_TILES_ALWAYS")) {
					pl_valid = false;
					pl_declined = true;
					break;
				}
			}
			const int K = pl_pick(std::min(n - done, pl_reach));
			const int Kn = pl_pick(n - done > K ? std::min(n - done - K, pl.km) : pl.km);
			actf::PlainLaunch p = pl_args(m);
			p.km = Kn;
			// compute blocks: two per CU (at K = 4: 1024 threads and 40 KiB of
			// LDS per whole tile, 576 threads and 23 KiB per quarter tile).
			// Quarter tiles while the list (its last count read back) leaves
			// a CU for each part: four blocks per tile then shorten the
			// launch; a longer list would pay for their halos without
			// gaining CUs.
			p.grid_comp = std::max(1, std::min(pl.n_tiles, pl.cus * 2));
			const int tpl = (actf::TILE / actf::PLAIN_PART) * (actf::TILE / actf::PLAIN_PART);
			const bool parts = pl_parts >= 0 ? pl_parts != 0 : *pl.h_cnt * tpl <= pl.cus;
			hipEvent_t ev;
			timer_begin(m, &ev);
			if (actf::launch_plain(m->stream, K, parts, p)) return -1;
			timer_end(m, ev);
			m->stats.launches++;
			flip(h);
			pl_L++;
			pl_e++;
			pl_reach = Kn;
			done += K;
			if ((pl_L & 7) == 0 && !pl_cnt_pending) {
				if (CALHIP_CHECK(hipMemcpyAsync(pl.h_cnt, pl.cnt + pl_L % 3, sizeof(int), hipMemcpyDeviceToHost,
				                                m->stream)) ||
				    CALHIP_CHECK(hipEventRecord(pl.ev, m->stream)))
					return -1;
				pl_cnt_pending = true;
			}
		}
		set_flow_state(Transient::Zero);
		h_prev_valid = false;
		return done;
	}

	int sync_active(ModelX* m) override
	{
		if (act_flush(m) || act_leave(m)) return -1;
		if (!list_stale) return 0;
		if (k::rebuild_active(m)) return -1;
		m->stats.rebuilds++;
		list_stale = false;
		return 0;
	}

	int launch_fused(ModelX* m, double r, double eps)
	{
		const int rows = m->pub.rows, cols = m->pub.columns;
		if (!d_trash && CALHIP_CHECK(hipMalloc(&d_trash, 128 * sizeof(double)))) return -1;
		// sweep shape: autotuned per grid shape (the sweeps it times write
		// into h's next layer, which this step overwrites)
		if (shape.rpw == 0)
			shape = sweep_shape(m->stream, rows, cols, (const double*)z->d_cur, (const double*)h->d_cur,
			                    (double*)h->d_nxt, d_trash, r, eps);
		launch_sweep(m->stream, rows, cols, (const double*)z->d_cur, (const double*)h->d_cur, (double*)h->d_nxt,
		             d_trash, r, eps, shape, &wet_order);
		return 0;
	}

	int check_entry(ModelX* m)
	{
		// flows reset by a next-layer-only init that no update followed are
		// written out; the step then starts from the exact reference state
		for (int q = 0; q < 4; ++q) {
			Transient t = f[q]->tstate;
			if (t == Transient::PendingZero || t == Transient::PendingActive ||
			    (!active && t == Transient::Derived))
				if (materialise(m, f[q])) return -1;
		}
		if (z->nxt_dirty || h->nxt_dirty) {
			set_error(10, "SciddicaT device program: z or h has next-layer writes pending at the start of a "
			              "step (call calUpdate2D first)");
			return -1;
		}
		for (int q = 0; q < 4; ++q)
			if (f[q]->nxt_dirty) flip(f[q]);   // the first update of the step copies them
		(void)m;
		return 0;
	}

	// ---- full grid ----
	int step_full(ModelX* m)
	{
		const double r = *p_r, eps = *p_eps;
		const int rows = m->pub.rows, cols = m->pub.columns;
		bool fused = true;
		for (int q = 0; q < 4; ++q) fused = fused && f[q]->tstate == Transient::Zero;
		if (fused) {
			hipEvent_t ev;
			timer_begin(m, &ev);
			if (launch_fused(m, r, eps)) return -1;
			timer_end(m, ev);
			m->stats.launches++;
			if (CALHIP_CHECK(hipGetLastError())) return -1;
			flip(h);
			h_prev_valid = true;
			set_flow_state(Transient::Derived);
			return 0;
		}
		// exact per-process sequence with the flows in memory
		for (int q = 0; q < 4; ++q)
			if (f[q]->tstate != Transient::None) {
				if (materialise(m, f[q])) return -1;
			}
		size_t n = m->cells();
		FlowPtrs fc = cur_flows(), fn;
		for (int q = 0; q < 4; ++q) fn.f[q] = (double*)f[q]->d_nxt;
		hipEvent_t ev;
		timer_begin(m, &ev);
		hipLaunchKernelGGL(flows_full, dim3(stride_grid(n, 256)), dim3(256), 0, m->stream, (const double*)z->d_cur,
		                   (const double*)h->d_cur, fc, fn, rows, cols, r, eps);
		for (int q = 0; q < 4; ++q) flip(f[q]);
		hipLaunchKernelGGL(width_full, dim3(stride_grid(n, 256)), dim3(256), 0, m->stream, (const double*)h->d_cur,
		                   cur_flows(), (double*)h->d_nxt, rows, cols);
		timer_end(m, ev);
		m->stats.launches += 2;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		flip(h);
		h_prev_valid = false;
		return 0;
	}

	// ---- active cells ----
	int step_active(ModelX* m)
	{
		// the host changed the flags since the last update: the list (not the
		// flags) is what this step's first process visits
		if (m->flags_dirty) {
			ctr_valid = false;
			return step_active_list(m);
		}
		const double r = *p_r, eps = *p_eps;
		const int rows = m->pub.rows, cols = m->pub.columns;
		if (!tiles.ctr) {
			tiles.ty = (rows + 3) / 4;
			tiles.tx = (cols + 63) / 64;
			tiles.n_tiles = tiles.ty * tiles.tx;
			const size_t nt = (size_t)tiles.n_tiles;
			const size_t bytes = 4 * sizeof(int) + 4 * nt * sizeof(int) + nt * sizeof(unsigned) + m->cells();
			char* b = nullptr;
			if (CALHIP_CHECK(hipMalloc(&b, bytes))) return -1;
			if (CALHIP_CHECK(hipMemsetAsync(b, 0, bytes, m->stream))) {
				hipFree(b);
				return -1;
			}
			tiles.ctr = reinterpret_cast<int*>(b);
			int* lists = tiles.ctr + 4;
			tiles.set_list[0] = lists;
			tiles.set_list[1] = lists + nt;
			tiles.x_list[0] = lists + 2 * nt;
			tiles.x_list[1] = lists + 3 * nt;
			tiles.tag = reinterpret_cast<unsigned*>(lists + 4 * nt);
			tiles.recv = reinterpret_cast<uint8_t*>(tiles.tag + nt);
			ctr_valid = false;
		}
		const int grid = std::min(tiles.n_tiles, 1024);
		hipEvent_t ev;
		timer_begin(m, &ev);
		if (tile_epoch >= 0x7ffffff0u) {
			// the tags (2e + 1) would wrap: start again from the zero state
			if (CALHIP_CHECK(hipMemsetAsync(tiles.tag, 0, (size_t)tiles.n_tiles * sizeof(unsigned), m->stream)))
				return -1;
			tile_epoch = 0;
			ctr_valid = false;
		}
		// a rebuilt list needs an epoch no tag holds yet: the last width
		// process tagged its tiles for (last epoch + 1)
		if (!ctr_valid) tile_epoch++;
		const unsigned e = ++tile_epoch;
		if (!ctr_valid) {
			// the flags are current: list the tiles that hold cells of the set
			if (CALHIP_CHECK(hipMemsetAsync(tiles.ctr, 0, 4 * sizeof(int), m->stream))) return -1;
			hipLaunchKernelGGL(tiles_from_flags, dim3(grid), dim3(256), 0, m->stream, (const uint8_t*)m->d_flags,
			                   tiles, parity, e, rows, cols);
			m->stats.launches++;
			ctr_valid = true;
		}
		int zero_dry = 1;
		for (int q = 0; q < 4; ++q)
			if (f[q]->tstate != Transient::ZeroAtActive && f[q]->tstate != Transient::Zero) zero_dry = 0;
		hipLaunchKernelGGL(flows_tiles, dim3(grid), dim3(256), 0, m->stream, (const double*)z->d_cur,
		                   (const double*)h->d_cur, cur_flows(), (const uint8_t*)m->d_flags, tiles, parity, e,
		                   rows, cols, r, eps, zero_dry);
		hipLaunchKernelGGL(width_tiles, dim3(grid), dim3(256), 0, m->stream, (double*)h->d_cur, cur_flows(),
		                   m->d_flags, tiles, parity, e, rows, cols, eps);
		timer_end(m, ev);
		m->stats.launches += 2;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		parity = 1 - parity;
		list_stale = true;
		m->flags_dirty = false;
		set_flow_state(Transient::DerivedActive);
		return 0;
	}

	// The same step over the row-major list, rebuilt after each process.
	int step_active_list(ModelX* m)
	{
		if (sync_active(m)) return -1;
		const double r = *p_r, eps = *p_eps;
		const int rows = m->pub.rows, cols = m->pub.columns;
		size_t n = m->cells();
		int zero_dry = 1;
		for (int q = 0; q < 4; ++q)
			if (f[q]->tstate != Transient::ZeroAtActive && f[q]->tstate != Transient::Zero) zero_dry = 0;
		int grid = stride_grid(n, 256);
		hipEvent_t ev;
		timer_begin(m, &ev);
		hipLaunchKernelGGL(flows_active, dim3(grid), dim3(256), 0, m->stream, (const double*)z->d_cur,
		                   (const double*)h->d_cur, cur_flows(), m->d_flags, m->d_cells, m->d_count, rows, cols, r,
		                   eps, zero_dry);
		m->flags_dirty = true;
		if (device_rebuild_active(m)) return -1;
		hipLaunchKernelGGL(width_active, dim3(grid), dim3(256), 0, m->stream, (const double*)h->d_cur,
		                   (double*)h->d_cur, cur_flows(), m->d_flags, m->d_cells, m->d_count, rows, cols, eps, 1);
		m->flags_dirty = true;
		if (device_rebuild_active(m)) return -1;
		timer_end(m, ev);
		m->stats.launches += 2;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		set_flow_state(Transient::DerivedActive);
		return 0;
	}

	int step(ModelX* m) override
	{
		if (act_flush(m) || act_leave(m)) return -1;
		pl_valid = false;
		if (check_entry(m)) return -1;
		return active ? step_active(m) : step_full(m);
	}

	// Device twin of the reference steering (sciddicaT.c:133-146; the active
	// variant resets the next layer only): the flows become logically zero.
	int steer(ModelX* m) override
	{
		if (act_flush(m) || act_leave(m)) return -1;
		for (int q = 0; q < 4; ++q)
			if (device_fill(m, f[q], 0.0, !active)) return -1;   // calInitSubstate(Next)2Dr(f[q], 0)
		return 0;
	}

	// Full grid with the steering twin: pairs of steps go to the two-step
	// sweep (the steering's reset between them is what makes a step a
	// function of z and h alone) when the tuner found it faster than two
	// one-step launches on this grid; the rest, one step at a time.
	int run_steps(ModelX* m, int n, bool with_steering) override
	{
		const int rows = m->pub.rows, cols = m->pub.columns;
		if (active && with_steering) {
			const int r = act_ready(m);
			if (r < 0) return -1;
			if (r) {
				n += act_pending;
				act_pending = 0;
				return act_steps(m, n);
			}
		}
		if (!active && with_steering) {
			n += act_pending;
			act_pending = 0;
			const int done = pl_run(m, n);
			if (done < 0) return -1;
			n -= done;
			if (n == 0) return 0;
		}
		if (act_flush(m)) return -1;
		if (!with_steering || active || !sweep2_supported(rows, cols) || getenv("CALHIP_SCIDDICAT_NO_2X"))
			return Program::run_steps(m, n, with_steering);
		while (n >= 2) {
			if (check_entry(m)) return -1;
			bool zero = true;
			for (int q = 0; q < 4; ++q) zero = zero && f[q]->tstate == Transient::Zero;
			if (!zero) break;
			const double r = *p_r, eps = *p_eps;
			if (!d_trash && CALHIP_CHECK(hipMalloc(&d_trash, 128 * sizeof(double)))) return -1;
			if (!shape2_known) {
				shape2 = sweep2_shape(m->stream, rows, cols, (const double*)z->d_cur, (const double*)h->d_cur,
				                      (double*)h->d_nxt, d_trash, r, eps);
				shape2_known = true;
			}
			if (shape2.s_rows == 0) break;   // one step per launch is faster on this grid
			hipEvent_t ev;
			timer_begin(m, &ev);
			launch_sweep2(m->stream, rows, cols, (const double*)z->d_cur, (const double*)h->d_cur,
			              (double*)h->d_nxt, d_trash, r, eps, shape2, &wet_order2);
			timer_end(m, ev);
			m->stats.launches++;
			if (CALHIP_CHECK(hipGetLastError())) return -1;
			flip(h);
			h_prev_valid = false;   // the flows stay zero: nothing to recover
			n -= 2;
		}
		return Program::run_steps(m, n, true);
	}

	// calApplyElementaryProcess2D (cal2D.c:717-734) with one of the bound
	// processes: that sweep alone, reading the current layers and writing the
	// next ones, over the grid or the active list (the whole grid while it is
	// empty); the caller's calUpdate2D then updates the written layers.  The
	// flows become ordinary substates first.
	int apply(ModelX* m, const void* ep) override
	{
		int which = -1;
		for (size_t k = 0; k < m->prog_eps.size(); ++k)
			if ((const void*)m->prog_eps[k] == ep) which = (int)k;
		if (which < 0) {
			set_error(12, "SciddicaT device program: calApplyElementaryProcess2D with a process that is not one "
			              "of the bound flows, width or remove");
			return -1;
		}
		if (act_flush(m) || act_leave(m)) return -1;
		pl_valid = false;
		for (int q = 0; q < 4; ++q)
			if (f[q]->tstate != Transient::None) {
				if (materialise(m, f[q])) return -1;
				m->stats.materialisations++;
			}
		h_prev_valid = false;
		const double r = *p_r, eps = *p_eps;
		const int rows = m->pub.rows, cols = m->pub.columns;
		const size_t n = m->cells();
		const int grid = stride_grid(n, 256);
		if (active) {
			if (sync_active(m)) return -1;   // the list of the last update
			ctr_valid = false;               // the tile lists are rebuilt from the flags
		}
		FlowPtrs fn;
		for (int q = 0; q < 4; ++q) fn.f[q] = (double*)f[q]->d_nxt;
		hipEvent_t ev;
		timer_begin(m, &ev);
		if (which == 0) {
			// flows: the next layer at wet cells; elsewhere it keeps its value
			// (equal to current unless an earlier process wrote it)
			bool stale = true;
			for (int q = 0; q < 4; ++q) stale = stale && f[q]->nxt_stale;
			if (active || !stale)
				for (int q = 0; q < 4; ++q)
					if (materialise_next(m, f[q])) return -1;
			if (active) {
				hipLaunchKernelGGL(flows_active, dim3(grid), dim3(256), 0, m->stream, (const double*)z->d_cur,
				                   (const double*)h->d_cur, fn, m->d_flags, m->d_cells, m->d_count, rows, cols, r,
				                   eps, 0);
				m->flags_dirty = true;   // receivers joined the set
			} else {
				// with stale next layers the kernel copies current into next at
				// dry cells; otherwise (fcur = fnxt) it leaves them
				hipLaunchKernelGGL(flows_full, dim3(grid), dim3(256), 0, m->stream, (const double*)z->d_cur,
				                   (const double*)h->d_cur, stale ? cur_flows() : fn, fn, rows, cols, r, eps);
			}
			for (int q = 0; q < 4; ++q) {
				f[q]->nxt_stale = false;
				f[q]->nxt_dirty = true;
			}
		} else if (which == 1) {
			if (active) {
				if (materialise_next(m, h)) return -1;
				hipLaunchKernelGGL(width_active, dim3(grid), dim3(256), 0, m->stream, (const double*)h->d_cur,
				                   (double*)h->d_nxt, cur_flows(), m->d_flags, m->d_cells, m->d_count, rows, cols,
				                   eps, 0);
			} else {
				hipLaunchKernelGGL(width_full, dim3(grid), dim3(256), 0, m->stream, (const double*)h->d_cur,
				                   cur_flows(), (double*)h->d_nxt, rows, cols);
			}
			h->nxt_stale = false;
			h->nxt_dirty = true;
		} else {
			hipLaunchKernelGGL(remove_active, dim3(grid), dim3(256), 0, m->stream, (const double*)h->d_cur,
			                   m->d_flags, m->d_cells, m->d_count, n, eps);
			m->flags_dirty = true;
		}
		timer_end(m, ev);
		m->stats.launches++;
		return CALHIP_CHECK(hipGetLastError());
	}

	int transient_init(ModelX* m, SubX* s, double v, bool both) override
	{
		if (act_flush(m)) return -1;
		if (v != 0.0 || std::signbit(v)) return 0;
		switch (s->tstate) {
		case Transient::Zero:
			return 1;
		case Transient::Derived:
		case Transient::PendingZero:
			s->tstate = both ? Transient::Zero : Transient::PendingZero;
			return 1;
		case Transient::ZeroAtActive:
			return 1;
		case Transient::DerivedActive:
		case Transient::PendingActive:
			s->tstate = both ? Transient::ZeroAtActive : Transient::PendingActive;
			return 1;
		default:
			return 0;
		}
	}

	void transient_update(ModelX*) override
	{
		for (int q = 0; q < 4; ++q) {
			if (f[q]->tstate == Transient::PendingZero) f[q]->tstate = Transient::Zero;
			if (f[q]->tstate == Transient::PendingActive) f[q]->tstate = Transient::ZeroAtActive;
		}
	}

	// Write the logical content of a transient flow substate into its buffers
	// and make it an ordinary substate.
	int materialise(ModelX* m, SubX* s) override
	{
		if (act_flush(m) || act_leave(m)) return -1;
		size_t n = m->cells();
		int grid = stride_grid(n, 256);
		switch (s->tstate) {
		case Transient::None:
			return 0;
		case Transient::Zero:
			for (int q = 0; q < 4; ++q)
				if (f[q] == s) {
					if (k::fill(m, s->d_cur, 8, 0.0, false)) return -1;
					s->nxt_stale = true;
					s->nxt_dirty = false;
				}
			s->tstate = Transient::None;
			return 0;
		case Transient::Derived:
		case Transient::PendingZero: {
			if (!h_prev_valid) {
				set_error(11, "SciddicaT: flows of the last fused step are no longer recoverable");
				return -1;
			}
			// recompute all four flows of the last step from z and the pre-step h
			FlowPtrs zero_src, dst;
			for (int q = 0; q < 4; ++q) {
				if (k::fill(m, f[q]->d_nxt, 8, 0.0, false)) return -1;
				zero_src.f[q] = (double*)f[q]->d_nxt;
				dst.f[q] = (double*)f[q]->d_cur;
			}
			// dry cells copy the (zero) source, wet cells get their flows
			hipLaunchKernelGGL(flows_full, dim3(grid), dim3(256), 0, m->stream, (const double*)z->d_cur,
			                   (const double*)h->d_nxt, zero_src, dst, m->pub.rows, m->pub.columns, *p_r, *p_eps);
			m->stats.launches++;
			if (CALHIP_CHECK(hipGetLastError())) return -1;
			for (int q = 0; q < 4; ++q) {
				bool pending = f[q]->tstate == Transient::PendingZero;
				f[q]->nxt_stale = !pending;
				f[q]->nxt_dirty = pending;   // next layer already zero
				f[q]->tstate = Transient::None;
			}
			return 0;
		}
		case Transient::ZeroAtActive:
		case Transient::DerivedActive:
		case Transient::PendingActive: {
			bool zero_cur = s->tstate == Transient::ZeroAtActive;
			bool zero_nxt = s->tstate != Transient::DerivedActive;
			for (int q = 0; q < 4; ++q) {
				if (f[q]->tstate == Transient::None) continue;
				if (zero_nxt && !zero_cur) {
					// next differs from current: next = current zeroed at the active cells
					if (k::copy(m, f[q]->d_nxt, f[q]->d_cur, n * 8)) return -1;
					f[q]->nxt_stale = false;
					f[q]->nxt_dirty = true;
				} else {
					f[q]->nxt_stale = true;
					f[q]->nxt_dirty = false;
				}
			}
			FlowPtrs tgt;
			for (int q = 0; q < 4; ++q) tgt.f[q] = (double*)(zero_cur ? f[q]->d_cur : f[q]->d_nxt);
			if (zero_nxt && sync_active(m)) return -1;
			if (zero_nxt) {
				hipLaunchKernelGGL(zero_at_active, dim3(grid), dim3(256), 0, m->stream, tgt, m->d_cells, m->d_count, n);
				m->stats.launches++;
			}
			for (int q = 0; q < 4; ++q) f[q]->tstate = Transient::None;
			return CALHIP_CHECK(hipGetLastError());
		}
		}
		return 0;
	}

	int before_modify(ModelX* m, SubX* s) override
	{
		if (act_flush(m) || act_leave(m)) return -1;
		if (s == h || s == z) {
			pl_valid = false;
			pl_declined = false;
		}
		if (s == h || s == z) {
			for (int q = 0; q < 4; ++q)
				if (f[q]->tstate == Transient::Derived || f[q]->tstate == Transient::PendingZero) {
					if (materialise(m, f[q])) return -1;
					break;
				}
			h_prev_valid = false;
		}
		return 0;
	}
};

}  // namespace calhip

using namespace calhip;

extern "C" int calhipBindSciddicaT2D(struct CALModel2D* ca2D, CALCallbackFunc2D flows, CALCallbackFunc2D width,
                                     CALCallbackFunc2D remove, struct CALSubstate2Dr* z, struct CALSubstate2Dr* h,
                                     struct CALSubstate2Dr* f[4], const CALreal* p_r, const CALreal* p_epsilon)
{
	ModelX* m = X(ca2D);
	if (ca2D->X_id != CAL_VON_NEUMANN_NEIGHBORHOOD_2D || ca2D->T != CAL_SPACE_TOROIDAL) {
		set_error(20, "calhipBindSciddicaT2D: the SciddicaT program needs a von Neumann, toroidal model");
		return -1;
	}
	if (ca2D->rows < 3 || ca2D->columns < 3) {
		set_error(23, "calhipBindSciddicaT2D: the grid must be at least 3 x 3");
		return -1;
	}
	if (!flows || !width || !z || !h || !f || !p_r || !p_epsilon) {
		set_error(21, "calhipBindSciddicaT2D: missing argument");
		return -1;
	}
	if (ca2D->OPTIMIZATION == CAL_OPT_ACTIVE_CELLS && !remove) {
		set_error(22, "calhipBindSciddicaT2D: the active-cells model needs its removal process");
		return -1;
	}
	auto* p = new SciddicaTProgram();
	p->z = S(z);
	p->h = S(h);
	for (int q = 0; q < 4; ++q) p->f[q] = S(f[q]);
	p->p_r = p_r;
	p->p_eps = p_epsilon;
	p->active = ca2D->OPTIMIZATION == CAL_OPT_ACTIVE_CELLS;
	delete m->prog;
	m->prog = p;
	m->prog_eps.clear();
	m->prog_steering = nullptr;
	m->prog_eps.push_back(flows);
	m->prog_eps.push_back(width);
	if (remove) m->prog_eps.push_back(remove);
	return 0;
}

extern "C" int calhipBindSciddicaTSteering2D(struct CALModel2D* ca2D, void (*steering)(struct CALModel2D*))
{
	ModelX* m = X(ca2D);
	if (!steering || !dynamic_cast<SciddicaTProgram*>(m->prog)) {
		set_error(24, "calhipBindSciddicaTSteering2D: bind the SciddicaT program (calhipBindSciddicaT2D) first "
		              "and give its steering callback");
		return -1;
	}
	m->prog_steering = steering;
	return 0;
}

extern "C" long long calhipSciddicaTCheckDivision(long long n, unsigned long long seed)
{
	if (n < 0) {
		set_error(25, "calhipSciddicaTCheckDivision: negative sample count");
		return -1;
	}
	unsigned long long* d_bad = nullptr;
	unsigned long long bad = 0;
	if (CALHIP_CHECK(hipMalloc(&d_bad, sizeof bad))) return -1;
	int rc = CALHIP_CHECK(hipMemset(d_bad, 0, sizeof bad));
	if (!rc) {
		hipLaunchKernelGGL(check_div_count, dim3(2048), dim3(256), 0, 0, n, seed, d_bad);
		rc = CALHIP_CHECK(hipGetLastError());
	}
	if (!rc) rc = CALHIP_CHECK(hipMemcpy(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost));
	hipFree(d_bad);
	return rc ? -1 : (long long)bad;
}
