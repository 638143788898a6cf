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
//   step_fused    full grid, both processes in one sweep.  Each wave owns a
//                 64-column strip (60 output columns plus a 2-column halo on
//                 each side) and walks down a chunk of rows with a three-row
//                 register window of z and h; horizontal neighbours come from
//                 the adjacent lanes.  The outflows of a cell are computed once,
//                 consumed by the width update of its four neighbours through
//                 lane shuffles (east/west) and the window (north/south), and
//                 never written to memory.  HBM traffic per cell: z and h read,
//                 h written = 24 B.
//   flows_full / width_full
//                 one process per sweep with the flows written out, the exact
//                 reference sequence for models whose flows are observed.
//   flows_active / width_active
//                 over the row-major active list (or the whole grid when it is
//                 empty, as OpenCAL does when A.cells == NULL).
//
// Arithmetic follows the reference operation by operation (no contraction:
// the library is built with -ffp-contract=off), so results are bit-identical
// to the CPU path.
#include <stdio.h>

#include <cmath>

#include "engine.hpp"
#include "device_common.hpp"

namespace calhip {

namespace {

// avg / cnt exactly as the reference divides (sciddicaT.c:72-73): one IEEE
// division for every count, which keeps the wave's lanes on one path.
__device__ __forceinline__ double div_count(double a, int cnt) { return a / (double)cnt; }

// Outflows of a wet cell (h0 > eps) with altitude z0 and neighbour levels
// u1..u4 (N, W, E, S) = z + h of each neighbour; keep = mask of the
// non-eliminated neighbours (bits 1..4) for the active-cell variant.
struct FlowOut {
	double f0, f1, f2, f3;
	unsigned keep;
};

__device__ __forceinline__ FlowOut flows_wet_body(double h0, double z0, double u1, double u2, double u3, double u4,
                                                  double r, double eps)
{
	const double m = h0 - eps;
	double u[5];
	u[0] = z0 + eps;
	u[1] = u1;
	u[2] = u2;
	u[3] = u3;
	u[4] = u4;
	unsigned elim = 0u;
	double avg;
	bool again;
	do {
		again = false;
		avg = m;
		int cnt = 0;
#pragma unroll
		for (int n = 0; n < 5; ++n)
			if (!((elim >> n) & 1u)) {
				avg += u[n];
				cnt++;
			}
		if (cnt != 0) avg = div_count(avg, cnt);
#pragma unroll
		for (int n = 0; n < 5; ++n)
			if (avg <= u[n] && !((elim >> n) & 1u)) {
				elim |= 1u << n;
				again = true;
			}
	} while (again);
	FlowOut o;
	o.f0 = (elim & 2u) ? 0.0 : (avg - u[1]) * r;
	o.f1 = (elim & 4u) ? 0.0 : (avg - u[2]) * r;
	o.f2 = (elim & 8u) ? 0.0 : (avg - u[3]) * r;
	o.f3 = (elim & 16u) ? 0.0 : (avg - u[4]) * r;
	o.keep = (~elim) & 0x1eu;
	return o;
}

// Out-of-line copy used by the streaming kernels: wet cells are rare, and
// keeping the elimination loop out of their unrolled row loop keeps that loop
// small enough to stay in the instruction cache.
__device__ __noinline__ FlowOut flows_wet(double h0, double z0, double u1, double u2, double u3, double u4,
                                          double r, double eps)
{
	return flows_wet_body(h0, z0, u1, u2, u3, u4, r, eps);
}

// Outflows of two cells at once (the two cells of a lane in the two-column
// kernels): one loop runs both minimisations until both have converged, which
// halves the code and interleaves two independent dependency chains.
struct PairFlows {
	double a[4], b[4];
};

__device__ __forceinline__ void avg_pass(const double u[5], double m, unsigned& elim, double& avg, bool& again)
{
	again = false;
	avg = m;
	int cnt = 0;
#pragma unroll
	for (int n = 0; n < 5; ++n)
		if (!((elim >> n) & 1u)) {
			avg += u[n];
			cnt++;
		}
	if (cnt != 0) avg = div_count(avg, cnt);
#pragma unroll
	for (int n = 0; n < 5; ++n)
		if (avg <= u[n] && !((elim >> n) & 1u)) {
			elim |= 1u << n;
			again = true;
		}
}

__device__ __forceinline__ void pair_flows(double ha, double za, double ua1, double ua2, double ua3, double ua4,
                                           double hb, double zb, double ub1, double ub2, double ub3, double ub4,
                                           double r, double eps, PairFlows& o)
{
#pragma unroll
	for (int q = 0; q < 4; ++q) o.a[q] = o.b[q] = 0.0;
	bool wa = !(ha <= eps), wb = !(hb <= eps);
	if (!(wa || wb)) return;
	const double ma = ha - eps, mb = hb - eps;
	double u_a[5] = {za + eps, ua1, ua2, ua3, ua4};
	double u_b[5] = {zb + eps, ub1, ub2, ub3, ub4};
	unsigned ea = 0u, eb = 0u;
	double avg_a = 0.0, avg_b = 0.0;
	bool ga = wa, gb = wb;
	// A pass after convergence changes nothing (same set, same average, no
	// new elimination), so both cells run every pass: straight-line code
	// with two independent chains.
	while (ga || gb) {
		bool na, nb;
		avg_pass(u_a, ma, ea, avg_a, na);
		avg_pass(u_b, mb, eb, avg_b, nb);
		ga = ga && na;
		gb = gb && nb;
	}
	if (wa) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.a[n - 1] = ((ea >> n) & 1u) ? 0.0 : (avg_a - u_a[n]) * r;
	}
	if (wb) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.b[n - 1] = ((eb >> n) & 1u) ? 0.0 : (avg_b - u_b[n]) * r;
	}
}

// Outflows f[0..3] (towards N, W, E, S) of any cell: zero when h0 <= eps
// (the process returns before writing them; the steering keeps them zero).
template <bool OUTLINE>
__device__ __forceinline__ unsigned cell_flows_t(double h0, double z0, double u1, double u2, double u3, double u4,
                                                 double r, double eps, double f[4])
{
	f[0] = f[1] = f[2] = f[3] = 0.0;
	if (h0 <= eps) return 0u;
	FlowOut o = OUTLINE ? flows_wet(h0, z0, u1, u2, u3, u4, r, eps)
	                    : flows_wet_body(h0, z0, u1, u2, u3, u4, r, eps);
	f[0] = o.f0;
	f[1] = o.f1;
	f[2] = o.f2;
	f[3] = o.f3;
	return o.keep;
}

__device__ __forceinline__ unsigned cell_flows(double h0, double z0, double u1, double u2, double u3, double u4,
                                               double r, double eps, double f[4])
{
	return cell_flows_t<false>(h0, z0, u1, u2, u3, u4, r, eps, f);
}

constexpr int FUSED_WAVES = 4;   // waves per block
constexpr int STRIP_OUT = 60;    // output columns per wave

// Full-grid fused step (see file header).  Block = FUSED_WAVES waves stacked
// vertically on one column strip; blocks are remapped so each XCD gets a
// contiguous band of the grid.
__global__ __launch_bounds__(64 * FUSED_WAVES) void step_fused(
    const double* __restrict__ z, const double* __restrict__ h, double* __restrict__ hn, int rows, int cols,
    int n_strips, int n_logical, int per_xcd, int rpw, double r, double eps)
{
	const int L = dev::xcd_logical_block(blockIdx.x, per_xcd);
	if (L >= n_logical) return;
	const int strip = L % n_strips;
	const int group = L / n_strips;
	const int lane = threadIdx.x & 63;
	const int chunk = group * FUSED_WAVES + (threadIdx.x >> 6);
	const int r0 = chunk * rpw;
	if (r0 >= rows) return;
	const int r1 = min(r0 + rpw, rows);

	const int jj = strip * STRIP_OUT - 2 + lane;           // unwrapped column
	const int j = ((jj % cols) + cols) % cols;
	const bool out_lane = lane >= 2 && lane < 2 + STRIP_OUT && jj < cols;

	// window: rows k-1 (zm,hm), k (zc,hc), k+1 (zs,hs)
	int rm = dev::wrap1(r0 - 2, rows), rc = dev::wrap1(r0 - 1, rows);
	double zm = z[(size_t)rm * cols + j], hm = h[(size_t)rm * cols + j];
	double zc = z[(size_t)rc * cols + j], hc = h[(size_t)rc * cols + j];
	int rs = dev::wrap1(r0, rows);
	double zs = z[(size_t)rs * cols + j], hs = h[(size_t)rs * cols + j];

	double fp0 = 0.0, fp1 = 0.0, fp2 = 0.0, fp3 = 0.0;  // f(k-1)
	double fpp3 = 0.0;                                   // f(k-2).S
	double hp = 0.0;                                     // h(k-1)

	for (int k = r0 - 1; k <= r1; ++k) {
		// prefetch row k+2
		int rn = dev::wrap1(k + 2, rows);
		double zn = 0.0, hn2 = 0.0;
		if (k + 2 <= r1 + 1) {
			zn = z[(size_t)rn * cols + j];
			hn2 = h[(size_t)rn * cols + j];
		}
		const double uc = zc + hc;
		const double uw = dev::from_left(uc);
		const double ue = dev::from_right(uc);
		double f[4];
		cell_flows(hc, zc, zm + hm, uw, ue, zs + hs, r, eps, f);
		if (k - 1 >= r0) {   // wave-uniform
			const double f2w = dev::from_left(fp2);
			const double f1e = dev::from_right(fp1);
			double o = hp;
			o += fpp3 - fp0;
			o += f2w - fp1;
			o += f1e - fp2;
			o += f[0] - fp3;
			if (out_lane) hn[(size_t)(k - 1) * cols + j] = o;
		}
		fpp3 = fp3;
		fp0 = f[0];
		fp1 = f[1];
		fp2 = f[2];
		fp3 = f[3];
		hp = hc;
		zm = zc;
		hm = hc;
		zc = zs;
		hc = hs;
		zs = zn;
		hs = hn2;
	}
}

// Two columns per lane: 16-byte loads and stores, 128-column strips with 124
// output columns.  Rows live in a circular register buffer of W = D + 4 rows
// and the row loop is unrolled W times, so every row keeps its registers for
// its whole lifetime: the loop has no register copies of in-flight loads and
// the compiler's vmcnt waits cover only the row being consumed, D rows after
// its load was issued; row stores stay in flight meanwhile (on gfx9-class
// GPUs vmcnt counts stores too).  Loads are unconditional (rows past the
// chunk are clamped to its last halo row) to keep the body straight-line.
// Needs an even column count (a lane's pair never straddles the wrap).
constexpr int STRIP2_OUT = 124;

template <int D>
__global__ __launch_bounds__(64 * FUSED_WAVES) void step_fused2(
    const double* __restrict__ z, const double* __restrict__ h, double* __restrict__ hn, double* __restrict__ trash,
    int rows, int cols, int n_strips, int n_logical, int per_xcd, int rpw, double r, double eps)
{
	constexpr int W = D + 4;
	const int L = dev::xcd_logical_block(blockIdx.x, per_xcd);
	if (L >= n_logical) return;
	// a block is FUSED_WAVES horizontally adjacent strips of one row chunk (their
	// halo columns overlap in the CU's cache); each XCD gets a band of chunks
	const int n_sgroups = (n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
	const int strip = (L % n_sgroups) * FUSED_WAVES + (threadIdx.x >> 6);
	const int chunk = L / n_sgroups;
	if (strip >= n_strips) return;
	const int lane = threadIdx.x & 63;
	const int r0 = chunk * rpw;
	if (r0 >= rows) return;
	const int r1 = min(r0 + rpw, rows);
	const int last_row = r1 + 1;   // last halo row needed (unwrapped)

	const int jj = strip * STRIP2_OUT - 2 + 2 * lane;   // unwrapped column of the left element
	const int j = ((jj % cols) + cols) % cols;           // even, j+1 < cols
	const bool out_lane = lane >= 1 && lane <= 62 && jj < cols;

	auto ld = [&](const double* a, int row) -> double2 {
		row = dev::wrap1(min(row, last_row), rows);
		return *reinterpret_cast<const double2*>(a + (size_t)row * cols + j);
	};

	// slot of row x: (x - (r0 - 2)) mod W; rows r0-2 .. r0+D preloaded
	double2 bz[W], bh[W];
#pragma unroll
	for (int q = 0; q < W - 1; ++q) {
		bz[q] = ld(z, r0 - 2 + q);
		bh[q] = ld(h, r0 - 2 + q);
		__builtin_amdgcn_sched_barrier(0);   // issue in the order the rows are used
	}
	bz[W - 1] = make_double2(0.0, 0.0);
	bh[W - 1] = make_double2(0.0, 0.0);

	// flows of row k-1 for both elements (a = left, b = right), f(k-2).S
	double pa0 = 0.0, pa1 = 0.0, pa2 = 0.0, pa3 = 0.0, pb0 = 0.0, pb1 = 0.0, pb2 = 0.0, pb3 = 0.0;
	double ppa3 = 0.0, ppb3 = 0.0;

	for (int base = r0 - 1; base <= r1; base += W) {
#pragma unroll
		for (int d = 0; d < W; ++d) {
			const int k = base + d;
			const int sm = d % W, sc = (d + 1) % W, ss = (d + 2) % W, sp = (d + W - 1) % W;
			// prefetch row k+2+D into the slot of row k-2 (no longer needed)
			bz[sp] = ld(z, k + 2 + D);
			bh[sp] = ld(h, k + 2 + D);
			const double2 zm = bz[sm], hm = bh[sm], zc = bz[sc], hc = bh[sc], zs = bz[ss], hs = bh[ss];
			const double ua = zc.x + hc.x, ub = zc.y + hc.y;
			const double uw = dev::from_left(ub);    // west of a: right element of lane-1
			const double ue = dev::from_right(ua);   // east of b: left element of lane+1
			PairFlows pf;
			pair_flows(hc.x, zc.x, zm.x + hm.x, uw, ub, zs.x + hs.x, hc.y, zc.y, zm.y + hm.y, ua, ue, zs.y + hs.y,
			           r, eps, pf);
			const double* fa = pf.a;
			const double* fb = pf.b;
			const double f2w = dev::from_left(pb2);   // W neighbour of a: b of lane-1, its E outflow
			const double f1e = dev::from_right(pa1);  // E neighbour of b: a of lane+1, its W outflow
			double oa = hm.x, ob = hm.y;             // h of row k-1
			oa += ppa3 - pa0;
			oa += f2w - pa1;
			oa += pb1 - pa2;
			oa += fa[0] - pa3;
			ob += ppb3 - pb0;
			ob += pa2 - pb1;
			ob += f1e - pb2;
			ob += fb[0] - pb3;
			// unconditional store (lanes and rows that own no output write to a
			// per-lane slot of the trash row) keeps the body branch-free
			const bool st = out_lane && k - 1 >= r0 && k <= r1;
			double* dst = st ? hn + (size_t)(k - 1) * cols + j : trash + 2 * lane;
			*reinterpret_cast<double2*>(dst) = make_double2(oa, ob);
			ppa3 = pa3;
			ppb3 = pb3;
			pa0 = fa[0];
			pa1 = fa[1];
			pa2 = fa[2];
			pa3 = fa[3];
			pb0 = fb[0];
			pb1 = fb[1];
			pb2 = fb[2];
			pb3 = fb[3];
			// keep each row's prefetch where it is written (the scheduler
			// would otherwise sink loads towards their use)
			__builtin_amdgcn_sched_barrier(0);
		}
	}
}

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

// width over the active list, in place (the process reads only its own h),
// followed by the removal of cells with h <= eps.
__global__ void width_active(double* __restrict__ h, FlowPtrs f, uint8_t* __restrict__ flags,
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
		double o = h[c];
		o += f.f[3][cn] - f.f[0][c];
		o += f.f[2][cw] - f.f[1][c];
		o += f.f[1][ce] - f.f[2][c];
		o += f.f[0][cs] - f.f[3][c];
		h[c] = o;
		if (remove && o <= eps) flags[c] = 0;
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

int stride_grid(size_t n, int block)
{
	size_t b = (n + block - 1) / block;
	return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

}  // namespace

// Launch geometry of the fused kernel for a rows x cols grid.
struct FusedGeom {
	int n_strips, n_groups, n_logical, per_xcd, blocks, rpw;
};

static FusedGeom fused_geom(int rows, int cols, int rpw, int strip_out)
{
	FusedGeom g;
	g.rpw = rpw;
	g.n_strips = (cols + strip_out - 1) / strip_out;
	int chunks = (rows + rpw - 1) / rpw;
	g.n_groups = (chunks + FUSED_WAVES - 1) / FUSED_WAVES;
	g.n_logical = g.n_strips * g.n_groups;
	g.per_xcd = (g.n_logical + 7) / 8;
	g.blocks = g.per_xcd * 8;
	return g;
}

// Two-column kernel: a logical block is FUSED_WAVES adjacent strips of one
// row chunk.
static FusedGeom fused_geom2(int rows, int cols, int rpw)
{
	FusedGeom g;
	g.rpw = rpw;
	g.n_strips = (cols + STRIP2_OUT - 1) / STRIP2_OUT;
	int sgroups = (g.n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
	g.n_groups = (rows + rpw - 1) / rpw;
	g.n_logical = sgroups * g.n_groups;
	g.per_xcd = (g.n_logical + 7) / 8;
	g.blocks = g.per_xcd * 8;
	return g;
}

// Rows per wave: the smallest chunk height (>= 16) for which the whole grid is
// covered by one round of resident blocks (blocks_resident), so no block
// starts after the first ones finish.
static int pick_rows_per_wave(int rows, int cols, int strip_out, int blocks_resident)
{
	for (int rpw = 16; rpw < 4096; ++rpw) {
		FusedGeom g = fused_geom(rows, cols, rpw, strip_out);
		if (g.n_logical <= blocks_resident) return rpw;
	}
	return 4096;
}

// ------------------------------------------------------------------ program

struct SciddicaTProgram : Program {
	SubX* z = nullptr;
	SubX* h = nullptr;
	SubX* f[4] = {nullptr, nullptr, nullptr, nullptr};
	const CALreal* p_r = nullptr;
	const CALreal* p_eps = nullptr;
	bool active = false;
	int rpw_override = 0;
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
		return 0;
	}

	// fused sweep launch: two columns per lane when the column count is even
	int prefetch = 1;
	int resident_blocks = 0;
	double* d_trash = nullptr;   // sink for the stores of halo lanes

	~SciddicaTProgram() override
	{
		if (d_trash) hipFree(d_trash);
	}

	template <typename K>
	int resident(K kernel)
	{
		if (resident_blocks > 0) return resident_blocks;
		int dev = 0, cus = 0, per_cu = 0;
		if (CALHIP_CHECK(hipGetDevice(&dev))) return -1;
		if (CALHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev))) return -1;
		if (CALHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * FUSED_WAVES, 0)))
			return -1;
		resident_blocks = cus * (per_cu > 0 ? per_cu : 1);
		return resident_blocks;
	}

	template <int D>
	int launch2(ModelX* m, double r, double eps)
	{
		const int rows = m->pub.rows, cols = m->pub.columns;
		// short row chunks: the few waves that meet wet cells run the
		// minimisation on every row of their chunk, so long chunks make them
		// the critical path of the sweep
		int rpw = rpw_override > 0 ? rpw_override : 8;
		FusedGeom g = fused_geom2(rows, cols, rpw);
		if (!d_trash && CALHIP_CHECK(hipMalloc(&d_trash, 128 * sizeof(double)))) return -1;
		hipLaunchKernelGGL(step_fused2<D>, dim3(g.blocks), dim3(64 * FUSED_WAVES), 0, m->stream,
		                   (const double*)z->d_cur, (const double*)h->d_cur, (double*)h->d_nxt, d_trash, rows,
		                   cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps);
		return 0;
	}

	int launch_fused(ModelX* m, double r, double eps)
	{
		const int rows = m->pub.rows, cols = m->pub.columns;
		if (cols % 2 == 0 && prefetch > 0) {
			switch (prefetch) {
			case 1: return launch2<1>(m, r, eps);
			case 2: return launch2<2>(m, r, eps);
			case 3: return launch2<3>(m, r, eps);
			case 6: return launch2<6>(m, r, eps);
			default: return launch2<4>(m, r, eps);
			}
		}
		int rb = resident(step_fused);
		if (rb < 0) return -1;
		int rpw = rpw_override > 0 ? rpw_override : pick_rows_per_wave(rows, cols, STRIP_OUT, rb);
		FusedGeom g = fused_geom(rows, cols, rpw, STRIP_OUT);
		hipLaunchKernelGGL(step_fused, dim3(g.blocks), dim3(64 * FUSED_WAVES), 0, m->stream,
		                   (const double*)z->d_cur, (const double*)h->d_cur, (double*)h->d_nxt, rows, cols,
		                   g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps);
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
		hipLaunchKernelGGL(width_active, dim3(grid), dim3(256), 0, m->stream, (double*)h->d_cur, cur_flows(),
		                   m->d_flags, m->d_cells, m->d_count, rows, cols, eps, 1);
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
		if (check_entry(m)) return -1;
		return active ? step_active(m) : step_full(m);
	}

	int transient_init(ModelX* m, SubX* s, double v, bool both) override
	{
		(void)m;
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
	if (const char* e = getenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE")) p->rpw_override = atoi(e);
	if (const char* e = getenv("CALHIP_SCIDDICAT_PREFETCH")) p->prefetch = atoi(e);
	delete m->prog;
	m->prog = p;
	m->prog_eps.clear();
	m->prog_eps.push_back(flows);
	m->prog_eps.push_back(width);
	if (remove) m->prog_eps.push_back(remove);
	return 0;
}
