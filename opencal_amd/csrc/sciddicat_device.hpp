// SciddicaT device code shared by the single-model program, the batched
// ensemble and the adjoint: the outflow minimisation and the fused
// full-grid sweep of one strip (see sciddicat.hip for the reference mapping).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace calhip {
namespace sct {

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
__device__ __forceinline__ unsigned cell_flows(double h0, double z0, double u1, double u2, double u3, double u4,
                                               double r, double eps, double f[4])
{
	f[0] = f[1] = f[2] = f[3] = 0.0;
	if (h0 <= eps) return 0u;
	FlowOut o = flows_wet_body(h0, z0, u1, u2, u3, u4, r, eps);
	f[0] = o.f0;
	f[1] = o.f1;
	f[2] = o.f2;
	f[3] = o.f3;
	return o.keep;
}

constexpr int FUSED_WAVES = 4;   // waves per block

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

// One wave sweeps rows [r0, r1) of the 128-column strip `strip` of a
// rows x cols toroidal grid.  EVEN: cols is even, so a lane's two cells are
// one aligned 16-byte access; otherwise each cell is loaded and stored on its
// own (the pair may straddle the wrap).  UP: walk the chunk bottom-up.
// Neighbouring chunks walk in opposite directions, so the halo rows they share
// are loaded by both at about the same time and the second load hits L2.
//
// Rows are consumed as a stream: position s is row r0-2+s walking down,
// r1+1-s walking up; iteration q computes the flows of position q+1 from
// positions q..q+2 and writes row(q) for q >= 2 (its flows, those of the row
// behind it, q-1, and those of the row ahead of it, q+1, are then known).
template <int D, bool EVEN, bool UP>
__device__ __forceinline__ void sweep_strip(const double* __restrict__ z, const double* __restrict__ h,
                                            double* __restrict__ hn, double* __restrict__ trash, int rows, int cols,
                                            int strip, int r0, int r1, double r, double eps)
{
	constexpr int W = D + 4;
	const int lane = threadIdx.x & 63;
	const int n_it = r1 - r0 + 2;   // flows computed for rows r0-1 .. r1
	const int last_pos = n_it + 1;
	const int jj = strip * STRIP2_OUT - 2 + 2 * lane;   // unwrapped column of the left element
	const int j = ((jj % cols) + cols) % cols;
	const int j2 = j + 1 == cols ? 0 : j + 1;
	const bool out_a = lane >= 1 && lane <= 62 && jj < cols;
	const bool out_b = lane >= 1 && lane <= 62 && jj + 1 < cols;

	auto row_of = [&](int pos) -> int { return UP ? r1 + 1 - pos : r0 - 2 + pos; };
	auto ld = [&](const double* a, int pos) -> double2 {
		const int row = dev::wrap1(row_of(min(pos, last_pos)), rows);
		const double* p = a + (size_t)row * cols;
		if (EVEN) return *reinterpret_cast<const double2*>(p + j);
		return make_double2(p[j], p[j2]);
	};

	// slot of stream position s: s mod W; positions 0 .. W-2 preloaded
	double2 bz[W], bh[W];
#pragma unroll
	for (int q = 0; q < W - 1; ++q) {
		bz[q] = ld(z, q);
		bh[q] = ld(h, q);
		__builtin_amdgcn_sched_barrier(0);   // issue in the order the rows are used
	}
	bz[W - 1] = make_double2(0.0, 0.0);
	bh[W - 1] = make_double2(0.0, 0.0);

	// flows of the row behind (position q) for both elements (a = left,
	// b = right), and the one component of position q-1 that row(q) needs:
	// the outflow towards row(q) (S outflow walking down, N outflow walking up)
	double pa0 = 0.0, pa1 = 0.0, pa2 = 0.0, pa3 = 0.0, pb0 = 0.0, pb1 = 0.0, pb2 = 0.0, pb3 = 0.0;
	double ppa = 0.0, ppb = 0.0;

	for (int base = 0; base < n_it; base += W) {
#pragma unroll
		for (int d = 0; d < W; ++d) {
			const int q = base + d;
			const int sb = d % W, sc = (d + 1) % W, sa = (d + 2) % W, sp = (d + W - 1) % W;
			// prefetch position q+W-1 into the slot of position q-1 (no longer needed)
			bz[sp] = ld(z, q + W - 1);
			bh[sp] = ld(h, q + W - 1);
			// behind / current / ahead rows; north is behind when walking down
			const double2 zb = bz[sb], hb = bh[sb], zc = bz[sc], hc = bh[sc], za = bz[sa], ha = bh[sa];
			const double2 zN = UP ? za : zb, hN = UP ? ha : hb, zS = UP ? zb : za, hS = UP ? hb : ha;
			const double ua = zc.x + hc.x, ub = zc.y + hc.y;
			const double uw = dev::from_left(ub);    // west of a: right element of lane-1
			const double ue = dev::from_right(ua);   // east of b: left element of lane+1
			PairFlows pf;
			pair_flows(hc.x, zc.x, zN.x + hN.x, uw, ub, zS.x + hS.x, hc.y, zc.y, zN.y + hN.y, ua, ue, zS.y + hS.y,
			           r, eps, pf);
			const double* fa = pf.a;
			const double* fb = pf.b;
			// row(q) = the row behind: its north / south neighbour terms
			const double nA = UP ? fa[3] : ppa, sA = UP ? ppa : fa[0];
			const double nB = UP ? fb[3] : ppb, sB = UP ? ppb : fb[0];
			const double f2w = dev::from_left(pb2);   // W neighbour of a: b of lane-1, its E outflow
			const double f1e = dev::from_right(pa1);  // E neighbour of b: a of lane+1, its W outflow
			double oa = hb.x, ob = hb.y;             // h of row(q)
			oa += nA - pa0;
			oa += f2w - pa1;
			oa += pb1 - pa2;
			oa += sA - pa3;
			ob += nB - pb0;
			ob += pa2 - pb1;
			ob += f1e - pb2;
			ob += sB - pb3;
			// unconditional stores (lanes and rows that own no output write to
			// a per-lane slot of the trash row) keep the body branch-free
			const bool row_ok = q >= 2 && q < n_it;
			double* orow = hn + (size_t)row_of(q) * cols;
			if (EVEN) {
				double* dst = (row_ok && out_a) ? orow + j : trash + 2 * lane;
				*reinterpret_cast<double2*>(dst) = make_double2(oa, ob);
			} else {
				*((row_ok && out_a) ? orow + j : trash + 2 * lane) = oa;
				*((row_ok && out_b) ? orow + j2 : trash + 2 * lane + 1) = ob;
			}
			ppa = UP ? pa0 : pa3;
			ppb = UP ? pb0 : pb3;
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

// Both walking directions; chunks alternate.
template <int D, bool EVEN>
__device__ __forceinline__ void sweep_chunk(const double* __restrict__ z, const double* __restrict__ h,
                                            double* __restrict__ hn, double* __restrict__ trash, int rows, int cols,
                                            int strip, int chunk, int r0, int r1, double r, double eps)
{
	if (chunk & 1)
		sweep_strip<D, EVEN, true>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps);
	else
		sweep_strip<D, EVEN, false>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps);
}

template <int D, bool EVEN>
__global__ __launch_bounds__(64 * FUSED_WAVES) void step_fused2(
    const double* __restrict__ z, const double* __restrict__ h, double* __restrict__ hn, double* __restrict__ trash,
    int rows, int cols, int n_strips, int n_logical, int per_xcd, int rpw, double r, double eps)
{
	const int L = dev::xcd_logical_block(blockIdx.x, per_xcd);
	if (L >= n_logical) return;
	// a block is FUSED_WAVES horizontally adjacent strips of one row chunk (their
	// halo columns overlap in the CU's cache); each XCD gets a band of chunks
	const int n_sgroups = (n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
	const int strip = (L % n_sgroups) * FUSED_WAVES + (threadIdx.x >> 6);
	const int chunk = L / n_sgroups;
	if (strip >= n_strips) return;
	const int r0 = chunk * rpw;
	if (r0 >= rows) return;
	sweep_chunk<D, EVEN>(z, h, hn, trash, rows, cols, strip, chunk, r0, min(r0 + rpw, rows), r, eps);
}

// Launch shape of step_fused2 for a rows x cols grid with row chunks of rpw
// (sweep_tune.hip).
struct SweepGeom {
	int n_strips, n_logical, per_xcd, blocks, rpw;
};
SweepGeom sweep_geom(int rows, int cols, int rpw);
// One fused step h -> hn (prefetch depth 1) on stream st.
void launch_sweep(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                  double r, double eps, int rpw);
// The row-chunk length for this grid shape: the environment's
// CALHIP_SCIDDICAT_ROWS_PER_WAVE, else the autotuned one (timed on first use
// by writing sweeps of h into `scratch`, then cached per shape).
int sweep_rpw(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch, double* trash,
              double r, double eps);

}  // namespace sct
}  // namespace calhip

namespace calhip {
namespace sct {

// Forward state of a cell's minimisation, kept for the adjoint: wet flag,
// eliminated mask (bit n = neighbour n, bit 0 = the cell itself), count of
// the remaining levels, their average and the levels u[0..4].
struct FlowState {
	double u[5];
	double avg;
	int cnt;
	unsigned elim;
	bool wet;
};

__device__ __forceinline__ void flow_state(double h0, double z0, double u1, double u2, double u3, double u4,
                                           double eps, FlowState& s)
{
	s.wet = !(h0 <= eps);
	s.elim = 0u;
	s.cnt = 0;
	s.avg = 0.0;
	if (!s.wet) return;
	const double m = h0 - eps;
	s.u[0] = z0 + eps;
	s.u[1] = u1;
	s.u[2] = u2;
	s.u[3] = u3;
	s.u[4] = u4;
	bool again;
	do {
		again = false;
		double avg = m;
		int cnt = 0;
#pragma unroll
		for (int n = 0; n < 5; ++n)
			if (!((s.elim >> n) & 1u)) {
				avg += s.u[n];
				cnt++;
			}
		if (cnt != 0) avg = div_count(avg, cnt);
#pragma unroll
		for (int n = 0; n < 5; ++n)
			if (avg <= s.u[n] && !((s.elim >> n) & 1u)) {
				s.elim |= 1u << n;
				again = true;
			}
		s.avg = avg;
		s.cnt = cnt;
	} while (again);
}

}  // namespace sct
}  // namespace calhip
