// SciddicaT device code shared by the single-model program, the batched
// ensemble and the adjoint: the outflow minimisation and the fused
// full-grid sweep of one strip (see sciddicat.hip for the reference mapping).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"

namespace calhip {
namespace sct {

// a / cnt for cnt = 0..5 (0 leaves a unchanged), bit-equal to the IEEE
// division of the reference (sciddicaT.c:72-73) for every double a, in a
// chain of one multiply and two FMAs instead of the division's scale /
// reciprocal / Newton / fix-up sequence.  y is the hardware reciprocal of
// b = cnt refined by one Newton step, within 2^-40 of 1/b (checked on the
// device for b = 1..5 by calhipSciddicaTCheckDivision).  Then q = RN(a*y) is
// within 2^13 ulp of Q = a/cnt, so rem = a - b*q is a multiple of q's ulp of
// at most 18 bits and the FMA gets it exactly, and q + rem*y differs from Q
// by less than 2^-27 ulp.  Q sits at a multiple of 1/cnt of the ulp grid, at
// least a tenth of an ulp from every rounding midpoint (or on the grid), so
// rounding q + rem*y once gives RN(Q) (Markstein's correction step; the same
// holds on the subnormal grid).  rem == 0: q is exact (and keeps the sign of
// a zero a); rem NaN: a is infinite or NaN and q is its quotient.
__device__ __forceinline__ double count_reciprocal(double b)
{
	const double y0 = __builtin_amdgcn_rcp(b);
	return __builtin_fma(y0, __builtin_fma(-b, y0, 1.0), y0);
}

__device__ __forceinline__ double div_count(double a, int cnt)
{
	const double b = (double)max(cnt, 1);
	const double y = count_reciprocal(b);
	const double q = a * y;
	const double rem = __builtin_fma(-q, b, a);
	const double qc = __builtin_fma(rem, y, q);
	// class mask 0x63: signalling NaN, quiet NaN, -0, +0
	return __builtin_amdgcn_class(rem, 0x63) ? q : qc;
}

// One pass of the reference's do-while (sciddicaT.c:62-79) on one cell:
// avg = (m + the levels u[n] still in the set, in index order) / their count,
// then every level of the set with avg <= u[n] leaves it (bit n of elim, the
// reference's eliminated_cells); again = some level left.
//
// The sum starts at m = h - eps > 0, so it is never -0 and adding +-0.0 for
// an eliminated level leaves it unchanged (for a dry cell nothing is used).
// FIN (every level finite): each term is one FMA, u[n] * (0.0 or 1.0) + s,
// which rounds exactly as s + u[n] or returns s.  Otherwise an eliminated
// level enters as a selected +0.0.  Levels that were already eliminated may
// be marked again: only the new ones make `again`.
template <bool FIN>
__device__ __forceinline__ void avg_pass(const double u[5], double m, unsigned& elim, double& avg, bool& again)
{
	double s = m;
#pragma unroll
	for (int n = 0; n < 5; ++n) {
		if (FIN)
			s = __builtin_fma(u[n], (double)__builtin_amdgcn_ubfe(~elim, n, 1), s);
		else
			s += ((elim >> n) & 1u) ? 0.0 : u[n];
	}
	avg = div_count(s, 5 - __builtin_popcount(elim));
	unsigned out = elim;
#pragma unroll
	for (int n = 0; n < 5; ++n) out |= avg <= u[n] ? 1u << n : 0u;
	again = out != elim;
	elim = out;
}

__device__ __forceinline__ bool levels_finite(const double u[5])
{
	return __builtin_isfinite(u[0]) && __builtin_isfinite(u[1]) && __builtin_isfinite(u[2]) &&
	       __builtin_isfinite(u[3]) && __builtin_isfinite(u[4]);
}

// Runs body(std::true_type) when every lane's levels (of the cells that take
// part) are finite, body(std::false_type) otherwise; uniform per wave.
template <class Body>
__device__ __forceinline__ void with_finiteness(bool lane_finite, Body body)
{
	if (__all((int)lane_finite))
		body(std::true_type{});
	else
		body(std::false_type{});
}

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
	const double u[5] = {z0 + eps, u1, u2, u3, u4};
	unsigned elim = 0u;
	double avg = 0.0;
	with_finiteness(levels_finite(u), [&](auto fin) {
		bool again;
		do avg_pass<decltype(fin)::value>(u, m, elim, avg, again);
		while (again);
	});
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
	with_finiteness((!wa || levels_finite(u_a)) && (!wb || levels_finite(u_b)), [&](auto fin) {
		constexpr bool F = decltype(fin)::value;
		while (ga || gb) {
			bool na, nb;
			avg_pass<F>(u_a, ma, ea, avg_a, na);
			avg_pass<F>(u_b, mb, eb, avg_b, nb);
			ga = ga && na;
			gb = gb && nb;
		}
	});
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
// rows x cols toroidal grid and returns (wave-uniform) whether any cell whose
// outflows it computed (the strip and its halo) was wet.  EVEN: cols is even, so a lane's two cells are
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
__device__ __forceinline__ bool sweep_strip(const double* __restrict__ z, const double* __restrict__ h,
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
	bool seen = false;

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
			seen = seen || !(hc.x <= eps) || !(hc.y <= eps);
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
	return __any((int)seen) != 0;
}

// Both walking directions; chunks alternate.
template <int D, bool EVEN>
__device__ __forceinline__ bool sweep_chunk(const double* __restrict__ z, const double* __restrict__ h,
                                            double* __restrict__ hn, double* __restrict__ trash, int rows, int cols,
                                            int strip, int chunk, int r0, int r1, double r, double eps)
{
	if (chunk & 1) return sweep_strip<D, EVEN, true>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps);
	return sweep_strip<D, EVEN, false>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps);
}

// ---------------------------------------------------------------------------
// Aligned strips (even column counts).  A wave owns the output columns
// [128*strip, 128*strip + 128), two per lane, so every row access of the
// strip is one aligned 1 KiB run and every store covers whole cache lines (a
// strip that started mid-line would share partial lines with its neighbour
// and cost about 12% of the sweep's bandwidth).  The two columns on either
// side come with one more 16-byte load per row and array: lane 0 reads
// columns j0-2, j0-1 and lane 1 reads jR, jR+1, where jR is the column after
// the strip's last one (both wrap on the torus).  The outflows of the two
// halo cells j0-1 (lane 0) and jR (lane 1) are a third minimisation chain;
// the strip's first and last cells take their horizontal neighbour terms
// from those lanes instead of from a lane shuffle.  `pitch` is the row
// stride in elements (>= cols).

struct TriFlows {
	double a[4], b[4], c[4];
};

// pair_flows with a third cell c (computed only where c_on).
__device__ __forceinline__ void tri_flows(double ha, double za, double ua1, double ua2, double ua3, double ua4,
                                          double hb, double zb, double ub1, double ub2, double ub3, double ub4,
                                          bool c_on, double hc, double zc, double uc1, double uc2, double uc3,
                                          double uc4, double r, double eps, TriFlows& o, bool& seen)
{
#pragma unroll
	for (int q = 0; q < 4; ++q) o.a[q] = o.b[q] = o.c[q] = 0.0;
	const bool wa = !(ha <= eps), wb = !(hb <= eps), wc = c_on && !(hc <= eps);
	if (!(wa || wb || wc)) return;
	seen = true;
	const double ma = ha - eps, mb = hb - eps, mc = hc - eps;
	double u_a[5] = {za + eps, ua1, ua2, ua3, ua4};
	double u_b[5] = {zb + eps, ub1, ub2, ub3, ub4};
	double u_c[5] = {zc + eps, uc1, uc2, uc3, uc4};
	unsigned ea = 0u, eb = 0u, ec = 0u;
	double avg_a = 0.0, avg_b = 0.0, avg_c = 0.0;
	bool ga = wa, gb = wb, gc = wc;
	// as in pair_flows (chains a and b interleave); the halo chain c is wet
	// in at most two lanes and its pass is skipped while no lane needs it
	const bool fin = (!wa || levels_finite(u_a)) && (!wb || levels_finite(u_b)) && (!wc || levels_finite(u_c));
	with_finiteness(fin, [&](auto fin_t) {
		constexpr bool F = decltype(fin_t)::value;
		while (ga || gb || gc) {
			bool na, nb;
			avg_pass<F>(u_a, ma, ea, avg_a, na);
			avg_pass<F>(u_b, mb, eb, avg_b, nb);
			ga = ga && na;
			gb = gb && nb;
			if (__any((int)gc)) {
				bool nc;
				avg_pass<F>(u_c, mc, ec, avg_c, nc);
				gc = gc && nc;
			}
		}
	});
	if (wa) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.a[n - 1] = ((ea >> n) & 1u) ? 0.0 : (avg_a - u_a[n]) * r;
	}
	if (wb) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.b[n - 1] = ((eb >> n) & 1u) ? 0.0 : (avg_b - u_b[n]) * r;
	}
	if (wc) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.c[n - 1] = ((ec >> n) & 1u) ? 0.0 : (avg_c - u_c[n]) * r;
	}
}

// Value of v in lane l (l wave-uniform).
__device__ __forceinline__ double lane_value(double v, int l)
{
	const long long b = __double_as_longlong(v);
	const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
	const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int STRIP_AL = 128;

// Same stream of rows and the same arithmetic per cell as sweep_strip.
// Returns (wave-uniform) whether any computed cell was wet.
template <int D, bool UP>
__device__ __forceinline__ bool sweep_strip_al(const double* __restrict__ z, const double* __restrict__ h,
                                               double* __restrict__ hn, double* __restrict__ trash, int rows,
                                               int cols, int pitch, int strip, int r0, int r1, double r, double eps)
{
	constexpr int W = D + 4;
	const int lane = threadIdx.x & 63;
	const int n_it = r1 - r0 + 2;   // flows computed for rows r0-1 .. r1
	const int last_pos = n_it + 1;
	const int j0 = strip * STRIP_AL;
	const int ncol = min(STRIP_AL, cols - j0);   // even
	const int kR = ncol / 2 - 1;                 // lane holding the strip's last column
	const bool own = 2 * lane < ncol;
	const int j = own ? j0 + 2 * lane : j0;      // lanes past the grid's edge read a valid pair (unused)
	const int jR = j0 + ncol == cols ? 0 : j0 + ncol;
	const int jL = j0 == 0 ? cols - 2 : j0 - 2;
	const int jh = lane == 1 ? jR : jL;
	const bool lane0 = lane == 0;
	const bool c_on = lane < 2;

	auto row_of = [&](int pos) -> int { return UP ? r1 + 1 - pos : r0 - 2 + pos; };
	auto row_off = [&](int pos) -> size_t {
		return (size_t)dev::wrap1(row_of(min(pos, last_pos)), rows) * (size_t)pitch;
	};

	// ring of rows: the lane's pair (bz, bh) and its halo pair (cz, ch)
	double2 bz[W], bh[W], cz[W], ch[W];
#pragma unroll
	for (int q = 0; q < W - 1; ++q) {
		const size_t o = row_off(q);
		bz[q] = *reinterpret_cast<const double2*>(z + o + j);
		bh[q] = *reinterpret_cast<const double2*>(h + o + j);
		cz[q] = *reinterpret_cast<const double2*>(z + o + jh);
		ch[q] = *reinterpret_cast<const double2*>(h + o + jh);
		__builtin_amdgcn_sched_barrier(0);
	}
	bz[W - 1] = bh[W - 1] = cz[W - 1] = ch[W - 1] = make_double2(0.0, 0.0);

	double pa0 = 0.0, pa1 = 0.0, pa2 = 0.0, pa3 = 0.0, pb0 = 0.0, pb1 = 0.0, pb2 = 0.0, pb3 = 0.0;
	double pc1 = 0.0, pc2 = 0.0;   // W and E outflows of the halo cell of the row behind
	double ppa = 0.0, ppb = 0.0;
	bool seen = false;

	for (int base = 0; base < n_it; base += W) {
#pragma unroll
		for (int d = 0; d < W; ++d) {
			const int q = base + d;
			const int sb = d % W, sc = (d + 1) % W, sa = (d + 2) % W, sp = (d + W - 1) % W;
			{
				const size_t o = row_off(q + W - 1);
				bz[sp] = *reinterpret_cast<const double2*>(z + o + j);
				bh[sp] = *reinterpret_cast<const double2*>(h + o + j);
				cz[sp] = *reinterpret_cast<const double2*>(z + o + jh);
				ch[sp] = *reinterpret_cast<const double2*>(h + o + jh);
			}
			const double2 zb = bz[sb], hb = bh[sb], zc = bz[sc], hc = bh[sc], za = bz[sa], ha = bh[sa];
			const double2 zN = UP ? za : zb, hN = UP ? ha : hb, zS = UP ? zb : za, hS = UP ? hb : ha;
			const double2 czc = cz[sc], chc = ch[sc];
			const double2 czN = UP ? cz[sa] : cz[sb], chN = UP ? ch[sa] : ch[sb];
			const double2 czS = UP ? cz[sb] : cz[sa], chS = UP ? ch[sb] : ch[sa];
			const double ua = zc.x + hc.x, ub = zc.y + hc.y;
			const double uhx = czc.x + chc.x, uhy = czc.y + chc.y;
			const double ub_last = lane_value(ub, kR);   // level of the strip's last cell
			const double u_right = lane_value(uhx, 1);   // level at column jR
			double uw = dev::dpp_from_left(ub);
			if (lane0) uw = uhy;
			double ue = dev::dpp_from_right(ua);
			if (lane == kR) ue = u_right;
			// halo cell: column j0-1 in lane 0 (y of its pair), jR in lane 1 (x)
			const double hc_c = lane0 ? chc.y : chc.x, zc_c = lane0 ? czc.y : czc.x;
			const double uN_c = lane0 ? czN.y + chN.y : czN.x + chN.x;
			const double uS_c = lane0 ? czS.y + chS.y : czS.x + chS.x;
			const double uW_c = lane0 ? uhx : ub_last;
			const double uE_c = lane0 ? ua : uhy;
			TriFlows tf;
			tri_flows(hc.x, zc.x, zN.x + hN.x, uw, ub, zS.x + hS.x, hc.y, zc.y, zN.y + hN.y, ua, ue, zS.y + hS.y,
			          c_on, hc_c, zc_c, uN_c, uW_c, uE_c, uS_c, r, eps, tf, seen);
			const double nA = UP ? tf.a[3] : ppa, sA = UP ? ppa : tf.a[0];
			const double nB = UP ? tf.b[3] : ppb, sB = UP ? ppb : tf.b[0];
			double f2w = dev::dpp_from_left(pb2);   // E outflow of a's west neighbour
			if (lane0) f2w = pc2;
			const double pc1_right = lane_value(pc1, 1);
			double f1e = dev::dpp_from_right(pa1);  // W outflow of b's east neighbour
			if (lane == kR) f1e = pc1_right;
			double oa = hb.x, ob = hb.y;
			oa += nA - pa0;
			oa += f2w - pa1;
			oa += pb1 - pa2;
			oa += sA - pa3;
			ob += nB - pb0;
			ob += pa2 - pb1;
			ob += f1e - pb2;
			ob += sB - pb3;
			const bool row_ok = q >= 2 && q < n_it;
			double* dst = (row_ok && own) ? hn + (size_t)row_of(q) * (size_t)pitch + j : trash + 2 * lane;
			*reinterpret_cast<double2*>(dst) = make_double2(oa, ob);
			ppa = UP ? pa0 : pa3;
			ppb = UP ? pb0 : pb3;
			pa0 = tf.a[0];
			pa1 = tf.a[1];
			pa2 = tf.a[2];
			pa3 = tf.a[3];
			pb0 = tf.b[0];
			pb1 = tf.b[1];
			pb2 = tf.b[2];
			pb3 = tf.b[3];
			pc1 = tf.c[1];
			pc2 = tf.c[2];
			__builtin_amdgcn_sched_barrier(0);
		}
	}
	return __any((int)seen) != 0;
}

template <int D>
__device__ __forceinline__ bool sweep_chunk_al(const double* __restrict__ z, const double* __restrict__ h,
                                               double* __restrict__ hn, double* __restrict__ trash, int rows,
                                               int cols, int pitch, int strip, int chunk, int r0, int r1, double r,
                                               double eps)
{
	if (chunk & 1) return sweep_strip_al<D, true>(z, h, hn, trash, rows, cols, pitch, strip, r0, r1, r, eps);
	return sweep_strip_al<D, false>(z, h, hn, trash, rows, cols, pitch, strip, r0, r1, r, eps);
}

// Wet-first tile order.  A tile (FUSED_WAVES strips x one row chunk) that
// met a wet cell costs several times a dry one: the minimisation runs for the
// whole wave.  Dispatched in grid order, a wet tile that starts late ends the
// launch late, so each launch records the tiles that met wet cells and the
// next launch dispatches them first.  Per XCD group g (blockIdx & 7, which
// owns the band of tiles [g*per_xcd, (g+1)*per_xcd)), block l < n runs wet
// tile l of the group's list, and the following blocks walk the band in grid
// order skipping the tiles of the list.  Every tile runs exactly once per
// launch, and the order never changes a result.  The lists rotate over three
// sets: a launch reads set e%3, appends to set (e+1)%3 and clears the lengths
// of set (e+2)%3, which the next launch appends to.
struct WetSets {
	const unsigned* n_in;        // [8] list lengths (read)
	const int* l_in;             // [8][cap] tiles (read)
	unsigned* n_out;             // [8] lengths (appended)
	int* l_out;                  // [8][cap]
	unsigned* n_zero;            // [8] lengths cleared for the launch after
	const unsigned char* f_in;   // [n_logical] 1 = tile is in l_in
	unsigned char* f_out;        // [n_logical]
	int cap;                     // list capacity per group
};

// The tile of this block under the wet-first order (ws.f_out set) or the
// plain XCD-banded grid order; -1: the block has no tile.  Uniform per block.
__device__ __forceinline__ int wet_first_tile(const WetSets& ws, int n_tiles, int per_xcd)
{
	const int g = blockIdx.x & 7, l = blockIdx.x >> 3;
	if (!ws.f_out) {
		const int L = dev::xcd_logical_block(blockIdx.x, per_xcd);
		return L < n_tiles ? L : -1;
	}
	if (blockIdx.x == 0 && threadIdx.x < 8) ws.n_zero[threadIdx.x] = 0u;
	const int nw = min((int)ws.n_in[g], ws.cap);
	if (l < nw) return ws.l_in[g * ws.cap + l];
	const int k = l - nw;
	if (k >= per_xcd) return -1;
	const int L = g * per_xcd + k;
	if (L >= n_tiles || ws.f_in[L]) return -1;
	return L;
}

// Records whether tile L met a wet cell (any wave of the block); every
// thread of the block calls it.
__device__ __forceinline__ void wet_first_record(const WetSets& ws, int L, bool wave_wet)
{
	if (!ws.f_out) return;
	__shared__ int blk_wet;
	if (threadIdx.x == 0) blk_wet = 0;
	__syncthreads();
	if (wave_wet && (threadIdx.x & 63) == 0) atomicOr(&blk_wet, 1);
	__syncthreads();
	if (threadIdx.x == 0) {
		int w = blk_wet;
		if (w) {
			const int g = blockIdx.x & 7;
			const unsigned j = atomicAdd(ws.n_out + g, 1u);
			if ((int)j < ws.cap)
				ws.l_out[g * ws.cap + j] = L;
			else
				w = 0;   // list full: the tile stays in grid order
		}
		ws.f_out[L] = (unsigned char)w;
	}
}

// Dry regions.  When every cell within S rows and S columns of the output
// rows [r0, r1) of a strip has h <= eps, no cell that can reach an output
// cell in S steps has an outflow: the flows process returns before reading z
// for a dry cell, every flow a cell of the region exchanges is +0.0, and
// each of an output cell's width terms adds (0.0 - 0.0): after S steps
// h = h0 + 0.0, which turns -0.0 into +0.0 and leaves every other value as
// it is.  dry_copy streams h over the region with a halo of two rows and two
// columns (enough for S = 1 and S = 2; DRY_ROWS rows at a time, every load
// of a batch in flight together), writes h + 0.0 to the output rows as it
// goes and returns (wave-uniform) whether the region was dry; at the first
// batch that holds a wet cell it stops and returns false, and the caller
// runs the full sweep, which rewrites those rows.  z is not read.  Needs an
// even column count.
constexpr int DRY_ROWS = 16;

__device__ __forceinline__ bool dry_copy(const double* __restrict__ h, double* __restrict__ hn, int rows, int cols,
                                         int pitch, int strip, int r0, int r1, double eps)
{
	const int lane = threadIdx.x & 63;
	const int j0 = strip * STRIP_AL;
	const int ncol = min(STRIP_AL, cols - j0);   // even
	const bool own = 2 * lane < ncol;
	const int j = own ? j0 + 2 * lane : j0;
	const int jR = j0 + ncol == cols ? 0 : j0 + ncol;
	const int jL2 = j0 - 2 < 0 ? j0 - 2 + cols : j0 - 2;
	// lane 0 also checks the pair (j0-2, j0-1), lane 1 the pair (jR, jR+1)
	const int jh = lane == 0 ? jL2 : jR;
	const bool h_on = lane < 2;
	for (int base = r0 - 2; base < r1 + 2; base += DRY_ROWS) {
		double2 v[DRY_ROWS], x[DRY_ROWS];
#pragma unroll
		for (int k = 0; k < DRY_ROWS; ++k) {
			int rr = min(base + k, r1 + 1) % rows;   // wave-uniform; any row count
			rr = rr < 0 ? rr + rows : rr;
			const size_t o = (size_t)rr * (size_t)pitch;
			v[k] = *reinterpret_cast<const double2*>(h + o + j);
			x[k] = h_on ? *reinterpret_cast<const double2*>(h + o + jh) : make_double2(0.0, 0.0);
		}
		bool wet = false;
#pragma unroll
		for (int k = 0; k < DRY_ROWS; ++k) {
			wet = wet || !(v[k].x <= eps) || !(v[k].y <= eps) || (h_on && (!(x[k].x <= eps) || !(x[k].y <= eps)));
			const int rr = base + k;
			if (own && rr >= r0 && rr < r1)
				*reinterpret_cast<double2*>(hn + (size_t)rr * (size_t)pitch + j) =
				    make_double2(v[k].x + 0.0, v[k].y + 0.0);
		}
		if (__any((int)wet)) return false;
	}
	return true;
}

// ws.f_out == nullptr: plain grid order (gridDim.x = 8 * per_xcd); otherwise
// wet-first order (gridDim.x = 8 * (per_xcd + cap)).
template <int D>
__global__ __launch_bounds__(64 * FUSED_WAVES) __attribute__((amdgpu_waves_per_eu(D == 1 ? 3 : 2))) void step_fused_al(
    const double* __restrict__ z, const double* __restrict__ h, double* __restrict__ hn, double* __restrict__ trash,
    int rows, int cols, int n_strips, int n_logical, int per_xcd, int rpw, double r, double eps, WetSets ws)
{
	const int L = wet_first_tile(ws, n_logical, per_xcd);
	if (L < 0) return;
	// wave items in row-major (chunk, strip) order; a block is FUSED_WAVES
	// consecutive items (the item index is wave-uniform: a scalar register)
	const int item = __builtin_amdgcn_readfirstlane(L * FUSED_WAVES + (int)(threadIdx.x >> 6));
	const int strip = item % n_strips, chunk = item / n_strips;
	const int r0 = chunk * rpw;
	bool wet = false;
	if (r0 < rows)
		wet = dry_copy(h, hn, rows, cols, cols, strip, r0, min(r0 + rpw, rows), eps)
		          ? false
		          : sweep_chunk_al<D>(z, h, hn, trash, rows, cols, cols, strip, chunk, r0, min(r0 + rpw, rows), r,
		                              eps);
	wet_first_record(ws, L, wet);
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
// Shape of the sweep: row-chunk length and prefetch depth.
struct SweepShape {
	int rpw, depth;
};
// Device state of the wet-first order for one grid shape (owned by a model,
// an adjoint run, ...); allocated on first use, rebuilt when the shape changes.
struct WetOrder {
	void* buf = nullptr;
	int n_logical = 0, cap = 0;
	unsigned long long epoch = 0;
	~WetOrder();
	// the sets of the next launch (advances the epoch); -1 on allocation failure
	int next(hipStream_t st, int n_logical, WetSets& ws, int max_cap = 128);
};
// One fused step h -> hn on stream st (wo: wet-first order, or nullptr).
void launch_sweep(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                  double r, double eps, SweepShape shape, WetOrder* wo);
// The sweep shape for this grid: the environment's
// CALHIP_SCIDDICAT_ROWS_PER_WAVE / CALHIP_SCIDDICAT_PREFETCH, else the
// autotuned one (timed on first use by writing sweeps of h into `scratch`,
// then cached per grid shape).
SweepShape sweep_shape(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch,
                       double* trash, double r, double eps);
// The shape cached for this grid by an earlier sweep_shape call, else 16
// rows at prefetch depth 1; times nothing and allocates nothing.
SweepShape sweep_shape_cached(int rows, int cols);
// Two steps per sweep (sciddicat_2step.hpp, flows reset between the steps):
// the grids it supports (even cols >= 8, rows >= 8), one launch z, h -> hn,
// and its tile shape: CALHIP_SCIDDICAT_ROWS_PER_WAVE_2X (small-tile rows) /
// CALHIP_SCIDDICAT_2X_GROUP (small tiles per big tile), else the shape the
// tuner times fastest per grid shape, or {0, 0} when two launches of the
// one-step sweep beat every two-step shape (cached per grid shape; the
// timing sweeps write into `scratch`).
struct Sweep2Shape {
	int s_rows, bigk;   // small-tile rows, small tiles per big tile
};
bool sweep2_supported(int rows, int cols);
void launch_sweep2(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                   double r, double eps, Sweep2Shape shape, WetOrder* wo);
Sweep2Shape sweep2_shape(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch,
                         double* trash, double r, double eps);

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
	with_finiteness(levels_finite(s.u), [&](auto fin) {
		do {
			// the count and average of the last pass (which eliminated nothing)
			s.cnt = 5 - __builtin_popcount(s.elim);
			avg_pass<decltype(fin)::value>(s.u, m, s.elim, s.avg, again);
		} while (again);
	});
}

}  // namespace sct
}  // namespace calhip
