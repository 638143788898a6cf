// Two SciddicaT steps per sweep (temporal blocking) for even column counts.
//
// Between two steps the reference's steering resets the flows to zero
// (sciddicaT.c:133-146), so a step is a function of z and h alone and two
// steps can be chained inside one wave: h0 -> h1 -> h2, with h1 held in
// registers for the few rows the second step needs and never written to
// memory.  HBM traffic per cell and step pair: z and h0 read once, h2 written
// once (24 B), against 48 B for two single-step sweeps.
//
// A wave owns the 128 output columns of an aligned strip (two per lane) and a
// chunk of rows [r0, r1), as in sweep_strip_al (sciddicat_device.hpp).  The
// first step needs h1 two rows and two columns beyond the output, so it
// reads h0 and z four rows and four columns beyond it:
//   * rows: one stream of positions p (row r0-4+p walking down, r1+3-p
//     walking up); iteration q computes the step-1 outflows of position q+1
//     and h1 of position q (stage A, the single-step pipeline shifted by two
//     rows), then the step-2 outflows of position q-1 and h2 of position q-2
//     from the last three h1 rows (stage B, the single-step pipeline on h1);
//   * columns: lanes 0..3 also hold the halo pairs (j0-4, j0-3),
//     (j0-2, j0-1), (jR, jR+1), (jR+2, jR+3); stage A computes the outflows
//     of the six halo cells j0-3 .. j0-1 and jR .. jR+2 and h1 of the four
//     halo cells j0-2, j0-1, jR, jR+1, which stage B then reads (shifted to
//     lanes 0 and 1) exactly as the single-step sweep reads its halo pairs.
// Every cell's arithmetic is the reference's, in its order, so h2 equals two
// single steps bit for bit.  Rings: z and h0 six rows (positions q-2 ..
// q+3; the row loaded at iteration q is used at iteration q+1), h1 three rows.
#pragma once
#include "sciddicat_device.hpp"

namespace calhip {
namespace sct {

// pair_flows (cells a, b) plus the two halo cells x, y of lanes 0..3 (each
// computed only where its flag is set; their passes are skipped while no
// lane needs them).
struct QuadFlows {
	double a[4], b[4], x[4], y[4];
};

__device__ __forceinline__ void quad_flows(double ha, double za, double ua1, double ua2, double ua3, double ua4,
                                           double hb, double zb, double ub1, double ub2, double ub3, double ub4,
                                           bool x_on, double hx, double zx, double ux1, double ux2, double ux3,
                                           double ux4, bool y_on, double hy, double zy, double uy1, double uy2,
                                           double uy3, double uy4, double r, double eps, QuadFlows& o, bool& seen)
{
#pragma unroll
	for (int q = 0; q < 4; ++q) o.a[q] = o.b[q] = o.x[q] = o.y[q] = 0.0;
	const bool wa = !(ha <= eps), wb = !(hb <= eps), wx = x_on && !(hx <= eps), wy = y_on && !(hy <= eps);
	if (!(wa || wb || wx || wy)) return;
	seen = true;
	const double ma = ha - eps, mb = hb - eps, mx = hx - eps, my = hy - eps;
	double u_a[5] = {za + eps, ua1, ua2, ua3, ua4};
	double u_b[5] = {zb + eps, ub1, ub2, ub3, ub4};
	double u_x[5] = {zx + eps, ux1, ux2, ux3, ux4};
	double u_y[5] = {zy + eps, uy1, uy2, uy3, uy4};
	unsigned ea = 0u, eb = 0u, ex = 0u, ey = 0u;
	double avg_a = 0.0, avg_b = 0.0, avg_x = 0.0, avg_y = 0.0;
	bool ga = wa, gb = wb, gx = wx, gy = wy;
	const bool fin = (!wa || levels_finite(u_a)) && (!wb || levels_finite(u_b)) && (!wx || levels_finite(u_x)) &&
	                 (!wy || levels_finite(u_y));
	with_finiteness(fin, [&](auto fin_t) {
		constexpr bool F = decltype(fin_t)::value;
		while (ga || gb || gx || gy) {
			bool na, nb;
			avg_pass<F>(u_a, ma, ea, avg_a, na);
			avg_pass<F>(u_b, mb, eb, avg_b, nb);
			ga = ga && na;
			gb = gb && nb;
			if (__any((int)(gx || gy))) {
				bool nx, ny;
				avg_pass<F>(u_x, mx, ex, avg_x, nx);
				avg_pass<F>(u_y, my, ey, avg_y, ny);
				gx = gx && nx;
				gy = gy && ny;
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
	if (wx) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.x[n - 1] = ((ex >> n) & 1u) ? 0.0 : (avg_x - u_x[n]) * r;
	}
	if (wy) {
#pragma unroll
		for (int n = 1; n < 5; ++n) o.y[n - 1] = ((ey >> n) & 1u) ? 0.0 : (avg_y - u_y[n]) * r;
	}
}

// Rows [r0, r1) of strip `strip` advanced two steps: z, h -> hn (row stride
// `pitch` >= cols elements, a multiple of 2).  Needs an even column count,
// cols >= 8 and rows >= 8 (one wrap reaches every halo row and column).
// Returns (wave-uniform) whether any computed cell was wet.
template <bool UP>
__device__ __forceinline__ bool sweep2_strip_al(const double* __restrict__ z, const double* __restrict__ h,
                                                double* __restrict__ hn, double* __restrict__ trash, int rows,
                                                int cols, int pitch, int strip, int r0, int r1, double r,
                                                double eps)
{
	constexpr int U = 6;   // unroll = z/h0 ring length; the h1 ring (3) divides it
	const int lane = threadIdx.x & 63;
	const int na = r1 - r0 + 6;        // stage-A iterations: h1 for rows r0-2 .. r1+1
	const int nb = r1 - r0 + 2;        // stage-B positions: flows for rows r0-1 .. r1
	const int last_pos = na + 1;       // rows past it are clamped to it
	const int j0 = strip * STRIP_AL;
	const int ncol = min(STRIP_AL, cols - j0);   // even
	const int kR = ncol / 2 - 1;                 // lane holding the strip's last column
	const bool own = 2 * lane < ncol;
	const int j = own ? j0 + 2 * lane : j0;
	const int jR = j0 + ncol == cols ? 0 : j0 + ncol;
	const int jL4 = j0 - 4 < 0 ? j0 - 4 + cols : j0 - 4;
	const int jL2 = j0 - 2 < 0 ? j0 - 2 + cols : j0 - 2;
	const int jR2 = jR + 2 >= cols ? jR + 2 - cols : jR + 2;
	const int jh = lane == 0 ? jL4 : lane == 1 ? jL2 : lane == 2 ? jR : lane == 3 ? jR2 : jL2;
	const bool lane0 = lane == 0;
	const bool x_on = lane >= 1 && lane <= 3, y_on = lane <= 2;
	const bool c_on = lane < 2;   // stage-B halo cell: j0-1 in lane 0, jR in lane 1

	auto row_of = [&](int pos) -> int { return UP ? r1 + 3 - pos : r0 - 4 + pos; };
	auto row_off = [&](int pos) -> size_t {
		return (size_t)dev::wrap1(row_of(min(pos, last_pos)), rows) * (size_t)pitch;
	};

	// z / h0 rings: the lane's pair (bz, bh) and its halo pair (cz, ch)
	double2 bz[U], bh[U], cz[U], ch[U];
#pragma unroll
	for (int q = 0; q < U; ++q) bz[q] = bh[q] = cz[q] = ch[q] = make_double2(0.0, 0.0);
#pragma unroll
	for (int q = 0; q < 3; ++q) {
		const size_t o = row_off(q);
		bz[q] = *reinterpret_cast<const double2*>(z + o + j);
		bh[q] = *reinterpret_cast<const double2*>(h + o + j);
		cz[q] = *reinterpret_cast<const double2*>(z + o + jh);
		ch[q] = *reinterpret_cast<const double2*>(h + o + jh);
		__builtin_amdgcn_sched_barrier(0);
	}
	// h1 ring (slot = position % 3): main pair and halo pair
	double2 g1[3], g1h[3];
#pragma unroll
	for (int q = 0; q < 3; ++q) g1[q] = g1h[q] = make_double2(0.0, 0.0);

	// stage A carries: step-1 outflows of the row behind (a, b, x, y) and the
	// component of the row before it that the row behind needs
	double pa0 = 0.0, pa1 = 0.0, pa2 = 0.0, pa3 = 0.0, pb0 = 0.0, pb1 = 0.0, pb2 = 0.0, pb3 = 0.0;
	double px0 = 0.0, px1 = 0.0, px2 = 0.0, px3 = 0.0, py0 = 0.0, py1 = 0.0, py2 = 0.0, py3 = 0.0;
	double ppa = 0.0, ppb = 0.0, ppx = 0.0, ppy = 0.0;
	// stage B carries (as in sweep_strip_al)
	double sa0 = 0.0, sa1 = 0.0, sa2 = 0.0, sa3 = 0.0, sb0 = 0.0, sb1 = 0.0, sb2 = 0.0, sb3 = 0.0;
	double sc1 = 0.0, sc2 = 0.0, spa = 0.0, spb = 0.0;
	bool seen = false;
	// Dry rows.  A cell with h <= eps has no outflow (the process returns
	// before writing them), so a row whose cells in this wave are all dry has
	// zero flows, and a row whose own flows and those of the rows on either
	// side are zero keeps its level: h' = h + (0 - 0) + ... = h + 0.0, the
	// same bits the full update gives.  Bit k of hist_a / hist_b: the flows
	// of the stage's position k rows back may be non-zero (wave-uniform).
	unsigned hist_a = 0u, hist_b = 0u;

	for (int base = 0; base < na; base += U) {
#pragma unroll
		for (int d = 0; d < U; ++d) {
			const int q = base + d;
			{   // prefetch position q+3 into the slot of position q-3
				const int sp = (d + 3) % U;
				const size_t o = row_off(q + 3);
				bz[sp] = *reinterpret_cast<const double2*>(z + o + j);
				bh[sp] = *reinterpret_cast<const double2*>(h + o + j);
				cz[sp] = *reinterpret_cast<const double2*>(z + o + jh);
				ch[sp] = *reinterpret_cast<const double2*>(h + o + jh);
			}
			// ---------------- stage A: step 1 on positions q .. q+2
			{
				const int sb = d % U, sc = (d + 1) % U, sa = (d + 2) % U;
				const double2 zb = bz[sb], hb = bh[sb], zc = bz[sc], hc = bh[sc], za = bz[sa], ha = bh[sa];
				const double2 xzb = cz[sb], xhb = ch[sb], xzc = cz[sc], xhc = ch[sc], xza = cz[sa], xha = ch[sa];
				const bool wet_a = __any((int)(!(hc.x <= eps) || !(hc.y <= eps) || (x_on && !(xhc.x <= eps)) ||
				                               (y_on && !(xhc.y <= eps)))) != 0;
				hist_a = ((hist_a << 1) | (wet_a ? 1u : 0u)) & 7u;
				QuadFlows f;
				if (wet_a) {
					const double2 zN = UP ? za : zb, hN = UP ? ha : hb, zS = UP ? zb : za, hS = UP ? hb : ha;
					const double2 xzN = UP ? xza : xzb, xhN = UP ? xha : xhb, xzS = UP ? xzb : xza,
					              xhS = UP ? xhb : xha;
					const double ua = zc.x + hc.x, ub = zc.y + hc.y;
					const double ux = xzc.x + xhc.x, uy = xzc.y + xhc.y;
					// horizontal neighbour levels
					const double ua_first = lane_value(ua, 0);    // column j0
					const double ub_last = lane_value(ub, kR);    // the strip's last column
					const double uy_l1 = lane_value(uy, 1);       // column j0-1
					const double ux_l2 = lane_value(ux, 2);       // column jR
					double uw = dev::dpp_from_left(ub);
					if (lane0) uw = uy_l1;
					double ue = dev::dpp_from_right(ua);
					if (lane == kR) ue = ux_l2;
					double uxw = dev::dpp_from_left(uy);         // x's west: y of lane-1
					if (lane == 2) uxw = ub_last;
					double uye = dev::dpp_from_right(ux);        // y's east: x of lane+1
					if (lane == 1) uye = ua_first;
					quad_flows(hc.x, zc.x, zN.x + hN.x, uw, ub, zS.x + hS.x, hc.y, zc.y, zN.y + hN.y, ua, ue,
					           zS.y + hS.y, x_on, xhc.x, xzc.x, xzN.x + xhN.x, uxw, uy, xzS.x + xhS.x, y_on,
					           xhc.y, xzc.y, xzN.y + xhN.y, ux, uye, xzS.y + xhS.y, r, eps, f, seen);
				} else {
#pragma unroll
					for (int k = 0; k < 4; ++k) f.a[k] = f.b[k] = f.x[k] = f.y[k] = 0.0;
				}
				// h1 of position q (the row behind)
				if (hist_a == 0u) {
					g1[d % 3] = make_double2(hb.x + 0.0, hb.y + 0.0);
					g1h[d % 3] = make_double2(xhb.x + 0.0, xhb.y + 0.0);
				} else {
				const double nA = UP ? f.a[3] : ppa, sA = UP ? ppa : f.a[0];
				const double nB = UP ? f.b[3] : ppb, sB = UP ? ppb : f.b[0];
				const double nX = UP ? f.x[3] : ppx, sX = UP ? ppx : f.x[0];
				const double nY = UP ? f.y[3] : ppy, sY = UP ? ppy : f.y[0];
				const double py2_l1 = lane_value(py2, 1);    // E outflow of column j0-1
				const double px1_l2 = lane_value(px1, 2);    // W outflow of column jR
				const double pb2_last = lane_value(pb2, kR); // E outflow of the strip's last column
				const double pa1_first = lane_value(pa1, 0); // W outflow of column j0
				double f2w = dev::dpp_from_left(pb2);
				if (lane0) f2w = py2_l1;
				double f1e = dev::dpp_from_right(pa1);
				if (lane == kR) f1e = px1_l2;
				double f2wx = dev::dpp_from_left(py2);
				if (lane == 2) f2wx = pb2_last;
				double f1ey = dev::dpp_from_right(px1);
				if (lane == 1) f1ey = pa1_first;
				double oa = hb.x, ob = hb.y, ox = xhb.x, oy = xhb.y;
				oa += nA - pa0;
				oa += f2w - pa1;
				oa += pb1 - pa2;
				oa += sA - pa3;
				ob += nB - pb0;
				ob += pa2 - pb1;
				ob += f1e - pb2;
				ob += sB - pb3;
				ox += nX - px0;
				ox += f2wx - px1;
				ox += py1 - px2;
				ox += sX - px3;
				oy += nY - py0;
				oy += px2 - py1;
				oy += f1ey - py2;
				oy += sY - py3;
				g1[d % 3] = make_double2(oa, ob);   // slot of position q (U is a multiple of 3)
				g1h[d % 3] = make_double2(ox, oy);
				}
				ppa = UP ? pa0 : pa3;
				ppb = UP ? pb0 : pb3;
				ppx = UP ? px0 : px3;
				ppy = UP ? py0 : py3;
				pa0 = f.a[0];
				pa1 = f.a[1];
				pa2 = f.a[2];
				pa3 = f.a[3];
				pb0 = f.b[0];
				pb1 = f.b[1];
				pb2 = f.b[2];
				pb3 = f.b[3];
				px0 = f.x[0];
				px1 = f.x[1];
				px2 = f.x[2];
				px3 = f.x[3];
				py0 = f.y[0];
				py1 = f.y[1];
				py2 = f.y[2];
				py3 = f.y[3];
			}
			// ---------------- stage B: step 2 on h1 positions q-2 .. q
			const int t = q - 4;   // stage-B position of h1 position q-2
			if (t >= 0 && t < nb) {
				const int sb = (d + U - 2) % U, sc = (d + U - 1) % U, sa = d % U;   // z slots of q-2, q-1, q
				const int gb = (d + 1) % 3, gc = (d + 2) % 3, ga = d % 3;           // h1 slots of q-2, q-1, q
				const double2 zb = bz[sb], zc = bz[sc], za = bz[sa];
				const double2 hb = g1[gb], hc = g1[gc], ha = g1[ga];
				// the stage-B halo cells are j0-1 (y of lane 1's halo pair) and jR
				// (x of lane 2's), moved to lanes 0 and 1 below
				const bool wet_b = __any((int)(!(hc.x <= eps) || !(hc.y <= eps) ||
				                               (lane == 1 && !(g1h[gc].y <= eps)) ||
				                               (lane == 2 && !(g1h[gc].x <= eps)))) != 0;
				hist_b = ((hist_b << 1) | (wet_b ? 1u : 0u)) & 7u;
				TriFlows tf;
				if (wet_b) {
				const double2 zN = UP ? za : zb, hN = UP ? ha : hb, zS = UP ? zb : za, hS = UP ? hb : ha;
				// halo pairs (j0-2, j0-1) to lane 0 and (jR, jR+1) to lane 1
				const double2 czb = make_double2(dev::dpp_from_right(cz[sb].x), dev::dpp_from_right(cz[sb].y));
				const double2 czc = make_double2(dev::dpp_from_right(cz[sc].x), dev::dpp_from_right(cz[sc].y));
				const double2 cza = make_double2(dev::dpp_from_right(cz[sa].x), dev::dpp_from_right(cz[sa].y));
				const double2 chb = make_double2(dev::dpp_from_right(g1h[gb].x), dev::dpp_from_right(g1h[gb].y));
				const double2 chc = make_double2(dev::dpp_from_right(g1h[gc].x), dev::dpp_from_right(g1h[gc].y));
				const double2 cha = make_double2(dev::dpp_from_right(g1h[ga].x), dev::dpp_from_right(g1h[ga].y));
				const double2 czN = UP ? cza : czb, chN = UP ? cha : chb, czS = UP ? czb : cza, chS = UP ? chb : cha;
				const double ua = zc.x + hc.x, ub = zc.y + hc.y;
				const double uhx = czc.x + chc.x, uhy = czc.y + chc.y;
				const double ub_last = lane_value(ub, kR);
				const double u_right = lane_value(uhx, 1);
				double uw = dev::dpp_from_left(ub);
				if (lane0) uw = uhy;
				double ue = dev::dpp_from_right(ua);
				if (lane == kR) ue = u_right;
				const double hc_c = lane0 ? chc.y : chc.x, zc_c = lane0 ? czc.y : czc.x;
				const double uN_c = lane0 ? czN.y + chN.y : czN.x + chN.x;
				const double uS_c = lane0 ? czS.y + chS.y : czS.x + chS.x;
				const double uW_c = lane0 ? uhx : ub_last;
				const double uE_c = lane0 ? ua : uhy;
				tri_flows(hc.x, zc.x, zN.x + hN.x, uw, ub, zS.x + hS.x, hc.y, zc.y, zN.y + hN.y, ua, ue,
				          zS.y + hS.y, c_on, hc_c, zc_c, uN_c, uW_c, uE_c, uS_c, r, eps, tf, seen);
				} else {
#pragma unroll
					for (int k = 0; k < 4; ++k) tf.a[k] = tf.b[k] = tf.c[k] = 0.0;
				}
				double oa, ob;
				if (hist_b == 0u) {
					oa = hb.x + 0.0;
					ob = hb.y + 0.0;
				} else {
				const double nA = UP ? tf.a[3] : spa, sA = UP ? spa : tf.a[0];
				const double nB = UP ? tf.b[3] : spb, sB = UP ? spb : tf.b[0];
				double f2w = dev::dpp_from_left(sb2);
				if (lane0) f2w = sc2;
				const double sc1_right = lane_value(sc1, 1);
				double f1e = dev::dpp_from_right(sa1);
				if (lane == kR) f1e = sc1_right;
				oa = hb.x;
				ob = hb.y;
				oa += nA - sa0;
				oa += f2w - sa1;
				oa += sb1 - sa2;
				oa += sA - sa3;
				ob += nB - sb0;
				ob += sa2 - sb1;
				ob += f1e - sb2;
				ob += sB - sb3;
				}
				const bool row_ok = t >= 2;
				// output row: stage-B position t = stream position q-2
				double* dst = (row_ok && own) ? hn + (size_t)row_of(q - 2) * (size_t)pitch + j : trash + 2 * lane;
				*reinterpret_cast<double2*>(dst) = make_double2(oa, ob);
				spa = UP ? sa0 : sa3;
				spb = UP ? sb0 : sb3;
				sa0 = tf.a[0];
				sa1 = tf.a[1];
				sa2 = tf.a[2];
				sa3 = tf.a[3];
				sb0 = tf.b[0];
				sb1 = tf.b[1];
				sb2 = tf.b[2];
				sb3 = tf.b[3];
				sc1 = tf.c[1];
				sc2 = tf.c[2];
			}
			__builtin_amdgcn_sched_barrier(0);
		}
	}
	return __any((int)seen) != 0;
}

}  // namespace sct
}  // namespace calhip
