// SCIARA-fv2 lava-flow program: the processes of
// OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp (vents :3-16,
// outflowsMin/empiricalFlows :18-105, width_update :106-135,
// updateTemperature :137-169, removeActiveCells :171-175) and a device twin
// of its steering (:185-197), on a Moore, CAL_SPACE_FLAT model.
//
// One transition is four launches:
//   sc_vents    emission at the vent cells that are in the active set (host
//               computes each vent's thickness from the elapsed time exactly
//               as TVent::thickness does, include/vent.h:74-82)
//   sc_flows    outflows of every listed lava cell; a receiver whose flag
//               goes 0 -> 1 is appended to the list (the set after the
//               rebuild that follows the process in OpenCAL)
//   sc_width    mass and heat balance, radiative cooling, solidification and
//               removal, fused (all three read only the cell itself apart
//               from the neighbours' flows and temperature, which no launch
//               of this kernel writes); the new temperature goes to a
//               separate layer because neighbours still read the old one
//   sc_finish   copies the new temperature back, compacts the survivors into
//               the next list and, fused with the steering when the run loop
//               calls it right after the transition, resets the flows at the
//               survivors and clears lava and temperature on the border mask
// The active list is kept unordered on the device (every process is
// independent of the visiting order); the row-major order of OpenCAL's
// A.cells is restored whenever the host reads the set.  Lists and counters:
//   count[0] list length, count[1] cells appended by sc_flows,
//   count[2] survivors written by sc_finish, count[3] finished blocks (the
//   last block moves count[2] to count[0] and clears the others).
// Elementary functions are the correctly rounded ones of crmath.h (10^x,
// atan and cos through their first evaluation with Ziv's rounding test,
// the same bits as the full one); the oracle built with the same source
// reproduces every bit, and a glibc build differs only where glibc itself
// rounds the wrong way (about one call in 10^3 to 10^4, by one ulp).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "device_common.hpp"
#include "crmath.h"

namespace calhip {

namespace {

constexpr int NF = 8;
constexpr int SC_BLOCK = 256;
constexpr int VENTS_PER_LAUNCH = 16;

struct Flows8 {
	double* f[NF];
};

struct SciaraParams {
	double Pc, Pac, a, b, c, d, rad2, PTsol, PTvent, Pclock, Pcool, Prho, Peps, Psigma, Pcv;
};

struct VentArgs {
	int n;
	int cell[VENTS_PER_LAUNCH];
	double e[VENTS_PER_LAUNCH];
};

// Moore neighbourhood in OpenCAL's order: centre, N, W, E, S, NW, SW, SE, NE
// (calDefineMooreNeighborhood2D, cal2D.c).
__device__ __forceinline__ int moore_di(int k)
{
	return k == 1 || k == 5 || k == 8 ? -1 : (k == 4 || k == 6 || k == 7 ? 1 : 0);
}
__device__ __forceinline__ int moore_dj(int k)
{
	return k == 2 || k == 5 || k == 6 ? -1 : (k == 3 || k == 7 || k == 8 ? 1 : 0);
}

// linear index of neighbour k of (i, j), -1 outside the grid
__device__ __forceinline__ int nb_flat(int i, int j, int k, int rows, int cols)
{
	const int ii = i + moore_di(k), jj = j + moore_dj(k);
	return (ii < 0 || ii >= rows || jj < 0 || jj >= cols) ? -1 : ii * cols + jj;
}

// grid-stride total: the listed cells, or the whole grid when the list is
// empty (OpenCAL's A.cells == NULL) or the model has no active cells
__device__ __forceinline__ size_t sweep_total(int cnt, size_t n) { return cnt > 0 ? (size_t)cnt : n; }

// emission at vent cells; with `flags`, only at cells of the active set
// (flags equal the list at the start of a transition; an empty list means
// the whole grid)
__global__ void sc_vents(double* __restrict__ h, double* __restrict__ T, const uint8_t* __restrict__ flags,
                         const int* __restrict__ count, VentArgs v, double PTvent)
{
	const int k = threadIdx.x;
	if (k >= v.n) return;
	const int c = v.cell[k];
	if (flags && count[0] > 0 && !flags[c]) return;
	h[c] = h[c] + v.e[k];
	T[c] = PTvent;
}

// the same, membership from the list itself (used when the flags were changed
// by the host after the last rebuild)
__global__ void sc_vents_list(double* __restrict__ h, double* __restrict__ T, const int* __restrict__ cells,
                              const int* __restrict__ count, VentArgs v, double PTvent, size_t n)
{
	const int cnt = count[0];
	const size_t total = sweep_total(cnt, n);
	for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
	     idx += (size_t)gridDim.x * blockDim.x) {
		const int c = cnt > 0 ? cells[idx] : (int)idx;
		for (int k = 0; k < v.n; ++k)
			if (c == v.cell[k]) {
				h[c] = h[c] + v.e[k];
				T[c] = PTvent;
			}
	}
}

// set the flag of cell c; true when it was clear
__device__ __forceinline__ bool set_flag(uint8_t* flags, int c)
{
	unsigned* w = reinterpret_cast<unsigned*>(flags + (c & ~3));
	const unsigned bit = 1u << ((c & 3) * 8);
	return !(atomicOr(w, bit) & bit);
}

// outflowsMin + empiricalFlows (Sciara.cpp:23-105), eight lanes per cell:
// lane k-1 of an aligned group of eight handles neighbour k (its level, slope
// angle, critical thickness and outflow), so a cell's eight atan / cos
// evaluations run side by side instead of one after another.  The group
// shares the eight levels through lane shuffles and every lane runs the same
// minimisation on them (same operations, same order: the same average).
// The grid-stride loop walks cells by group; a group's lanes always agree on
// its cell, so a shuffle only reads lanes that take part.
template <bool ACTIVE>
__global__ __launch_bounds__(SC_BLOCK) void sc_flows(const double* __restrict__ z, const double* __restrict__ h,
                                                     const double* __restrict__ T, Flows8 f, uint8_t* flags,
                                                     int* cells, int* count, int rows, int cols, SciaraParams P,
                                                     int append)
{
	const size_t n = (size_t)rows * cols;
	const int cnt0 = ACTIVE ? count[0] : 0;
	const size_t total = sweep_total(cnt0, n);
	const int lane = (int)(threadIdx.x & 63);
	const int g0 = lane & ~7;          // first lane of this lane's group
	const int k = (lane & 7) + 1;      // the neighbour this lane handles
	const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 3;
	const size_t idx0 = (((size_t)blockIdx.x * blockDim.x) >> 3) + (threadIdx.x >> 3);
	// the first list entry, loaded beside the count (clamped into the list's
	// n entries; used only when the count covers it)
	const int c_first = ACTIVE ? cells[min(idx0, n - 1)] : 0;
	for (size_t idx = idx0; idx < total; idx += stride) {
		const int c = cnt0 > 0 ? (idx == idx0 ? c_first : cells[idx]) : (int)idx;
		const double h0 = h[c];
		if (!(h0 > 0)) continue;
		const int i = c / cols, j = c - (c / cols) * cols;
		const double t = T[c];
		const double Pr = crz_pow10(P.a + P.b * t);
		const double hc = crz_pow10(P.c + P.d * t);
		const double sz0 = z[c];
		const int nbk = nb_flat(i, j, k, rows, cols);
		const double hk = nbk >= 0 ? h[nbk] : 0.0;
		const double zr = nbk >= 0 ? z[nbk] : 0.0;
		const double zk = k < 5 ? zr : sz0 - (sz0 - zr) / P.rad2;
		const double lvl0 = sz0 + h0;
		const double lk = zk + hk;
		const bool down = lvl0 > lk;
		const double Hk = down ? lk : 0.0;
		const double hcck = hc * crz_cos(crz_atan((lvl0 - lk) / P.Pc));
		double H[9];
		H[0] = sz0;
#pragma unroll
		for (int q = 1; q < 9; ++q) H[q] = __shfl(Hk, g0 + q - 1, 64);
		unsigned keep = 1u | ((unsigned)((__ballot((int)down) >> g0) & 0xffull) << 1);
		double avg;
		bool loop;
		do {
			loop = false;
			avg = h0;
			int counter = 0;
#pragma unroll
			for (int q = 0; q < 9; ++q)
				if ((keep >> q) & 1u) {
					avg += H[q];
					counter++;
				}
			if (counter != 0) avg = avg / (double)counter;
#pragma unroll
			for (int q = 0; q < 9; ++q)
				if (((keep >> q) & 1u) && avg <= H[q]) {
					keep &= ~(1u << q);
					loop = true;
				}
		} while (loop);
		if (((keep >> k) & 1u) && h0 > hcck) {
			const double fk = Pr * (avg - Hk);
			if (fk > 0) {
				f.f[k - 1][c] = fk;
				if (ACTIVE && nbk >= 0 && set_flag(flags, nbk) && append)
					cells[cnt0 + atomicAdd(&count[1], 1)] = nbk;
			}
		}
	}
}

// width_update + updateTemperature + removeActiveCells (Sciara.cpp:106-175)
template <bool ACTIVE>
__global__ __launch_bounds__(SC_BLOCK) void sc_width(double* __restrict__ z, double* __restrict__ h,
                                                     const double* __restrict__ T, double* __restrict__ Tn,
                                                     Flows8 f, const uint8_t* __restrict__ Mb,
                                                     double* __restrict__ Msl, uint8_t* __restrict__ flags,
                                                     const int* __restrict__ cells, const int* __restrict__ count,
                                                     int rows, int cols, SciaraParams P)
{
	// the flow that enters from neighbour n is that neighbour's outflow
	// towards this cell (Sciara.cpp:107)
	constexpr int opposite[NF] = {3, 2, 1, 0, 6, 7, 4, 5};
	const size_t n = (size_t)rows * cols;
	const int cnt = ACTIVE ? count[0] + count[1] : 0;
	const size_t total = sweep_total(cnt, n);
	// eight lanes per cell, as in sc_flows: lane k-1 of a group loads
	// neighbour k's inflow, temperature and the cell's outflow towards k;
	// every lane then sums the group's eight terms in k order
	const int lane = (int)(threadIdx.x & 63);
	const int g0 = lane & ~7;
	const int k = (lane & 7) + 1;
	const bool writer = (lane & 7) == 0;
	const double* fin = f.f[opposite[0]];
	const double* fout = f.f[0];
#pragma unroll
	for (int q = 1; q < NF; ++q)
		if (k - 1 == q) {
			fin = f.f[opposite[q]];
			fout = f.f[q];
		}
	const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 3;
	const size_t idx0 = (((size_t)blockIdx.x * blockDim.x) >> 3) + (threadIdx.x >> 3);
	const int c_first = ACTIVE ? cells[min(idx0, n - 1)] : 0;   // as in sc_flows
	for (size_t idx = idx0; idx < total; idx += stride) {
		const int c = cnt > 0 ? (idx == idx0 ? c_first : cells[idx]) : (int)idx;
		const int i = c / cols, j = c - (c / cols) * cols;
		const double ih = h[c], it = T[c];
		const int qk = nb_flat(i, j, k, rows, cols);
		const double in_k = qk >= 0 ? fin[qk] : 0.0;
		const double out_k = fout[c];
		const double nt_k = qk >= 0 ? T[qk] : 0.0;
		const double ht_k = in_k * nt_k;
		double ht = 0.0, in_sum = 0.0, out_sum = 0.0;
#pragma unroll
		for (int m = 0; m < 8; ++m) {
			ht += __shfl(ht_k, g0 + m, 64);
			in_sum += __shfl(in_k, g0 + m, 64);
			out_sum += __shfl(out_k, g0 + m, 64);
		}
		double h1 = ih;
		h1 += in_sum - out_sum;
		double t1 = it;
		if (in_sum > 0 || out_sum > 0) {
			double residual = ih;
			residual -= out_sum;
			t1 = (residual * it + ht) / (residual + in_sum);
		}
		double h2 = h1, t2 = t1;
		const bool mb = Mb[c] != 0;
		if (h1 > 0 && !mb) {
			const double nT = t1;
			const double num = (((3.0 * cr_cube(nT)) * P.Peps) * P.Psigma * P.Pclock) * P.Pcool;
			const double den = ((P.Prho * P.Pcv) * h1) * P.Pac;
			const double aus = 1.0 + num / den;
			const double st = nT / crz_pow_third(aus);
			t2 = st;
			if (st <= P.PTsol) {
				if (writer) {
					z[c] = z[c] + h1;
					Msl[c] = Msl[c] + h1;
				}
				h2 = 0.0;
				t2 = P.PTsol;
			}
		}
		if (writer) {
			h[c] = h2;
			Tn[c] = t2;
			if (ACTIVE && t2 <= P.PTsol && !mb) flags[c] = 0;
		}
	}
}

// End of a transition: T := Tn over the processed cells, survivors appended
// to `alt`; with STEER also the steering (flows := 0 at the active cells,
// h = T = 0 on the border mask).
template <bool ACTIVE, bool STEER>
__global__ __launch_bounds__(SC_BLOCK) void sc_finish(double* __restrict__ h, double* __restrict__ T,
                                                      const double* __restrict__ Tn, Flows8 f,
                                                      const uint8_t* __restrict__ Mb,
                                                      const uint8_t* __restrict__ flags,
                                                      const int* __restrict__ cells, int* __restrict__ alt,
                                                      int* count, const int* __restrict__ mb_cells, int n_mb,
                                                      int rows, int cols)
{
	const size_t n = (size_t)rows * cols;
	const int cnt = ACTIVE ? count[0] + count[1] : 0;
	const size_t total = sweep_total(cnt, n);
	const size_t stride = (size_t)gridDim.x * blockDim.x;
	const int lane = threadIdx.x & (dev::WAVE - 1);
	__shared__ int wave_off[SC_BLOCK / dev::WAVE];
	__shared__ int block_pos;
	// the steering's border reset first: its cells are independent of the
	// list, and where a border cell is also listed both write h = T = 0
	if (ACTIVE && STEER)
		for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n_mb; k += stride) {
			const int c = mb_cells[k];
			h[c] = 0.0;
			T[c] = 0.0;
		}
	const size_t idx0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int c_first = ACTIVE ? cells[min(idx0, n - 1)] : 0;   // as in sc_flows
	for (size_t base = (size_t)blockIdx.x * blockDim.x; base < total; base += stride) {
		const size_t idx = base + threadIdx.x;
		const bool valid = idx < total;
		const int c = valid ? (cnt > 0 ? (idx == idx0 ? c_first : cells[idx]) : (int)idx) : 0;
		bool surv = false;
		if (valid) {
			const bool mb = Mb[c] != 0;
			T[c] = (STEER && mb) ? 0.0 : Tn[c];
			if (ACTIVE) surv = flags[c] != 0;
			if (STEER && (!ACTIVE || surv))
#pragma unroll
				for (int q = 0; q < NF; ++q) f.f[q][c] = 0.0;
			if (STEER && mb) h[c] = 0.0;
		}
		if (ACTIVE) {
			// one atomic per block: waves' survivor counts scanned in LDS
			const unsigned long long m = __ballot(surv);
			const int wid = threadIdx.x >> 6;
			if (lane == 0) wave_off[wid] = __popcll(m);
			__syncthreads();
			if (threadIdx.x == 0) {
				int tot = 0;
				for (int w = 0; w < SC_BLOCK / dev::WAVE; ++w) {
					const int k = wave_off[w];
					wave_off[w] = tot;
					tot += k;
				}
				block_pos = tot ? atomicAdd(&count[2], tot) : 0;
			}
			__syncthreads();
			if (surv) alt[block_pos + wave_off[wid] + __popcll(m & ((1ull << lane) - 1ull))] = c;
			__syncthreads();
		}
	}
	if (!ACTIVE) return;
	// The last block to finish moves the survivor count to count[0].  Only
	// the counters pass between blocks here, all through device-scope atomics
	// (the list and the cells are read by later launches), so no fence is
	// needed before the block's arrival is counted.
	__shared__ int last_block;
	__shared__ int survivors;
	__syncthreads();
	if (threadIdx.x == 0) {
		const int done = atomicAdd(&count[3], 1);
		last_block = done == (int)gridDim.x - 1;
		if (last_block) {
			survivors = atomicAdd(&count[2], 0);
			atomicExch(&count[0], survivors);
			atomicExch(&count[1], 0);
			atomicExch(&count[2], 0);
			atomicExch(&count[3], 0);
		}
	}
	__syncthreads();
	// an empty set: the steering's reset covers the whole grid
	if (STEER && last_block && survivors == 0)
		for (size_t c = threadIdx.x; c < n; c += blockDim.x)
#pragma unroll
			for (int q = 0; q < NF; ++q) f.f[q][c] = 0.0;
}

// ---- two launches per transition (flows, width) ----
//
// A transition followed by the steering twin, in active mode, as two
// launches: sc_flows_f and sc_width_f.  Between two such transitions the
// engine keeps the state in a form the next sc_flows_f completes:
//   * the list P of the cells the last width process visited (the set
//     before it plus its receivers, less the set-aside border cells below),
//     count cnt[t % 3];
//   * flags: 1 for the cells of the set after the removal, 0 elsewhere;
//   * Tn: the new temperature of P's cells (T not yet copied back);
//   * the flows of the set's cells not yet reset by the steering.
// sc_flows_f copies Tn back into T over P, lists the set's cells (the
// members of P) and its receivers in the next list, applies the vents'
// emission where it reads h and T (a vent's cell, when it is in the set,
// reads as h + e and PTvent, the values sc_vents would have stored), writes
// all eight flows of every cell of the set (the steering's reset folded
// in) and marks its receivers outside the set with flag bit 1.  sc_width_f
// then runs the width, temperature and removal processes over the new
// list, applies the steering's border reset to the cells it visits (the
// others hold h = T = 0 already), writes flags 0/1, and marks the vents
// whose cell stayed in the set for the next transition's emission.
//
// Border cells of the set.  A cell of the border mask is never removed, and
// after every steering it holds h = T = 0 and zero flows; unless a vent
// emits into it, every later transition leaves it so: it sends nothing
// (h = 0), what it receives the steering removes, and the steering's reset
// rewrites the zeros it holds.  So sc_flows_f writes its zero flows once
// and then leaves it out of the next list (the reference's example keeps
// the whole border in its set, which would otherwise be revisited at every
// step); *skip counts those cells, which stay in the set (flag 1).  The
// engine turns this off when a vent lies on the border.
//
// When the set is empty OpenCAL's processes cover the whole grid.  The set
// at the start of transition t is empty exactly when no cell of P is a
// member and no border cell was set aside: then the new list is empty and
// *skip is 0, and sc_width_f records t in *stall.  sc_flows_f of t has then
// only finished transition t - 1 (T copied back), and every later fused
// launch does nothing: the host finishes the last completed transition and
// runs the stalled ones again through the four launches.
struct FusedCtl {
	const int* P;          // the list of the previous transition's width
	int* Lnew;             // the list of this transition
	int* cnt;              // [3] list lengths by transition mod 3
	int* skip;             // border cells of the set left out of the lists
	int* stall;            // transition at which the set was found empty (0: none)
	int* vent_act;         // [VENTS_PER_LAUNCH] transition of each vent's last emission
	const int* count0;     // first fused transition: P is the set list, count0[0] entries
	int t;                 // transition number (>= 1)
	int entry;             // first fused transition: T is current, P = the set
	int skip_border;       // leave idle border cells of the set out of the lists
};

__device__ __forceinline__ int vent_index(const VentArgs& v, int c)
{
	int r = -1;
#pragma unroll
	for (int k = 0; k < VENTS_PER_LAUNCH; ++k)
		if (k < v.n && v.cell[k] == c) r = k;
	return r;
}

// What a lane of sc_flows_f reads of its cell c and of c's neighbour k (h,
// z and flags are not written by that launch except flag bit 1, which is
// not read there).
struct FlowsIn {
	int c, nbk;
	unsigned fl, fln;
	double tcur, hcur, sz0, hk, zr;
	bool mbc;
};

__device__ __forceinline__ FlowsIn flows_load(int c, int k, const double* __restrict__ z,
                                              const double* __restrict__ h, const double* T,
                                              const double* __restrict__ Tn, const uint8_t* flags,
                                              const uint8_t* __restrict__ Mb, int entry, int rows, int cols)
{
	FlowsIn r;
	r.c = c;
	const int i = c / cols, j = c - (c / cols) * cols;
	r.nbk = nb_flat(i, j, k, rows, cols);
	r.fl = flags[c];
	r.tcur = entry ? T[c] : Tn[c];
	r.hcur = h[c];
	r.sz0 = z[c];
	r.mbc = Mb[c] != 0;
	r.hk = 0.0;
	r.zr = 0.0;
	r.fln = 0u;
	if (r.nbk >= 0) {
		r.hk = h[r.nbk];
		r.zr = z[r.nbk];
		r.fln = flags[r.nbk];
	}
	return r;
}

// Eight lanes per cell (lane k - 1 of a group owns direction k); a block
// takes SC_BLOCK / 8 consecutive entries of P per pass, so that the
// positions of its cells in the next list come from one atomic per pass.
// A pass is latency-bound (a few cells per wave): the first pass's list
// entry and the cell's loads go out before the list length and the stall
// word are known (an entry beyond the list, stale or never written, is
// clamped into the grid and its values are not used), a pass's loads are
// issued together, the list atomic's return is awaited only after the
// flows, and the two power laws of a cell are split between the group's
// even and odd lanes.
__global__ __launch_bounds__(SC_BLOCK) void sc_flows_f(const double* __restrict__ z, const double* __restrict__ h,
                                                       double* __restrict__ T, const double* __restrict__ Tn, Flows8 f,
                                                       uint8_t* flags, const uint8_t* __restrict__ Mb, FusedCtl F,
                                                       VentArgs v, int rows, int cols, SciaraParams P)
{
	constexpr int CPB = SC_BLOCK / 8;            // cells per block and pass
	constexpr int NW = SC_BLOCK / dev::WAVE;     // waves per block
	__shared__ int s_off[2][NW];
	__shared__ int s_base[2];
	const int tc = F.t % 3, tn = (F.t + 1) % 3, tp = (F.t + 2) % 3;
	const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
	const int g0 = lane & ~7;
	const int k = (lane & 7) + 1;
	const bool writer = (lane & 7) == 0;
	const unsigned long long below = (1ull << lane) - 1ull;
	const int ncell = rows * cols;
	const size_t first = (size_t)blockIdx.x * CPB;
	int c_pre = F.P[min(first + (threadIdx.x >> 3), (size_t)ncell - 1)];
	c_pre = min(max(c_pre, 0), ncell - 1);
	const FlowsIn pre = flows_load(c_pre, k, z, h, T, Tn, flags, Mb, F.entry, rows, cols);
	if (*F.stall) return;
	const int np = F.entry ? F.count0[0] : F.cnt[tp];
	if (blockIdx.x == 0 && threadIdx.x == 0) F.cnt[tn] = 0;   // the next transition's list length
	int pass = 0;
	for (size_t bb = first; bb < (size_t)np; bb += (size_t)gridDim.x * CPB, pass ^= 1) {
		const size_t idx = bb + (threadIdx.x >> 3);
		const bool valid = idx < (size_t)np;
		const FlowsIn in = bb == first ? pre
		                               : flows_load(valid ? F.P[idx] : 0, k, z, h, T, Tn, flags, Mb, F.entry, rows, cols);
		const int c = in.c, nbk = valid ? in.nbk : -1;
		const unsigned fl = valid ? in.fl : 0u, fln = in.fln;
		const double tcur = in.tcur, hcur = in.hcur, sz0 = in.sz0, zr = in.zr;
		double hk = in.hk;
		const bool mbc = in.mbc;
		if (valid && !F.entry && writer) T[c] = tcur;
		const bool member = valid && (fl & 1u);
		const int vk = member ? vent_index(v, c) : -1;
		double h0 = member ? hcur : 0.0, t = tcur;
		if (vk >= 0) {
			h0 = h0 + v.e[vk];
			t = P.PTvent;
			if (writer) F.vent_act[vk] = F.t;
		}
		const bool wet = member && h0 > 0;
		const bool set_aside = member && F.skip_border && vk < 0 && !wet && mbc;
		const bool keep = member && !set_aside;
		// positions in the next list: one atomic per block and pass
		const unsigned long long km = __ballot(keep && writer);
		const unsigned long long sm = __ballot(set_aside && writer);
		if (lane == 0) s_off[pass][wid] = __popcll(km) | (__popcll(sm) << 16);
		__syncthreads();
		int base = 0;
		if (threadIdx.x == 0) {
			int tot = 0, sk = 0;
			for (int w = 0; w < NW; ++w) {
				const int x = s_off[pass][w];
				s_off[pass][w] = tot;
				tot += x & 0xffff;
				sk += x >> 16;
			}
			if (tot) base = atomicAdd(F.cnt + tc, tot);
			if (sk) atomicAdd(F.skip, sk);
		}
		if (member && !wet) f.f[k - 1][c] = 0.0;
		if (wet) {
			const double pw = crz_pow10((lane & 1) ? P.c + P.d * t : P.a + P.b * t);
			const double Pr = __shfl(pw, g0, 64);
			const double hc = __shfl(pw, g0 + 1, 64);
			if (nbk >= 0) {
				const int vn = vent_index(v, nbk);
				if (vn >= 0 && (fln & 1u)) hk = hk + v.e[vn];
			}
			const double zk = k < 5 ? zr : sz0 - (sz0 - zr) / P.rad2;
			const double lvl0 = sz0 + h0;
			const double lk = zk + hk;
			const bool down = lvl0 > lk;
			const double Hk = down ? lk : 0.0;
			const double hcck = hc * crz_cos(crz_atan((lvl0 - lk) / P.Pc));
			double H[9];
			H[0] = sz0;
#pragma unroll
			for (int q = 1; q < 9; ++q) H[q] = __shfl(Hk, g0 + q - 1, 64);
			unsigned keepm = 1u | ((unsigned)((__ballot((int)down) >> g0) & 0xffull) << 1);
			double avg;
			bool loop;
			do {
				loop = false;
				avg = h0;
				int counter = 0;
#pragma unroll
				for (int q = 0; q < 9; ++q)
					if ((keepm >> q) & 1u) {
						avg += H[q];
						counter++;
					}
				if (counter != 0) avg = avg / (double)counter;
#pragma unroll
				for (int q = 0; q < 9; ++q)
					if (((keepm >> q) & 1u) && avg <= H[q]) {
						keepm &= ~(1u << q);
						loop = true;
					}
			} while (loop);
			double fk = 0.0;
			if (((keepm >> k) & 1u) && h0 > hcck) fk = Pr * (avg - Hk);
			// the steering reset every flow of the set's cells; the process
			// writes the positive ones (Sciara.cpp:95-101)
			f.f[k - 1][c] = fk > 0 ? fk : 0.0;
			// a receiver outside the set joins the next list once (flag bit
			// 1), its slot from one atomic per wave; the set's cells are
			// listed below
			bool fresh = false;
			if (fk > 0 && nbk >= 0 && !(fln & 1u)) {
				unsigned* w = reinterpret_cast<unsigned*>(flags + (nbk & ~3));
				const int sh = (nbk & 3) * 8;
				const unsigned old = atomicOr(w, 2u << sh);
				fresh = ((old >> sh) & 3u) == 0u;
			}
			const unsigned long long rm = __ballot(fresh);
			if (rm) {
				const int leader = __ffsll((long long)rm) - 1;
				int rb = 0;
				if (lane == leader) rb = atomicAdd(F.cnt + tc, __popcll(rm));
				rb = __shfl(rb, leader, 64);
				if (fresh) F.Lnew[rb + __popcll(rm & below)] = nbk;
			}
		}
		if (threadIdx.x == 0) s_base[pass] = base;
		__syncthreads();
		if (keep && writer) F.Lnew[s_base[pass] + s_off[pass][wid] + __popcll(km & below)] = c;
	}
}

// What a lane of sc_width_f reads of its cell c and of c's neighbour k
// (the flows, T and the neighbours' h are not written by that launch).
struct WidthIn {
	int c, qk, vk, vn, vact, vnact;
	double ih, it, in_k, out_k, nt_k;
	bool mb;
};

__device__ __forceinline__ WidthIn width_load(int c, int k, const double* h, const double* __restrict__ T,
                                              const double* fin, const double* fout,
                                              const uint8_t* __restrict__ Mb, const int* vent_act,
                                              const VentArgs& v, int rows, int cols)
{
	WidthIn r;
	r.c = c;
	const int i = c / cols, j = c - (c / cols) * cols;
	r.ih = h[c];
	r.it = T[c];
	r.mb = Mb[c] != 0;
	r.vk = vent_index(v, c);
	r.vact = r.vk >= 0 ? vent_act[r.vk] : 0;
	r.qk = nb_flat(i, j, k, rows, cols);
	r.out_k = fout[c];
	r.in_k = 0.0;
	r.nt_k = 0.0;
	r.vn = -1;
	r.vnact = 0;
	if (r.qk >= 0) {
		r.in_k = fin[r.qk];
		r.nt_k = T[r.qk];
		r.vn = vent_index(v, r.qk);
		r.vnact = r.vn >= 0 ? vent_act[r.vn] : 0;
	}
	return r;
}

__global__ __launch_bounds__(SC_BLOCK) void sc_width_f(double* __restrict__ z, double* __restrict__ h,
                                                       const double* __restrict__ T, double* __restrict__ Tn,
                                                       Flows8 f, const uint8_t* __restrict__ Mb,
                                                       double* __restrict__ Msl, uint8_t* __restrict__ flags,
                                                       FusedCtl F, VentArgs v, int rows, int cols, SciaraParams P)
{
	constexpr int opposite[NF] = {3, 2, 1, 0, 6, 7, 4, 5};
	const int tc = F.t % 3;
	const int lane = (int)(threadIdx.x & 63);
	const int g0 = lane & ~7;
	const int k = (lane & 7) + 1;
	const bool writer = (lane & 7) == 0;
	const double* fin = f.f[opposite[0]];
	const double* fout = f.f[0];
#pragma unroll
	for (int q = 1; q < NF; ++q)
		if (k - 1 == q) {
			fin = f.f[opposite[q]];
			fout = f.f[q];
		}
	// the first list entry and its loads go out before the list length and
	// the stall word are known (an entry beyond the list is clamped into the
	// grid and its values are not used)
	const int ncell = rows * cols;
	const size_t first = (((size_t)blockIdx.x * blockDim.x) >> 3) + (threadIdx.x >> 3);
	int c_pre = F.Lnew[min(first, (size_t)ncell - 1)];
	c_pre = min(max(c_pre, 0), ncell - 1);
	const WidthIn pre = width_load(c_pre, k, h, T, fin, fout, Mb, F.vent_act, v, rows, cols);
	if (*F.stall) return;
	const int cnt = F.cnt[tc];
	if (cnt == 0) {
		// no member in P: the set is empty unless border cells were set aside
		if (*F.skip == 0 && blockIdx.x == 0 && threadIdx.x == 0) *F.stall = F.t;
		return;
	}
	const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 3;
	for (size_t idx = first; idx < (size_t)cnt; idx += stride) {
		const WidthIn in = idx == first ? pre : width_load(F.Lnew[idx], k, h, T, fin, fout, Mb, F.vent_act, v, rows, cols);
		const int c = in.c;
		double ih = in.ih, it = in.it;
		if (in.vk >= 0 && in.vact == F.t) {
			ih = ih + v.e[in.vk];
			it = P.PTvent;
		}
		const int qk = in.qk;
		const double in_k = in.in_k;
		const double out_k = in.out_k;
		double nt_k = in.nt_k;
		if (qk >= 0 && in.vn >= 0 && in.vnact == F.t) nt_k = P.PTvent;
		const double ht_k = in_k * nt_k;
		double ht = 0.0, in_sum = 0.0, out_sum = 0.0;
#pragma unroll
		for (int m = 0; m < 8; ++m) {
			ht += __shfl(ht_k, g0 + m, 64);
			in_sum += __shfl(in_k, g0 + m, 64);
			out_sum += __shfl(out_k, g0 + m, 64);
		}
		double h1 = ih;
		h1 += in_sum - out_sum;
		double t1 = it;
		if (in_sum > 0 || out_sum > 0) {
			double residual = ih;
			residual -= out_sum;
			t1 = (residual * it + ht) / (residual + in_sum);
		}
		double h2 = h1, t2 = t1;
		const bool mb = in.mb;
		if (h1 > 0 && !mb) {
			const double nT = t1;
			const double num = (((3.0 * cr_cube(nT)) * P.Peps) * P.Psigma * P.Pclock) * P.Pcool;
			const double den = ((P.Prho * P.Pcv) * h1) * P.Pac;
			const double aus = 1.0 + num / den;
			const double st = nT / crz_pow_third(aus);
			t2 = st;
			if (st <= P.PTsol) {
				if (writer) {
					z[c] = z[c] + h1;
					Msl[c] = Msl[c] + h1;
				}
				h2 = 0.0;
				t2 = P.PTsol;
			}
		}
		const bool surv = !(t2 <= P.PTsol && !mb);
		if (writer) {
			// the steering's border reset (Sciara.cpp:189-194) after the removal
			h[c] = mb ? 0.0 : h2;
			Tn[c] = mb ? 0.0 : t2;
			flags[c] = surv ? 1 : 0;
		}
	}
}

// finishing the fused transitions: count[0] = the last list's length, the
// other counters cleared (sc_finish then runs over that list)
__global__ void sc_fused_counts(int* count, const int* cnt_slot)
{
	count[0] = *cnt_slot;
	count[1] = 0;
	count[2] = 0;
	count[3] = 0;
}

// steering alone (the run loop called the twin separately from the transition)
template <bool ACTIVE>
__global__ __launch_bounds__(SC_BLOCK) void sc_steer(double* __restrict__ h, double* __restrict__ T, Flows8 f,
                                                     const uint8_t* __restrict__ Mb, const int* __restrict__ cells,
                                                     const int* __restrict__ count,
                                                     const int* __restrict__ mb_cells, int n_mb, int rows,
                                                     int cols)
{
	const size_t n = (size_t)rows * cols;
	const int cnt = ACTIVE ? count[0] : 0;
	const size_t total = sweep_total(cnt, n);
	const size_t stride = (size_t)gridDim.x * blockDim.x;
	for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
		const int c = cnt > 0 ? cells[idx] : (int)idx;
#pragma unroll
		for (int q = 0; q < NF; ++q) f.f[q][c] = 0.0;
		if (!ACTIVE && Mb[c]) {
			h[c] = 0.0;
			T[c] = 0.0;
		}
	}
	if (ACTIVE)
		for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n_mb; k += stride) {
			const int c = mb_cells[k];
			h[c] = 0.0;
			T[c] = 0.0;
		}
}

int clamp_grid(size_t work, int lo, int hi)
{
	size_t b = (work + SC_BLOCK - 1) / SC_BLOCK;
	return (int)std::max<size_t>((size_t)lo, std::min<size_t>((size_t)hi, b));
}

struct SciaraProgram : Program {
	SubX *z = nullptr, *h = nullptr, *T = nullptr, *mb = nullptr, *msl = nullptr;
	SubX* f[NF] = {};
	calhipSciaraBinding b{};
	std::vector<int> vent_rc;
	std::vector<double> rates;
	bool active = false;

	int* alt = nullptr;          // second list buffer (compaction target)
	int* d_mb = nullptr;         // border-mask cells (linear indices)
	int n_mb = 0, cap_mb = 0;
	bool unordered = false;      // the device list is not in row-major order
	// launch sizing: an estimate of the list length, refreshed from an
	// asynchronous read-back of count[0] every few transitions
	int* h_cnt = nullptr;
	hipEvent_t ev_cnt = nullptr;
	bool cnt_pending = false;
	long long est = 0;
	long long steps_done = 0;

	// fused transitions (sc_flows_f / sc_width_f)
	bool fz_on = false;
	int fz_t = 1;                  // number of the next fused transition (from 1)
	int* fz_buf = nullptr;         // cnt[3], skip, stall, vent_act[VENTS_PER_LAUNCH]
	bool fz_skip_border = false;   // no vent on the border mask: idle border cells are set aside
	int* fz_lists[2] = {nullptr, nullptr};   // the list buffers (d_cells, alt at entry)
	std::vector<double> fz_elapsed;           // elapsed time at the start of each fused transition
	bool fz_replaying = false;                // running stalled transitions again (four launches)

	~SciaraProgram() override
	{
		if (alt) (void)hipFree(alt);
		if (d_mb) (void)hipFree(d_mb);
		if (h_cnt) (void)hipHostFree(h_cnt);
		if (ev_cnt) (void)hipEventDestroy(ev_cnt);
		if (fz_buf) (void)hipFree(fz_buf);
	}

	const char* name() const override { return "SCIARA-fv2"; }

	SciaraParams params() const
	{
		SciaraParams P;
		P.Pc = *b.Pc;
		P.Pac = *b.Pac;
		P.a = *b.a;
		P.b = *b.b;
		P.c = *b.c;
		P.d = *b.d;
		P.rad2 = *b.rad2;
		P.PTsol = *b.PTsol;
		P.PTvent = *b.PTvent;
		P.Pclock = *b.Pclock;
		P.Pcool = *b.Pcool;
		P.Prho = *b.Prho;
		P.Peps = *b.Pepsilon;
		P.Psigma = *b.Psigma;
		P.Pcv = *b.Pcv;
		return P;
	}

	Flows8 flows() const
	{
		Flows8 F;
		for (int q = 0; q < NF; ++q) F.f[q] = (double*)f[q]->d_cur;
		return F;
	}

	int after_upload(ModelX* m) override
	{
		size_t n = m->cells();
		const CALbyte* hm = mb->pub.b.current;
		std::vector<int> lst;
		for (size_t c = 0; c < n; ++c)
			if (hm[c]) lst.push_back((int)c);
		n_mb = (int)lst.size();
		if (n_mb > cap_mb) {
			if (d_mb) (void)hipFree(d_mb);
			d_mb = nullptr;
			cap_mb = 0;
			if (CALHIP_CHECK(hipMalloc(&d_mb, sizeof(int) * (size_t)n_mb))) return -1;
			cap_mb = n_mb;
		}
		if (n_mb && CALHIP_CHECK(hipMemcpyAsync(d_mb, lst.data(), sizeof(int) * (size_t)n_mb,
		                                        hipMemcpyHostToDevice, m->stream)))
			return -1;
		if (active) {
			if (!alt && CALHIP_CHECK(hipMalloc(&alt, sizeof(int) * n))) return -1;
			if (!h_cnt && CALHIP_CHECK(hipHostMalloc((void**)&h_cnt, sizeof(int), hipHostMallocDefault)))
				return -1;
			if (!ev_cnt && CALHIP_CHECK(hipEventCreateWithFlags(&ev_cnt, hipEventDisableTiming))) return -1;
			// counters 1..3 start cleared; count[0] came with the list
			if (CALHIP_CHECK(hipMemsetAsync(m->d_count + 1, 0, 3 * sizeof(int), m->stream))) return -1;
			est = m->pub.A.cells ? m->pub.A.size_current : 0;
			cnt_pending = false;
		}
		unordered = false;
		fz_on = false;    // the upload replaced the state the fused transitions kept
		fz_elapsed.clear();
		return CALHIP_CHECK(hipStreamSynchronize(m->stream));
	}

	int sync_active(ModelX* m) override
	{
		if (fz_finish(m)) return -1;
		if (!unordered) return 0;
		// flags equal the set; rebuild the row-major list from them
		if (k::rebuild_active(m)) return -1;
		unordered = false;
		return 0;
	}

	int check_entry(ModelX* m)
	{
		SubX* dl[3 + NF] = {z, h, T};
		for (int q = 0; q < NF; ++q) dl[3 + q] = f[q];
		for (SubX* s : dl)
			if (s->nxt_dirty || s->tstate != Transient::None) {
				set_error(10, "SCIARA device program: a substate has next-layer writes pending at the start of a "
				              "step (call calUpdate2D first)");
				return -1;
			}
		(void)m;
		return 0;
	}

	// emission of every vent for the current elapsed time (vent.h:74-82),
	// one entry per distinct cell: the last listed vent with a positive
	// thickness sets it (Sciara.cpp:5-15 overwrite the next layer)
	void vent_list(std::vector<std::pair<int, double>>& out, int cols) const
	{
		out.clear();
		const double el = *b.elapsed_time;
		const unsigned et = *b.emission_time;
		const unsigned i = (unsigned)(el / et);
		for (int v = 0; v < b.n_vents; ++v) {
			double e = i >= (unsigned)b.n_rates ? 0.0 : rates[(size_t)v * b.n_rates + i] / *b.Pac * *b.Pclock;
			if (!(e > 0)) continue;
			int c = vent_rc[2 * v] * cols + vent_rc[2 * v + 1];
			bool found = false;
			for (auto& p : out)
				if (p.first == c) {
					p.second = e;
					found = true;
				}
			if (!found) out.push_back({c, e});
		}
	}

	int launch_vents(ModelX* m, bool by_list)
	{
		std::vector<std::pair<int, double>> vl;
		vent_list(vl, m->pub.columns);
		for (size_t s = 0; s < vl.size(); s += VENTS_PER_LAUNCH) {
			VentArgs va;
			va.n = (int)std::min<size_t>(VENTS_PER_LAUNCH, vl.size() - s);
			for (int k = 0; k < va.n; ++k) {
				va.cell[k] = vl[s + k].first;
				va.e[k] = vl[s + k].second;
			}
			if (by_list) {
				hipLaunchKernelGGL(sc_vents_list, dim3(clamp_grid((size_t)est + 1, 1, 1024)), dim3(SC_BLOCK), 0,
				                   m->stream, (double*)h->d_cur, (double*)T->d_cur, m->d_cells, m->d_count, va,
				                   *b.PTvent, m->cells());
			} else {
				hipLaunchKernelGGL(sc_vents, dim3(1), dim3(64), 0, m->stream, (double*)h->d_cur,
				                   (double*)T->d_cur, active ? (const uint8_t*)m->d_flags : nullptr,
				                   active ? (const int*)m->d_count : nullptr, va, *b.PTvent);
			}
			m->stats.launches++;
		}
		return 0;
	}

	void refresh_estimate()
	{
		if (cnt_pending && hipEventQuery(ev_cnt) == hipSuccess) {
			est = *h_cnt;
			cnt_pending = false;
		}
	}

	// blocks for `lanes` threads per listed cell
	int list_grid(ModelX* m, int lanes = 1) const
	{
		if (!active) return clamp_grid(m->cells() * (size_t)lanes, 1, 8192);
		// room for growth between read-backs; the kernels loop over any excess
		return clamp_grid((size_t)(est + est / 4 + 1024) * (size_t)lanes, 8, 8192);
	}

	// ---- fused transitions ----

	bool fz_eligible(ModelX* m) const
	{
		// CALHIP_SCIARA_TWO_LAUNCH=0 keeps every transition in four launches
		const char* sw = getenv("CALHIP_SCIARA_TWO_LAUNCH");
		if (!active || fz_replaying || m->flags_dirty || (sw && strcmp(sw, "0") == 0)) return false;
		std::vector<std::pair<int, double>> vl;
		vent_list(vl, m->pub.columns);
		return vl.size() <= (size_t)VENTS_PER_LAUNCH;
	}

	FusedCtl fz_ctl(ModelX* m, int t) const
	{
		FusedCtl F;
		const int s = t - 1;   // transitions since the entry
		F.P = fz_lists[s % 2];
		F.Lnew = fz_lists[(s + 1) % 2];
		F.cnt = fz_buf;
		F.skip = fz_buf + 3;
		F.stall = fz_buf + 4;
		F.vent_act = fz_buf + 5;
		F.count0 = m->d_count;
		F.t = t;
		F.entry = t == 1 ? 1 : 0;
		F.skip_border = fz_skip_border ? 1 : 0;
		return F;
	}

	int fused_transition(ModelX* m)
	{
		if (check_entry(m)) return -1;
		refresh_estimate();
		if (!fz_on) {
			const size_t bytes = sizeof(int) * (5 + VENTS_PER_LAUNCH);
			if (!fz_buf && CALHIP_CHECK(hipMalloc(&fz_buf, bytes))) return -1;
			if (CALHIP_CHECK(hipMemsetAsync(fz_buf, 0, bytes, m->stream))) return -1;
			fz_lists[0] = m->d_cells;
			fz_lists[1] = alt;
			// the border mask is read-only for the program: its host copy is
			// the one uploaded
			const CALbyte* hm = mb->pub.b.current;
			fz_skip_border = true;
			for (int v = 0; v < b.n_vents; ++v)
				if (hm[(size_t)vent_rc[2 * v] * m->pub.columns + vent_rc[2 * v + 1]]) fz_skip_border = false;
			fz_t = 1;
			fz_elapsed.clear();
			fz_on = true;
		}
		const SciaraParams P = params();
		const int rows = m->pub.rows, cols = m->pub.columns;
		std::vector<std::pair<int, double>> vl;
		vent_list(vl, cols);
		VentArgs va;
		va.n = (int)vl.size();
		for (int k = 0; k < va.n; ++k) {
			va.cell[k] = vl[k].first;
			va.e[k] = vl[k].second;
		}
		fz_elapsed.push_back(*b.elapsed_time);
		const FusedCtl F = fz_ctl(m, fz_t);
		const int grid8 = list_grid(m, 8);
		hipEvent_t ev;
		timer_begin(m, &ev);
		hipLaunchKernelGGL(sc_flows_f, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (const double*)z->d_cur,
		                   (const double*)h->d_cur, (double*)T->d_cur, (const double*)T->d_nxt, flows(), m->d_flags,
		                   (const uint8_t*)mb->d_cur, F, va, rows, cols, P);
		hipLaunchKernelGGL(sc_width_f, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (double*)z->d_cur,
		                   (double*)h->d_cur, (const double*)T->d_cur, (double*)T->d_nxt, flows(),
		                   (const uint8_t*)mb->d_cur, (double*)msl->d_cur, m->d_flags, F, va, rows, cols, P);
		timer_end(m, ev);
		m->stats.launches += 2;
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		*b.elapsed_time += *b.Pclock;
		if ((++steps_done & 15) == 0 && !cnt_pending) {
			if (CALHIP_CHECK(hipMemcpyAsync(h_cnt, fz_buf + fz_t % 3, sizeof(int), hipMemcpyDeviceToHost,
			                                m->stream)) ||
			    CALHIP_CHECK(hipEventRecord(ev_cnt, m->stream)))
				return -1;
			cnt_pending = true;
		}
		fz_t++;
		unordered = true;
		T->nxt_stale = true;
		return 0;
	}

	// Back to the four-launch representation: the last completed fused
	// transition finished by sc_finish (temperatures copied back, the
	// steering's flow reset, the set compacted into the list), the list
	// rebuilt from the flags when border cells were set aside; transitions
	// that found the set empty run again through the four launches.
	int fz_finish(ModelX* m)
	{
		if (!fz_on) return 0;
		fz_on = false;
		int skip_stall[2] = {0, 0};
		if (CALHIP_CHECK(hipMemcpyAsync(skip_stall, fz_buf + 3, sizeof(skip_stall), hipMemcpyDeviceToHost,
		                                m->stream)) ||
		    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
			return -1;
		const int skipped = skip_stall[0], stall = skip_stall[1];
		const int last = stall ? stall - 1 : fz_t - 1;   // last completed fused transition
		if (last >= 1) {
			int* done_list = fz_lists[last % 2];
			int* other = fz_lists[(last + 1) % 2];
			hipLaunchKernelGGL(sc_fused_counts, dim3(1), dim3(1), 0, m->stream, m->d_count, fz_buf + last % 3);
			hipLaunchKernelGGL((sc_finish<true, true>), dim3(list_grid(m)), dim3(SC_BLOCK), 0, m->stream,
			                   (double*)h->d_cur, (double*)T->d_cur, (const double*)T->d_nxt, flows(),
			                   (const uint8_t*)mb->d_cur, (const uint8_t*)m->d_flags, done_list, other, m->d_count,
			                   d_mb, n_mb, m->pub.rows, m->pub.columns);
			m->stats.launches += 2;
			if (CALHIP_CHECK(hipGetLastError())) return -1;
			m->d_cells = other;
			alt = done_list;
			unordered = true;
			if (skipped) {
				// the set-aside border cells back into the row-major list
				if (k::rebuild_active(m)) return -1;
				unordered = false;
			}
		} else {
			m->d_cells = fz_lists[0];
			alt = fz_lists[1];
		}
		if (stall) {
			// the transitions from `stall` on left the state as it was: run
			// them again, each with the elapsed time it started from
			const int redo = fz_t - stall;
			const double end = *b.elapsed_time;
			*b.elapsed_time = fz_elapsed[(size_t)(stall - 1)];
			fz_elapsed.clear();
			fz_replaying = true;
			m->stats.reruns += redo;
			for (int r = 0; r < redo; ++r)
				if (transition(m, true)) {
					fz_replaying = false;
					return -1;
				}
			fz_replaying = false;
			if (*b.elapsed_time != end) {
				set_error(12, "SCIARA fused transitions: elapsed time differs after the re-run");
				return -1;
			}
		}
		fz_elapsed.clear();
		return 0;
	}

	int before_read(ModelX* m) override { return fz_finish(m); }
	int before_modify(ModelX* m, SubX*) override { return fz_finish(m); }

	int transition(ModelX* m, bool steer)
	{
		if (steer && fz_eligible(m)) return fused_transition(m);
		if (fz_finish(m)) return -1;
		if (check_entry(m)) return -1;
		refresh_estimate();
		const SciaraParams P = params();
		const int rows = m->pub.rows, cols = m->pub.columns;
		const Flows8 F = flows();
		hipEvent_t ev;
		timer_begin(m, &ev);
		// host changes to the set since the last rebuild: the vent process
		// still sees the old list, then the rebuild applies them
		const bool pending_host = active && m->flags_dirty;
		if (launch_vents(m, pending_host)) return -1;
		if (pending_host) {
			if (device_rebuild_active(m)) return -1;
			unordered = false;
		}
		const int grid = list_grid(m);
		const int grid8 = list_grid(m, 8);   // eight lanes per cell
		if (active)
			hipLaunchKernelGGL(sc_flows<true>, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (const double*)z->d_cur,
			                   (const double*)h->d_cur, (const double*)T->d_cur, F, m->d_flags, m->d_cells,
			                   m->d_count, rows, cols, P, 1);
		else
			hipLaunchKernelGGL(sc_flows<false>, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (const double*)z->d_cur,
			                   (const double*)h->d_cur, (const double*)T->d_cur, F, nullptr, nullptr, nullptr,
			                   rows, cols, P, 0);
		double* Tn = (double*)T->d_nxt;
		if (active)
			hipLaunchKernelGGL(sc_width<true>, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (double*)z->d_cur,
			                   (double*)h->d_cur, (const double*)T->d_cur, Tn, F, (const uint8_t*)mb->d_cur,
			                   (double*)msl->d_cur, m->d_flags, m->d_cells, m->d_count, rows, cols, P);
		else
			hipLaunchKernelGGL(sc_width<false>, dim3(grid8), dim3(SC_BLOCK), 0, m->stream, (double*)z->d_cur,
			                   (double*)h->d_cur, (const double*)T->d_cur, Tn, F, (const uint8_t*)mb->d_cur,
			                   (double*)msl->d_cur, nullptr, nullptr, nullptr, rows, cols, P);
		if (active) {
			if (steer)
				hipLaunchKernelGGL((sc_finish<true, true>), dim3(grid), dim3(SC_BLOCK), 0, m->stream,
				                   (double*)h->d_cur, (double*)T->d_cur, (const double*)Tn, F,
				                   (const uint8_t*)mb->d_cur, (const uint8_t*)m->d_flags, m->d_cells, alt,
				                   m->d_count, d_mb, n_mb, rows, cols);
			else
				hipLaunchKernelGGL((sc_finish<true, false>), dim3(grid), dim3(SC_BLOCK), 0, m->stream,
				                   (double*)h->d_cur, (double*)T->d_cur, (const double*)Tn, F,
				                   (const uint8_t*)mb->d_cur, (const uint8_t*)m->d_flags, m->d_cells, alt,
				                   m->d_count, d_mb, n_mb, rows, cols);
			std::swap(m->d_cells, alt);
			unordered = true;
		} else {
			if (steer)
				hipLaunchKernelGGL((sc_finish<false, true>), dim3(grid), dim3(SC_BLOCK), 0, m->stream,
				                   (double*)h->d_cur, (double*)T->d_cur, (const double*)Tn, F,
				                   (const uint8_t*)mb->d_cur, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
				                   rows, cols);
			else
				hipLaunchKernelGGL((sc_finish<false, false>), dim3(grid), dim3(SC_BLOCK), 0, m->stream,
				                   (double*)h->d_cur, (double*)T->d_cur, (const double*)Tn, F,
				                   (const uint8_t*)mb->d_cur, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
				                   rows, cols);
		}
		m->stats.launches += 3;
		if (steer) *b.elapsed_time += *b.Pclock;
		timer_end(m, ev);
		if (CALHIP_CHECK(hipGetLastError())) return -1;
		if (active && (++steps_done & 15) == 0 && !cnt_pending) {
			if (CALHIP_CHECK(hipMemcpyAsync(h_cnt, m->d_count, sizeof(int), hipMemcpyDeviceToHost, m->stream)) ||
			    CALHIP_CHECK(hipEventRecord(ev_cnt, m->stream)))
				return -1;
			cnt_pending = true;
		}
		// every layer was written in place; the next layers are logically
		// equal to the current ones (St's next buffer held scratch)
		T->nxt_stale = true;
		return 0;
	}

	int step(ModelX* m) override { return transition(m, false); }
	int step_and_steer(ModelX* m) override { return transition(m, true); }

	// Sciara.cpp:185-197 on the device
	int steer(ModelX* m) override
	{
		if (fz_finish(m)) return -1;
		const int rows = m->pub.rows, cols = m->pub.columns;
		refresh_estimate();
		const int grid = list_grid(m);
		if (active && m->flags_dirty) {
			// the reset covers the list as OpenCAL holds it (rebuilt at updates)
			set_error(11, "SCIARA steering twin: the active set has changes pending (call calUpdate2D first)");
			return -1;
		}
		if (active)
			hipLaunchKernelGGL(sc_steer<true>, dim3(grid), dim3(SC_BLOCK), 0, m->stream, (double*)h->d_cur,
			                   (double*)T->d_cur, flows(), (const uint8_t*)mb->d_cur, m->d_cells, m->d_count, d_mb,
			                   n_mb, rows, cols);
		else
			hipLaunchKernelGGL(sc_steer<false>, dim3(grid), dim3(SC_BLOCK), 0, m->stream, (double*)h->d_cur,
			                   (double*)T->d_cur, flows(), (const uint8_t*)mb->d_cur, nullptr, nullptr, nullptr, 0,
			                   rows, cols);
		m->stats.launches++;
		*b.elapsed_time += *b.Pclock;
		return CALHIP_CHECK(hipGetLastError());
	}
};

bool real_layered(struct CALSubstate2Dr* q) { return q && S(q)->type == 'r' && !S(q)->single; }

}  // namespace

}  // namespace calhip

using namespace calhip;

extern "C" int calhipBindSciara2D(struct CALModel2D* ca2D, const struct calhipSciaraBinding* b)
{
	if (!ca2D || !b) {
		set_error(30, "calhipBindSciara2D: missing argument");
		return -1;
	}
	ModelX* m = X(ca2D);
	if (ca2D->X_id != CAL_MOORE_NEIGHBORHOOD_2D || ca2D->T != CAL_SPACE_FLAT) {
		set_error(31, "calhipBindSciara2D: the SCIARA program needs a Moore, CAL_SPACE_FLAT model");
		return -1;
	}
	bool ok = real_layered(b->Sz) && real_layered(b->Slt) && real_layered(b->St) && b->Mb &&
	          S(b->Mb)->type == 'b' && b->Msl && S(b->Msl)->type == 'r';
	for (int q = 0; q < NF; ++q) ok = ok && real_layered(b->f[q]);
	const CALreal* const prm[] = {b->Pclock, b->PTsol, b->PTvent, b->Pc, b->Pac, b->a, b->b, b->c,
	                              b->d, b->rad2, b->Pcool, b->Prho, b->Pepsilon, b->Psigma, b->Pcv};
	for (const CALreal* p : prm) ok = ok && p;
	ok = ok && b->vents && b->flows && b->width && b->temperature && b->elapsed_time && b->emission_time;
	ok = ok && b->n_vents >= 0 && b->n_rates >= 0 && (b->n_vents == 0 || b->vent_rc) &&
	     (b->n_vents * b->n_rates == 0 || b->rates);
	if (!ok) {
		set_error(32, "calhipBindSciara2D: missing or mistyped substate, process or parameter (Sz, Slt, St, f[8] "
		              "double-layer real; Mb byte; Msl real)");
		return -1;
	}
	const bool active = ca2D->OPTIMIZATION == CAL_OPT_ACTIVE_CELLS;
	if (active != (b->remove != nullptr)) {
		set_error(33, "calhipBindSciara2D: the removal process is required with active cells and only then");
		return -1;
	}
	for (int v = 0; v < b->n_vents; ++v)
		if (b->vent_rc[2 * v] < 0 || b->vent_rc[2 * v] >= ca2D->rows || b->vent_rc[2 * v + 1] < 0 ||
		    b->vent_rc[2 * v + 1] >= ca2D->columns) {
			set_error(34, "calhipBindSciara2D: a vent lies outside the grid");
			return -1;
		}
	// single-layer substates the program does not use (the reference's Mv and
	// Sz_t0) stay on the host: OpenCAL code replaces and frees their arrays
	// (io.cpp loadVents / saveConfiguration)
	if (m->phase == Phase::Device && ensure_host(m)) return -1;
	for (SubX* s : m->subs)
		s->host_only = s->single && s != S(b->Mb) && s != S(b->Msl);
	auto* p = new SciaraProgram();
	p->b = *b;
	p->z = S(b->Sz);
	p->h = S(b->Slt);
	p->T = S(b->St);
	for (int q = 0; q < NF; ++q) p->f[q] = S(b->f[q]);
	p->mb = S(b->Mb);
	p->msl = S(b->Msl);
	p->active = active;
	p->vent_rc.assign(b->vent_rc, b->vent_rc + 2 * (size_t)b->n_vents);
	p->rates.assign(b->rates, b->rates + (size_t)b->n_vents * b->n_rates);
	p->b.vent_rc = p->vent_rc.data();
	p->b.rates = p->rates.data();
	delete m->prog;
	m->prog = p;
	m->prog_eps.clear();
	m->prog_eps.push_back(b->vents);
	m->prog_eps.push_back(b->flows);
	m->prog_eps.push_back(b->width);
	m->prog_eps.push_back(b->temperature);
	if (b->remove) m->prog_eps.push_back(b->remove);
	m->prog_steering = b->steering;
	// a model already on the device gets the program's set-up now
	if (m->phase == Phase::Device && p->after_upload(m)) return -1;
	return 0;
}
