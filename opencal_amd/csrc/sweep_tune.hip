// Launch shape of the fused SciddicaT sweep, its wet-first tile order and its
// autotuner.
//
// The sweep (step_fused_al for even column counts, step_fused2 for odd ones)
// gives every wave one 128-column (124 for odd counts) strip of a chunk of
// `rpw` rows, prefetching `depth` rows ahead.  Short chunks make more waves
// and bound the time of a wave that meets wet cells, but re-read more halo
// rows; long chunks and deeper prefetch stream better on large grids.  The
// autotuner times each candidate shape on the caller's stream with HIP events
// (a few launches each, writing into a buffer the caller is about to
// overwrite anyway) and keeps the fastest per grid shape for the life of the
// process.  Every candidate computes the same bits, so the choice never
// changes a result.
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_2step.hpp"

namespace calhip {
namespace sct {

// Tiles of the two-step sweep.  A small tile is `s_rows` rows of a group of
// FUSED_WAVES strips (one wave per strip); a big tile is `bigk` small tiles
// stacked vertically.  A wave that meets wet cells pays the minimisation's
// latency on every row of its chunk, twice per step pair, while a dry chunk
// is cheaper per row the longer it is (its four halo rows above and below
// are re-read and recomputed).  So the launch runs the small tiles that met
// wet cells in the previous launch first, one block each (wet-first order,
// as in step_fused_al), and every big tile as one chunk per wave of its
// runs of small tiles that are not in that list.  Every small tile is
// computed exactly once per launch, so the order and the grouping never
// change a result.  Each small tile records whether it met a wet cell (a
// run that met one marks all its small tiles).
struct Sweep2Geom {
	int n_strips, n_sg, s_rows, bigk, n_small_chunks, n_small, n_big, per_xcd;
};

// Records whether small tile S met a wet cell; a wet tile is listed for the
// next launch's wet blocks (a full list leaves it in its big tile).
__device__ __forceinline__ void wet2_record(const WetSets& ws, int grp, int S, bool wet)
{
	if (wet) {
		const unsigned jx = atomicAdd(ws.n_out + grp, 1u);
		if ((int)jx < ws.cap)
			ws.l_out[grp * ws.cap + jx] = S;
		else
			wet = false;
	}
	ws.f_out[S] = (unsigned char)(wet ? 1 : 0);
}

template <bool UP>
__device__ __forceinline__ bool sweep2_rows(const double* z, const double* h, double* hn, double* trash, int rows,
                                            int cols, int strip, int r0, int r1, double r, double eps)
{
	return sweep2_strip_al<UP>(z, h, hn, trash, rows, cols, cols, strip, r0, r1, r, eps);
}

// ws.f_out == nullptr: big tiles only (gridDim.x = 8 * per_xcd); otherwise
// gridDim.x = 8 * (per_xcd + ws.cap), the first ws.cap blocks of each XCD
// group taking the listed small tiles.
__global__ __launch_bounds__(64 * FUSED_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2))) void step_fused_2x(
    const double* __restrict__ z, const double* __restrict__ h, double* __restrict__ hn, double* __restrict__ trash,
    int rows, int cols, Sweep2Geom g, double r, double eps, WetSets ws)
{
	const int grp = blockIdx.x & 7, l = blockIdx.x >> 3;
	const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
	__shared__ unsigned blk_mask;
	if (threadIdx.x == 0) blk_mask = 0u;
	if (ws.f_out && blockIdx.x == 0 && threadIdx.x < 8) ws.n_zero[threadIdx.x] = 0u;
	__syncthreads();
	const int nw = ws.f_out ? min((int)ws.n_in[grp], ws.cap) : 0;
	if (l < nw) {
		const int S = ws.l_in[grp * ws.cap + l];
		const int cs = S / g.n_sg, sg = S % g.n_sg;
		const int strip = sg * FUSED_WAVES + w;
		bool wet = false;
		if (strip < g.n_strips) {
			const int r0 = cs * g.s_rows, r1 = min(r0 + g.s_rows, rows);
			wet = (cs & 1) ? sweep2_rows<true>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps)
			               : sweep2_rows<false>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps);
		}
		if (wet && lane == 0) atomicOr(&blk_mask, 1u);
		__syncthreads();
		if (threadIdx.x == 0) wet2_record(ws, grp, S, blk_mask != 0u);
		return;
	}
	const int k = l - nw;
	if (k >= g.per_xcd) return;
	const int B = grp * g.per_xcd + k;
	if (B >= g.n_big) return;
	const int cb = B / g.n_sg, sg = B % g.n_sg;
	const int strip = sg * FUSED_WAVES + w;
	const int cs0 = cb * g.bigk, cs1 = min(cs0 + g.bigk, g.n_small_chunks);
	unsigned listed = 0u;   // small tiles of this big tile that the wet blocks take
	if (ws.f_out)
		for (int c = cs0; c < cs1; ++c)
			if (ws.f_in[c * g.n_sg + sg]) listed |= 1u << (c - cs0);
	unsigned my_wet = 0u;
	if (strip < g.n_strips) {
		int c = cs0;
		while (c < cs1) {
			if ((listed >> (c - cs0)) & 1u) {
				++c;
				continue;
			}
			int e = c + 1;
			while (e < cs1 && !((listed >> (e - cs0)) & 1u)) ++e;
			const int r0 = c * g.s_rows, r1 = min(e * g.s_rows, rows);
			// a run that is dry within two cells is h + 0.0 (dry_copy)
			const bool wet = dry_copy(h, hn, rows, cols, cols, strip, r0, r1, eps)
			                     ? false
			                     : ((c & 1) ? sweep2_rows<true>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps)
			                                : sweep2_rows<false>(z, h, hn, trash, rows, cols, strip, r0, r1, r, eps));
			if (wet) my_wet |= ((1u << (e - c)) - 1u) << (c - cs0);
			c = e;
		}
	}
	if (my_wet && lane == 0) atomicOr(&blk_mask, my_wet);
	__syncthreads();
	if (threadIdx.x == 0 && ws.f_out)
		for (int c = cs0; c < cs1; ++c)
			if (!((listed >> (c - cs0)) & 1u)) wet2_record(ws, grp, c * g.n_sg + sg, (blk_mask >> (c - cs0)) & 1u);
}


SweepGeom sweep_geom(int rows, int cols, int rpw)
{
	SweepGeom g;
	g.rpw = rpw;
	// even column counts: aligned 128-column strips; odd: 124-column strips
	g.n_strips = cols % 2 == 0 ? (cols + STRIP_AL - 1) / STRIP_AL : (cols + STRIP2_OUT - 1) / STRIP2_OUT;
	if (cols % 2 == 0) {
		// aligned sweep: blocks of FUSED_WAVES consecutive (chunk, strip) items
		g.n_logical = (g.n_strips * ((rows + rpw - 1) / rpw) + FUSED_WAVES - 1) / FUSED_WAVES;
	} else {
		// 124-column sweep: blocks of FUSED_WAVES strips of one chunk
		const int sgroups = (g.n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
		g.n_logical = sgroups * ((rows + rpw - 1) / rpw);
	}
	g.per_xcd = (g.n_logical + 7) / 8;
	g.blocks = g.per_xcd * 8;
	return g;
}

// ------------------------------------------------------------- wet order

namespace {
constexpr size_t WO_COUNTS = 3 * 8 * sizeof(unsigned);
size_t wo_lists_off() { return 256; }
size_t wo_flags_off(int cap) { return wo_lists_off() + (((size_t)3 * 8 * cap * sizeof(int) + 255) & ~(size_t)255); }
}  // namespace

WetOrder::~WetOrder()
{
	if (buf) (void)hipFree(buf);
}

int WetOrder::next(hipStream_t st, int nl, WetSets& ws, int max_cap)
{
	int want_cap = (nl + 7) / 8;
	if (want_cap > max_cap) want_cap = max_cap;
	if (!buf || nl != n_logical || want_cap != cap) {
		if (buf) (void)hipFree(buf);
		buf = nullptr;
		n_logical = nl;
		// wet tiles beyond the capacity run in grid order; each unused slot
		// costs one block that exits at once
		cap = want_cap;
		const size_t bytes = wo_flags_off(cap) + 2 * (size_t)nl;
		static_assert(WO_COUNTS <= 256, "counts fit the first 256 bytes");
		if (CALHIP_CHECK(hipMalloc(&buf, bytes))) {
			buf = nullptr;
			return -1;
		}
		if (CALHIP_CHECK(hipMemsetAsync(buf, 0, bytes, st))) return -1;
		epoch = 0;
	}
	char* b = (char*)buf;
	unsigned* counts = (unsigned*)b;
	int* lists = (int*)(b + wo_lists_off());
	unsigned char* flags = (unsigned char*)(b + wo_flags_off(cap));
	const int si = (int)(epoch % 3), so = (int)((epoch + 1) % 3), sz = (int)((epoch + 2) % 3);
	ws.n_in = counts + 8 * si;
	ws.l_in = lists + (size_t)8 * cap * si;
	ws.n_out = counts + 8 * so;
	ws.l_out = lists + (size_t)8 * cap * so;
	ws.n_zero = counts + 8 * sz;
	ws.f_in = flags + (size_t)nl * (epoch & 1);
	ws.f_out = flags + (size_t)nl * ((epoch + 1) & 1);
	ws.cap = cap;
	epoch++;
	return 0;
}

void launch_sweep(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                  double r, double eps, SweepShape shape, WetOrder* wo)
{
	const SweepGeom g = sweep_geom(rows, cols, shape.rpw);
	if (cols % 2 != 0) {
		hipLaunchKernelGGL((step_fused2<1, false>), dim3(g.blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash,
		                   rows, cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps);
		return;
	}
	WetSets ws = {};
	int blocks = g.blocks;
	if (wo && wo->next(st, g.n_logical, ws) == 0)
		blocks = 8 * (g.per_xcd + ws.cap);
	else
		ws = WetSets{};
	if (shape.depth >= 2)
		hipLaunchKernelGGL((step_fused_al<2>), dim3(blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash, rows,
		                   cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps, ws);
	else
		hipLaunchKernelGGL((step_fused_al<1>), dim3(blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash, rows,
		                   cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps, ws);
}

bool sweep2_supported(int rows, int cols) { return cols % 2 == 0 && cols >= 8 && rows >= 8; }

void launch_sweep2(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                   double r, double eps, Sweep2Shape shape, WetOrder* wo)
{
	Sweep2Geom g;
	g.n_strips = (cols + STRIP_AL - 1) / STRIP_AL;
	g.n_sg = (g.n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
	g.s_rows = shape.s_rows;
	g.bigk = shape.bigk;
	g.n_small_chunks = (rows + g.s_rows - 1) / g.s_rows;
	g.n_small = g.n_small_chunks * g.n_sg;
	g.n_big = ((g.n_small_chunks + g.bigk - 1) / g.bigk) * g.n_sg;
	g.per_xcd = (g.n_big + 7) / 8;
	WetSets ws = {};
	int blocks = 8 * g.per_xcd;
	if (wo && wo->next(st, g.n_small, ws, 256) == 0)
		blocks = 8 * (g.per_xcd + ws.cap);
	else
		ws = WetSets{};
	hipLaunchKernelGGL(step_fused_2x, dim3(blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash, rows, cols, g, r,
	                   eps, ws);
}

// --------------------------------------------------------------- tuner

namespace {
std::mutex tune_mu;
std::map<std::pair<int, int>, SweepShape> tuned;
}  // namespace

SweepShape sweep_shape(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch,
                       double* trash, double r, double eps)
{
	SweepShape fixed = {0, 0};
	if (const char* e = getenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE")) fixed.rpw = atoi(e);
	if (const char* e = getenv("CALHIP_SCIDDICAT_PREFETCH")) fixed.depth = atoi(e);
	if (fixed.rpw > 0 && fixed.depth > 0) return fixed;
	std::lock_guard<std::mutex> lk(tune_mu);
	const auto key = std::make_pair(rows, cols);
	auto it = tuned.find(key);
	if (it != tuned.end()) {
		SweepShape s = it->second;
		if (fixed.rpw > 0) s.rpw = fixed.rpw;
		if (fixed.depth > 0) s.depth = fixed.depth;
		return s;
	}
	// short chunks with deep prefetch keep a wet front's waves short (at
	// 1024^2 with a spreading front, 4 rows at depth 2 run a step in about
	// 60% of the time of 8 - 16 rows); long ones stream dry grids best
	static const SweepShape cand[] = {{4, 2}, {4, 1}, {6, 2}, {8, 1},  {12, 1}, {16, 1}, {24, 1},
	                                  {32, 1}, {8, 2}, {16, 2}, {24, 2}, {32, 2}};
	hipEvent_t e0 = nullptr, e1 = nullptr;
	SweepShape best = {16, 1};
	float best_ms = 0.f;
	if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
		for (SweepShape c : cand) {
			if (fixed.rpw > 0 && c.rpw != fixed.rpw) continue;
			if (fixed.depth > 0 && c.depth != fixed.depth) continue;
			if (cols % 2 != 0 && c.depth != 1) continue;   // odd counts have one kernel depth
			WetOrder wo;   // its own order state: the first launch learns the wet tiles
			launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, c, &wo);
			(void)hipEventRecord(e0, st);
			for (int k = 0; k < 3; ++k) launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, c, &wo);
			(void)hipEventRecord(e1, st);
			float ms = 0.f;
			if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
			if (best_ms == 0.f || ms < best_ms) {
				best_ms = ms;
				best = c;
			}
		}
	}
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipGetLastError();
	tuned[key] = best;
	return best;
}

SweepShape sweep_shape_cached(int rows, int cols)
{
	std::lock_guard<std::mutex> lk(tune_mu);
	const auto it = tuned.find(std::make_pair(rows, cols));
	return it != tuned.end() ? it->second : SweepShape{16, 1};
}

namespace {
std::mutex tune2_mu;
std::map<std::pair<int, int>, Sweep2Shape> tuned2;
}  // namespace

Sweep2Shape sweep2_shape(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch,
                         double* trash, double r, double eps)
{
	// The environment fixes the tile shape (and so the two-step sweep).
	Sweep2Shape fixed = {0, 0};
	if (const char* e = getenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE_2X"))
		if (atoi(e) > 0) fixed.s_rows = atoi(e);
	if (const char* e = getenv("CALHIP_SCIDDICAT_2X_GROUP"))
		if (atoi(e) > 0) fixed.bigk = atoi(e);
	if (fixed.s_rows > 0 || fixed.bigk > 0) {
		if (fixed.s_rows <= 0) fixed.s_rows = 12;
		if (fixed.bigk <= 0) fixed.bigk = 4;
		return fixed;
	}
	// The one-step sweep's shape is tuned first (it has its own lock).
	const SweepShape one = sweep_shape(st, rows, cols, z, h, scratch, trash, r, eps);
	std::lock_guard<std::mutex> lk(tune2_mu);
	const auto key = std::make_pair(rows, cols);
	auto it = tuned2.find(key);
	if (it != tuned2.end()) return it->second;
	// Small tiles of s_rows rows in big tiles of bigk: at 4096^2 with a
	// spreading front 12 x 4 ran 101 us per launch (12 x 3: 107; 8 x 4: 103;
	// 16 x 2: 112); on small grids the big tiles are too few to fill the
	// chip, and shorter tiles, or one step per launch, win.  Each candidate
	// is timed (its own wet-first state, one launch to learn the wet tiles,
	// then three) against two launches of the tuned one-step sweep; {0, 0}
	// means one step per launch.
	static const Sweep2Shape cand[] = {{12, 4}, {12, 2}, {12, 1}, {8, 4}, {8, 2}, {8, 1}, {16, 2}, {6, 2}, {24, 1}};
	Sweep2Shape best = {12, 4};
	float best_ms = 0.f;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
		for (Sweep2Shape c : cand) {
			WetOrder wo;
			launch_sweep2(st, rows, cols, z, h, scratch, trash, r, eps, c, &wo);
			(void)hipEventRecord(e0, st);
			for (int k = 0; k < 3; ++k) launch_sweep2(st, rows, cols, z, h, scratch, trash, r, eps, c, &wo);
			(void)hipEventRecord(e1, st);
			float ms = 0.f;
			if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
			if (best_ms == 0.f || ms < best_ms) {
				best_ms = ms;
				best = c;
			}
		}
		WetOrder wo;
		launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, one, &wo);
		(void)hipEventRecord(e0, st);
		for (int k = 0; k < 6; ++k) launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, one, &wo);
		(void)hipEventRecord(e1, st);
		float ms1 = 0.f;
		if (best_ms > 0.f && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms1, e0, e1) == hipSuccess &&
		    ms1 < best_ms)
			best = Sweep2Shape{0, 0};   // six one-step launches (six steps) beat three two-step launches
	}
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipGetLastError();
	tuned2[key] = best;
	return best;
}

}  // namespace sct
}  // namespace calhip
