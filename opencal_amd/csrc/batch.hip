// Batched SciddicaT ensemble: many independent full-grid SciddicaT instances
// of uneven sizes and parameters stepped together.
//
// Layout.  Every instance keeps its own z and double-buffered h in one
// packed device allocation.  Rows of an instance with an even column count
// are padded to a pitch of 16 columns, so every row starts on a 128-byte
// line and the aligned 128-column strip sweep applies; odd column counts keep
// packed rows and the 124-column strip sweep.  An instance is cut into tiles
// (a strip of columns x a chunk of rows); the tiles of all instances are
// numbered in one flat list.  The sweep of a tile is the fused two-process
// sweep of the single-model program, so each instance's result equals its
// own plain run bit for bit.
//
// Wet-tile queue.  A tile whose cells and von Neumann halo are all dry
// (h <= eps) has no outflow and receives none, so its level after the step
// is h + 0.0: unchanged except that a negative zero becomes +0.  Hence a tile
// that was quiet (tile and halo dry) at step t-1 and whose four neighbour
// tiles were quiet too is quiet again at step t with the same values, and
// from step 2 on (once every value has passed through one sweep) both h
// buffers already hold those values: the tile need not be swept.  So each
// step sweeps only the tiles of a packed queue; a swept tile that met a wet
// cell queues itself and its four neighbour tiles for the next step (an
// atomic max on a per-tile step tag keeps each tile once per queue), and
// steps 0 and 1 sweep every tile.  Skipped tiles keep bits that equal the
// full sweep's.
//
// Launches.  The steps run in one cooperative launch: its waves pull tiles
// of the step's queue with an atomic counter until it is empty, and a grid
// barrier orders the step's writes (levels, next queue) before the next
// step.  Without cooperative launches each step is one launch.  Three queues
// rotate: step t reads queue t%3, appends to (t+1)%3 and clears (t+2)%3,
// which nothing reads or appends to during step t.
//
// Tile rows.  The chunk length (rows per tile) is chosen per ensemble by
// timing the first steps with each candidate length; the choice never
// changes a result.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_device.hpp"

namespace calhip {

namespace {

using namespace sct;

constexpr int BATCH_BLOCK = 64 * FUSED_WAVES;

struct BatchInst {
	int rows, cols, strips, pitch;   // pitch: row stride in elements
	int chunks, first;               // row chunks; index of the instance's first tile
	long long off;                   // element offset of the instance's grids
	double r, eps;
};

// Queue state, one device allocation: counters, tags and three queues.
struct BatchQueues {
	int* count;   // [0..2] queue lengths, [3..5] pull counters
	int* tag;     // [n_tiles] last step the tile was queued for
	int* q[3];    // [n_tiles] each
};

__device__ __forceinline__ void queue_tile(const BatchQueues& Q, int tile, int step)
{
	if (atomicMax(&Q.tag[tile], step) < step) Q.q[step % 3][atomicAdd(&Q.count[step % 3], 1)] = tile;
}

// Step t of every instance.  Each wave pulls tiles of the step's queue
// until it is empty (the pull counter grows by one per pull, so every wave
// leaves the loop).  The loop's control is wave-uniform by construction: the
// queue length and the pulled index are scalars, and every lane takes part in
// the pull (lane 0 adds 1, the others 0), so no lane can leave the loop or
// re-enter it apart from the others.
template <int D>
__device__ __forceinline__ void batch_step_body(const BatchInst* __restrict__ inst,
                                                const int* __restrict__ tile_inst, int rpw,
                                                const double* __restrict__ z, double* __restrict__ h0,
                                                double* __restrict__ h1, double* __restrict__ trash, BatchQueues Q,
                                                int t)
{
	const int lane = threadIdx.x & 63;
	const int qi = t % 3;
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		Q.count[(t + 2) % 3] = 0;
		Q.count[3 + (t + 2) % 3] = 0;
	}
	const double* h = (t & 1) ? h1 : h0;
	double* hn = (t & 1) ? h0 : h1;
	const int n_q = __builtin_amdgcn_readfirstlane(*(volatile int*)&Q.count[qi]);
	for (;;) {
		const int k = __builtin_amdgcn_readfirstlane(atomicAdd(&Q.count[3 + qi], lane == 0 ? 1 : 0));
		if (k >= n_q) break;
		const int tile = __builtin_amdgcn_readfirstlane(Q.q[qi][k]);
		const int ii = __builtin_amdgcn_readfirstlane(tile_inst[tile]);
		const BatchInst I = inst[ii];
		const int local = tile - I.first;
		const int chunk = local / I.strips, strip = local - chunk * I.strips;
		const int r0 = chunk * rpw, r1 = min(r0 + rpw, I.rows);
		bool wet;
		if (I.cols % 2 == 0)
			wet = sweep_chunk_al<D>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, I.pitch, strip, chunk,
			                        r0, r1, I.r, I.eps);
		else
			wet = sweep_chunk<1, false>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, strip, chunk, r0,
			                            r1, I.r, I.eps);
		// the tiles to sweep at step t+1: this one while it is wet or
		// before step 2, its four neighbours while it is wet
		if (lane == 0) {
			if (wet || t + 1 < 2) queue_tile(Q, tile, t + 1);
			if (wet) {
				const int cn = chunk == 0 ? I.chunks - 1 : chunk - 1;
				const int cs = chunk + 1 == I.chunks ? 0 : chunk + 1;
				const int sw = strip == 0 ? I.strips - 1 : strip - 1;
				const int se = strip + 1 == I.strips ? 0 : strip + 1;
				queue_tile(Q, I.first + cn * I.strips + strip, t + 1);
				queue_tile(Q, I.first + cs * I.strips + strip, t + 1);
				queue_tile(Q, I.first + chunk * I.strips + sw, t + 1);
				queue_tile(Q, I.first + chunk * I.strips + se, t + 1);
			}
		}
	}
}


// One step per launch (the launch boundary orders the step's writes before
// the next step).
template <int D>
__global__ __launch_bounds__(BATCH_BLOCK) __attribute__((amdgpu_waves_per_eu(D == 1 ? 3 : 2))) void batch_step(
    const BatchInst* __restrict__ inst, const int* __restrict__ tile_inst, int rpw, const double* __restrict__ z,
    double* __restrict__ h0, double* __restrict__ h1, double* __restrict__ trash, BatchQueues Q, int t)
{
	batch_step_body<D>(inst, tile_inst, rpw, z, h0, h1, trash, Q, t);
}

// Steps [t0, t1) in one cooperative launch: every block is resident, and a
// grid barrier (with its device-scope release and acquire) separates the
// steps.  Every wave leaves each step's loop and reaches every barrier.
template <int D>
__global__ __launch_bounds__(BATCH_BLOCK) __attribute__((amdgpu_waves_per_eu(D == 1 ? 3 : 2))) void batch_steps_coop(
    const BatchInst* __restrict__ inst, const int* __restrict__ tile_inst, int rpw, const double* __restrict__ z,
    double* __restrict__ h0, double* __restrict__ h1, double* __restrict__ trash, BatchQueues Q, int t0, int t1)
{
	for (int t = t0; t < t1; ++t) {
		batch_step_body<D>(inst, tile_inst, rpw, z, h0, h1, trash, Q, t);
		cooperative_groups::this_grid().sync();
	}
}

struct DevBuf {
	void* p = nullptr;
	~DevBuf()
	{
		if (p) (void)hipFree(p);
	}
};

// The packed layout (independent of the chunk length) and the tile list of
// one ensemble for a chunk length.
struct Layout {
	std::vector<BatchInst> inst;
	std::vector<int> tile_inst;
	long long total = 0;   // elements of one packed field
};

Layout make_layout(int n, const int* rows, const int* cols, const double* p_r, const double* p_eps, int rpw)
{
	Layout L;
	L.inst.resize(n);
	for (int k = 0; k < n; ++k) {
		BatchInst& I = L.inst[k];
		const bool even = cols[k] % 2 == 0;
		I.rows = rows[k];
		I.cols = cols[k];
		I.strips = even ? (cols[k] + STRIP_AL - 1) / STRIP_AL : (cols[k] + STRIP2_OUT - 1) / STRIP2_OUT;
		I.pitch = even ? (cols[k] + 15) / 16 * 16 : cols[k];
		I.chunks = (rows[k] + rpw - 1) / rpw;
		I.first = (int)L.tile_inst.size();
		I.off = L.total;
		I.r = p_r[k];
		I.eps = p_eps[k];
		L.total += ((long long)rows[k] * I.pitch + 31) / 32 * 32;   // 256-byte aligned starts
		L.tile_inst.insert(L.tile_inst.end(), (size_t)I.chunks * I.strips, k);
	}
	return L;
}

// The ensemble on the device: z after the reference init (z -= h where
// h > 0) and the initial h uploaded once; each run restores both h buffers
// from the initial copy and steps with a given chunk length.
struct Runner {
	int n;
	const int *rows, *cols;
	const double *p_r, *p_eps;
	hipStream_t stream = nullptr;
	long long total = 0;
	DevBuf dz, dhi, dh0, dh1, dtrash;

	int prepare(const double* const* z_host, const double* const* h_host)
	{
		const Layout L = make_layout(n, rows, cols, p_r, p_eps, 32);
		total = L.total;
		std::vector<double> hz((size_t)total, 0.0), hh((size_t)total, 0.0);
		for (int k = 0; k < n; ++k)
			for (int i = 0; i < rows[k]; ++i) {
				const double* zs = z_host[k] + (size_t)i * cols[k];
				const double* hs = h_host[k] + (size_t)i * cols[k];
				double* zz = hz.data() + L.inst[k].off + (size_t)i * L.inst[k].pitch;
				double* h0 = hh.data() + L.inst[k].off + (size_t)i * L.inst[k].pitch;
				for (int j = 0; j < cols[k]; ++j) {
					const double hv = hs[j];
					h0[j] = hv;
					zz[j] = hv > 0.0 ? zs[j] - hv : zs[j];
				}
			}
		const size_t bytes = (size_t)total * sizeof(double);
		if (CALHIP_CHECK(hipMalloc(&dz.p, bytes)) || CALHIP_CHECK(hipMalloc(&dhi.p, bytes)) ||
		    CALHIP_CHECK(hipMalloc(&dh0.p, bytes)) || CALHIP_CHECK(hipMalloc(&dh1.p, bytes)) ||
		    CALHIP_CHECK(hipMalloc(&dtrash.p, 128 * sizeof(double))))
			return -1;
		if (CALHIP_CHECK(hipMemcpyAsync(dz.p, hz.data(), bytes, hipMemcpyHostToDevice, stream)) ||
		    CALHIP_CHECK(hipMemcpyAsync(dhi.p, hh.data(), bytes, hipMemcpyHostToDevice, stream)))
			return -1;
		return CALHIP_CHECK(hipStreamSynchronize(stream));
	}

	// `steps` steps with chunk length rpw; the stepping time in ms (< 0 on
	// error); with h_out the final thickness is copied out.
	float run(int rpw, int depth, int steps, int per_launch, double* const* h_out)
	{
		const Layout L = make_layout(n, rows, cols, p_r, p_eps, rpw);
		const int n_tiles = (int)L.tile_inst.size();
		// queues: counts (padded to 64 ints), tags, three queues; step 0's
		// queue holds every tile, whose tags are 0 (queued for step 0)
		std::vector<int> qinit((size_t)64 + 4 * (size_t)n_tiles, 0);
		qinit[0] = n_tiles;
		for (int k = 0; k < n_tiles; ++k) qinit[64 + (size_t)n_tiles + k] = k;
		const size_t bytes = (size_t)total * sizeof(double);
		DevBuf dinst, dti, dq;
		float ms = -1.f;
		hipEvent_t e0 = nullptr, e1 = nullptr;
		do {
			if (CALHIP_CHECK(hipMalloc(&dinst.p, sizeof(BatchInst) * (size_t)n)) ||
			    CALHIP_CHECK(hipMalloc(&dti.p, sizeof(int) * (size_t)n_tiles)) ||
			    CALHIP_CHECK(hipMalloc(&dq.p, sizeof(int) * qinit.size())))
				break;
			int* qb = (int*)dq.p;
			BatchQueues Q;
			Q.count = qb;
			Q.tag = qb + 64;
			for (int k = 0; k < 3; ++k) Q.q[k] = qb + 64 + (size_t)(k + 1) * n_tiles;
			if (CALHIP_CHECK(hipMemcpyAsync(dh0.p, dhi.p, bytes, hipMemcpyDeviceToDevice, stream)) ||
			    CALHIP_CHECK(hipMemcpyAsync(dh1.p, dhi.p, bytes, hipMemcpyDeviceToDevice, stream)) ||
			    CALHIP_CHECK(hipMemcpyAsync(dinst.p, L.inst.data(), sizeof(BatchInst) * (size_t)n,
			                                hipMemcpyHostToDevice, stream)) ||
			    CALHIP_CHECK(hipMemcpyAsync(dti.p, L.tile_inst.data(), sizeof(int) * (size_t)n_tiles,
			                                hipMemcpyHostToDevice, stream)) ||
			    CALHIP_CHECK(hipMemcpyAsync(dq.p, qinit.data(), sizeof(int) * qinit.size(), hipMemcpyHostToDevice,
			                                stream)))
				break;
			if (CALHIP_CHECK(hipEventCreate(&e0)) || CALHIP_CHECK(hipEventCreate(&e1))) break;
			if (launch_all(rpw, depth, steps, per_launch, (const double*)dz.p, (double*)dh0.p, (double*)dh1.p, (double*)dtrash.p,
			               (const BatchInst*)dinst.p, (const int*)dti.p, Q, e0, e1))
				break;
			std::vector<double> hh;
			if (h_out) {
				hh.resize((size_t)total);
				if (CALHIP_CHECK(hipMemcpyAsync(hh.data(), (steps & 1) ? dh1.p : dh0.p, bytes, hipMemcpyDeviceToHost,
				                                stream)))
					break;
			}
			if (CALHIP_CHECK(hipStreamSynchronize(stream))) break;
			float t = 0.f;
			if (CALHIP_CHECK(hipEventElapsedTime(&t, e0, e1))) break;
			if (h_out)
				for (int k = 0; k < n; ++k)
					for (int i = 0; i < rows[k]; ++i)
						memcpy(h_out[k] + (size_t)i * cols[k], hh.data() + L.inst[k].off + (size_t)i * L.inst[k].pitch,
						       sizeof(double) * (size_t)cols[k]);
			ms = t;
		} while (0);
		if (e0) (void)hipEventDestroy(e0);
		if (e1) (void)hipEventDestroy(e1);
		(void)hipStreamSynchronize(stream);
		return ms;
	}

	// The stepping launches between events e0 and e1: `per_launch` steps per
	// cooperative launch (0: all of them), or one ordinary launch per step
	// (per_launch 1, or when the device offers no cooperative launch).
	int launch_all(int rpw, int depth, int steps, int per_launch, const double* z, double* h0, double* h1,
	               double* trash, const BatchInst* dinst, const int* dti, BatchQueues Q, hipEvent_t e0, hipEvent_t e1)
	{
		int dev = 0, cus = 0, coop = 0, per_cu = 0;
		if (CALHIP_CHECK(hipGetDevice(&dev)) ||
		    CALHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) ||
		    CALHIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev)))
			return -1;
		const void* kcoop = depth >= 2 ? (const void*)batch_steps_coop<2> : (const void*)batch_steps_coop<1>;
		if (coop && per_launch != 1 &&
		    CALHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kcoop, BATCH_BLOCK, 0)))
			return -1;
		if (CALHIP_CHECK(hipEventRecord(e0, stream))) return -1;
		if (coop && per_launch != 1 && per_cu > 0) {
			// exactly the blocks that fit at once: the launch is refused
			// rather than started when they do not
			// two blocks per CU: 256 instances x 2000 steps took 772 ms against
			// 935 (one) and 845 (three, the most that fit); one launch per
			// step on the same grid takes the same 771 ms (an idle grid
			// barrier costs 34-78 us here against 2.5-2.8 us per empty
			// launch, but the waves reach it staggered and its cost hides)
			int bpc = std::min(per_cu, 2);
			if (const char* e = getenv("CALHIP_BATCH_BLOCKS_PER_CU"))
				if (atoi(e) > 0 && atoi(e) <= per_cu) bpc = atoi(e);
			const int grid = bpc * cus;
			const int chunk = per_launch > 0 ? per_launch : steps;
			for (int t0 = 0; t0 < steps; t0 += chunk) {
				int t1 = std::min(steps, t0 + chunk);
				void* args[] = {(void*)&dinst, (void*)&dti, (void*)&rpw, (void*)&z, (void*)&h0,
				                (void*)&h1,    (void*)&trash, (void*)&Q, (void*)&t0, (void*)&t1};
				if (CALHIP_CHECK(hipLaunchCooperativeKernel(kcoop, dim3(grid), dim3(BATCH_BLOCK), args, 0, stream)))
					return -1;
			}
		} else {
			int bpc = 2;   // as the cooperative grid (771 ms either way; 3 per CU: 798)
			if (const char* e = getenv("CALHIP_BATCH_BLOCKS_PER_CU"))
				if (atoi(e) > 0 && atoi(e) <= 8) bpc = atoi(e);
			const int grid = bpc * cus;
			for (int t = 0; t < steps; ++t) {
				if (depth >= 2)
					hipLaunchKernelGGL(batch_step<2>, dim3(grid), dim3(BATCH_BLOCK), 0, stream, dinst, dti, rpw, z, h0,
					                   h1, trash, Q, t);
				else
					hipLaunchKernelGGL(batch_step<1>, dim3(grid), dim3(BATCH_BLOCK), 0, stream, dinst, dti, rpw, z, h0,
					                   h1, trash, Q, t);
			}
			if (CALHIP_CHECK(hipGetLastError())) return -1;
		}
		return CALHIP_CHECK(hipEventRecord(e1, stream));
	}
};

// Chunk length per ensemble: the candidates run the ensemble's first
// TUNE_STEPS steps (the two full sweeps and the queue's first steps) and the
// fastest is kept, cached by the ensemble's shape.
constexpr int TUNE_STEPS = 12;
std::mutex tune_mu;
std::map<std::vector<int>, int> tuned_rpw;

}  // namespace

}  // namespace calhip

using namespace calhip;

extern "C" int calhipSciddicaTBatchRun(int n, const int* rows, const int* cols, const double* p_r,
                                       const double* p_eps, const double* const* z_host,
                                       const double* const* h_host, int steps, double* const* h_out,
                                       double* timing_ms)
{
	if (n <= 0 || !rows || !cols || !p_r || !p_eps || !z_host || !h_host || !h_out || steps < 0) {
		set_error(40, "calhipSciddicaTBatchRun: bad arguments");
		return -1;
	}
	for (int k = 0; k < n; ++k)
		if (rows[k] < 3 || cols[k] < 3 || !z_host[k] || !h_host[k] || !h_out[k]) {
			set_error(41, "calhipSciddicaTBatchRun: every instance needs a grid of at least 3 x 3 and its arrays");
			return -1;
		}
	int depth = 1, rpw = 0, per_launch = 0;
	if (const char* e = getenv("CALHIP_BATCH_ROWS_PER_WAVE")) rpw = atoi(e) > 0 ? atoi(e) : 0;
	if (const char* e = getenv("CALHIP_BATCH_PREFETCH")) depth = atoi(e) > 0 ? atoi(e) : depth;
	if (const char* e = getenv("CALHIP_BATCH_STEPS_PER_LAUNCH")) per_launch = atoi(e) > 0 ? atoi(e) : 0;
	Runner R{n, rows, cols, p_r, p_eps};
	if (CALHIP_CHECK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking))) return -1;
	int rc = -1;
	do {
		if (R.prepare(z_host, h_host)) break;
		if (rpw == 0) {
			std::vector<int> key(rows, rows + n);
			key.insert(key.end(), cols, cols + n);
			std::lock_guard<std::mutex> lk(tune_mu);
			auto it = tuned_rpw.find(key);
			if (it != tuned_rpw.end()) {
				rpw = it->second;
			} else {
				static const int cand[] = {8, 16, 32};
				float best = -1.f;
				for (int c : cand) {
					const float ms = R.run(c, depth, std::min(steps, TUNE_STEPS), per_launch, nullptr);
					if (ms < 0.f) break;
					if (best < 0.f || ms < best) {
						best = ms;
						rpw = c;
					}
				}
				if (best < 0.f) break;
				tuned_rpw[key] = rpw;
			}
		}
		const float ms = R.run(rpw, depth, steps, per_launch, h_out);
		if (ms < 0.f) break;
		if (timing_ms) *timing_ms = ms;
		rc = 0;
	} while (0);
	(void)hipStreamSynchronize(R.stream);
	(void)hipStreamDestroy(R.stream);
	return rc;
}
