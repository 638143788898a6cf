// Batched SciddicaT ensemble: many independent full-grid SciddicaT instances
// of uneven sizes and parameters stepped together.
//
// Every instance keeps its own z and double-buffered h in one packed device
// allocation.  Rows of an instance with an even column count are padded to a
// pitch of 16 columns, so every row starts on a 128-byte line and the aligned
// 128-column strip sweep applies; odd column counts keep packed rows and the
// 124-column strip sweep.  The work of one step over all instances is a
// single flat list of wave items (instance, strip, chunk of rows), run four
// to a block; one launch per step covers the whole ensemble, so the launch cost is
// paid once per step for all instances and the GPU is filled even when each
// instance alone is far too small to fill it.  The sweep of a tile is the
// fused two-process sweep of the single-model program, so each instance's
// result equals its own plain run bit for bit.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_device.hpp"

namespace calhip {

namespace {

using namespace sct;

// rows per chunk: with the whole ensemble in one launch the GPU is full and
// the sweep is bound by throughput, so long chunks (fewer re-read halo rows)
// win: 32 rows 154 GCUPS against 16: 139, 48: 150, 64: 143, 4 (prefetch 2): 67
constexpr int BATCH_RPW = 32;
constexpr int BATCH_DEPTH = 1;  // rows prefetched ahead by the aligned sweep

struct BatchInst {
	int rows, cols, strips, pitch;   // pitch: row stride in elements
	long long off;                   // element offset of the instance's grids
	double r, eps;
};

// One step of every instance.  Wave items (instance, strip, chunk) form one
// flat list; a block runs FUSED_WAVES consecutive items (so no wave idles at
// an instance's last strip group), and blocks run in wet-first order (ws)
// when given.
template <int D>
__global__ __launch_bounds__(64 * FUSED_WAVES) __attribute__((amdgpu_waves_per_eu(D == 1 ? 3 : 2))) void batch_step(const BatchInst* __restrict__ inst,
                                                                 const int4* __restrict__ items, int n_items,
                                                                 int n_blocks, int per_xcd, int rpw,
                                                                 const double* __restrict__ z,
                                                                 const double* __restrict__ h,
                                                                 double* __restrict__ hn, double* __restrict__ trash,
                                                                 WetSets ws)
{
	const int L = wet_first_tile(ws, n_blocks, per_xcd);
	if (L < 0) return;
	const int k = __builtin_amdgcn_readfirstlane(L * FUSED_WAVES + (int)(threadIdx.x >> 6));
	bool wet = false;
	if (k < n_items) {
		const int4 it = items[k];
		const BatchInst I = inst[it.x];
		const int strip = it.y, r0 = it.z * rpw;
		const int r1 = min(r0 + rpw, I.rows);
		if (I.cols % 2 == 0)
			wet = sweep_chunk_al<D>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, I.pitch, strip, it.z,
			                        r0, r1, I.r, I.eps);
		else   // odd column counts: the tile keeps its place in grid order
			sweep_chunk<1, false>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, strip, it.z, r0, r1,
			                      I.r, I.eps);
	}
	wet_first_record(ws, L, wet);
}

struct DevBuf {
	void* p = nullptr;
	~DevBuf()
	{
		if (p) (void)hipFree(p);
	}
};

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
	std::vector<BatchInst> inst(n);
	std::vector<int4> items;
	int rpw = BATCH_RPW, depth = BATCH_DEPTH;
	if (const char* e = getenv("CALHIP_BATCH_ROWS_PER_WAVE")) rpw = atoi(e) > 0 ? atoi(e) : rpw;
	if (const char* e = getenv("CALHIP_BATCH_PREFETCH")) depth = atoi(e) > 0 ? atoi(e) : depth;
	long long total = 0;
	for (int k = 0; k < n; ++k) {
		if (rows[k] < 3 || cols[k] < 3 || !z_host[k] || !h_host[k] || !h_out[k]) {
			set_error(41, "calhipSciddicaTBatchRun: every instance needs a grid of at least 3 x 3 and its arrays");
			return -1;
		}
		BatchInst& I = inst[k];
		const bool even = cols[k] % 2 == 0;
		I.rows = rows[k];
		I.cols = cols[k];
		I.strips = even ? (cols[k] + STRIP_AL - 1) / STRIP_AL : (cols[k] + STRIP2_OUT - 1) / STRIP2_OUT;
		I.pitch = even ? (cols[k] + 15) / 16 * 16 : cols[k];
		I.off = total;
		I.r = p_r[k];
		I.eps = p_eps[k];
		total += ((long long)rows[k] * I.pitch + 31) / 32 * 32;   // 256-byte aligned starts
		const int chunks = (rows[k] + rpw - 1) / rpw;
		for (int c = 0; c < chunks; ++c)
			for (int st = 0; st < I.strips; ++st) items.push_back(make_int4(k, st, c, 0));
	}
	const int n_items = (int)items.size();
	const int n_blocks = (n_items + FUSED_WAVES - 1) / FUSED_WAVES;
	const int per_xcd = (n_blocks + 7) / 8;

	// host staging: z after the reference init (z -= h where h > 0), h
	std::vector<double> hz((size_t)total, 0.0), hh((size_t)total, 0.0);
	for (int k = 0; k < n; ++k) {
		for (int i = 0; i < rows[k]; ++i) {
			const double* zs = z_host[k] + (size_t)i * cols[k];
			const double* hs = h_host[k] + (size_t)i * cols[k];
			double* zz = hz.data() + inst[k].off + (size_t)i * inst[k].pitch;
			double* h0 = hh.data() + inst[k].off + (size_t)i * inst[k].pitch;
			for (int j = 0; j < cols[k]; ++j) {
				const double hv = hs[j];
				h0[j] = hv;
				zz[j] = hv > 0.0 ? zs[j] - hv : zs[j];
			}
		}
	}

	hipStream_t stream = nullptr;
	if (CALHIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking))) return -1;
	DevBuf dz, dh0, dh1, dinst, ditems, dtrash;
	WetOrder wo;   // wet-first tile order (freed after the stream is synchronised)
	size_t bytes = (size_t)total * sizeof(double);
	int rc = -1;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	do {
		if (CALHIP_CHECK(hipMalloc(&dz.p, bytes)) || CALHIP_CHECK(hipMalloc(&dh0.p, bytes)) ||
		    CALHIP_CHECK(hipMalloc(&dh1.p, bytes)) ||
		    CALHIP_CHECK(hipMalloc(&dinst.p, sizeof(BatchInst) * (size_t)n)) ||
		    CALHIP_CHECK(hipMalloc(&ditems.p, sizeof(int4) * (size_t)n_items)) ||
		    CALHIP_CHECK(hipMalloc(&dtrash.p, 128 * sizeof(double))))
			break;
		if (CALHIP_CHECK(hipMemcpyAsync(dz.p, hz.data(), bytes, hipMemcpyHostToDevice, stream)) ||
		    CALHIP_CHECK(hipMemcpyAsync(dh0.p, hh.data(), bytes, hipMemcpyHostToDevice, stream)) ||
		    CALHIP_CHECK(hipMemcpyAsync(dh1.p, hh.data(), bytes, hipMemcpyHostToDevice, stream)) ||
		    CALHIP_CHECK(hipMemcpyAsync(dinst.p, inst.data(), sizeof(BatchInst) * (size_t)n, hipMemcpyHostToDevice,
		                                stream)) ||
		    CALHIP_CHECK(hipMemcpyAsync(ditems.p, items.data(), sizeof(int4) * (size_t)n_items,
		                                hipMemcpyHostToDevice, stream)))
			break;
		if (CALHIP_CHECK(hipEventCreate(&e0)) || CALHIP_CHECK(hipEventCreate(&e1))) break;
		if (CALHIP_CHECK(hipEventRecord(e0, stream))) break;
		double* bufs[2] = {(double*)dh0.p, (double*)dh1.p};
		for (int s = 0; s < steps; ++s) {
			WetSets ws = {};
			int blocks = per_xcd * 8;
			if (wo.next(stream, n_blocks, ws, 512) == 0)
				blocks = 8 * (per_xcd + ws.cap);
			else
				ws = WetSets{};
			if (depth >= 2)
				hipLaunchKernelGGL((batch_step<2>), dim3(blocks), dim3(64 * FUSED_WAVES), 0, stream,
				                   (const BatchInst*)dinst.p, (const int4*)ditems.p, n_items, n_blocks, per_xcd, rpw,
				                   (const double*)dz.p, (const double*)bufs[s & 1], bufs[(s + 1) & 1],
				                   (double*)dtrash.p, ws);
			else
				hipLaunchKernelGGL((batch_step<1>), dim3(blocks), dim3(64 * FUSED_WAVES), 0, stream,
				                   (const BatchInst*)dinst.p, (const int4*)ditems.p, n_items, n_blocks, per_xcd, rpw,
				                   (const double*)dz.p, (const double*)bufs[s & 1], bufs[(s + 1) & 1],
				                   (double*)dtrash.p, ws);
		}
		if (CALHIP_CHECK(hipGetLastError())) break;
		if (CALHIP_CHECK(hipEventRecord(e1, stream))) break;
		if (CALHIP_CHECK(hipMemcpyAsync(hh.data(), bufs[steps & 1], bytes, hipMemcpyDeviceToHost, stream))) break;
		if (CALHIP_CHECK(hipStreamSynchronize(stream))) break;
		float ms = 0.f;
		if (CALHIP_CHECK(hipEventElapsedTime(&ms, e0, e1))) break;
		if (timing_ms) *timing_ms = ms;
		for (int k = 0; k < n; ++k)
			for (int i = 0; i < rows[k]; ++i)
				memcpy(h_out[k] + (size_t)i * cols[k], hh.data() + inst[k].off + (size_t)i * inst[k].pitch,
				       sizeof(double) * (size_t)cols[k]);
		rc = 0;
	} while (0);
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipStreamSynchronize(stream);
	(void)hipStreamDestroy(stream);
	return rc;
}
