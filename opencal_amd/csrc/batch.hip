// Batched SciddicaT ensemble: many independent full-grid SciddicaT instances
// of uneven sizes and parameters stepped together.
//
// Every instance keeps its own z and double-buffered h in one packed device
// allocation (256-byte aligned offsets).  The work of one step over all
// instances is a single flat list of tiles (instance, group of 4 adjacent
// 124-column strips, chunk of 8 rows); one launch per step covers the whole
// ensemble, so the launch cost is paid once per step for all instances and
// the GPU is filled even when each instance alone is far too small to fill
// it.  The sweep of a tile is the fused two-process sweep of the single-model
// program, so each instance's result equals its own plain run bit for bit.
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_device.hpp"

namespace calhip {

namespace {

using namespace sct;

constexpr int BATCH_RPW = 8;

struct BatchInst {
	int rows, cols, strips, pad;
	long long off;   // element offset of the instance's grids
	double r, eps;
};

__global__ __launch_bounds__(64 * FUSED_WAVES) void batch_step(const BatchInst* __restrict__ inst,
                                                                 const int4* __restrict__ items, int n_items,
                                                                 int per_xcd, const double* __restrict__ z,
                                                                 const double* __restrict__ h,
                                                                 double* __restrict__ hn, double* __restrict__ trash)
{
	const int L = dev::xcd_logical_block(blockIdx.x, per_xcd);
	if (L >= n_items) return;
	const int4 it = items[L];
	const BatchInst I = inst[it.x];
	const int strip = it.y * FUSED_WAVES + (threadIdx.x >> 6);
	if (strip >= I.strips) return;
	const int r0 = it.z * BATCH_RPW;
	const int r1 = min(r0 + BATCH_RPW, I.rows);
	if (I.cols % 2 == 0)
		sweep_chunk<1, true>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, strip, it.z, r0, r1, I.r, I.eps);
	else
		sweep_chunk<1, false>(z + I.off, h + I.off, hn + I.off, trash, I.rows, I.cols, strip, it.z, r0, r1, I.r,
		                      I.eps);
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
	long long total = 0;
	for (int k = 0; k < n; ++k) {
		if (rows[k] < 3 || cols[k] < 3 || !z_host[k] || !h_host[k] || !h_out[k]) {
			set_error(41, "calhipSciddicaTBatchRun: every instance needs a grid of at least 3 x 3 and its arrays");
			return -1;
		}
		BatchInst& I = inst[k];
		I.rows = rows[k];
		I.cols = cols[k];
		I.strips = (cols[k] + STRIP2_OUT - 1) / STRIP2_OUT;
		I.pad = 0;
		I.off = total;
		I.r = p_r[k];
		I.eps = p_eps[k];
		total += ((long long)rows[k] * cols[k] + 31) / 32 * 32;   // 256-byte aligned starts
		int sgroups = (I.strips + FUSED_WAVES - 1) / FUSED_WAVES;
		int chunks = (rows[k] + BATCH_RPW - 1) / BATCH_RPW;
		for (int c = 0; c < chunks; ++c)
			for (int g = 0; g < sgroups; ++g) items.push_back(make_int4(k, g, c, 0));
	}
	int n_items = (int)items.size();
	int per_xcd = (n_items + 7) / 8;

	// host staging: z after the reference init (z -= h where h > 0), h
	std::vector<double> hz((size_t)total, 0.0), hh((size_t)total, 0.0);
	for (int k = 0; k < n; ++k) {
		size_t cells = (size_t)rows[k] * cols[k];
		double* zz = hz.data() + inst[k].off;
		double* h0 = hh.data() + inst[k].off;
		for (size_t c = 0; c < cells; ++c) {
			double hv = h_host[k][c];
			h0[c] = hv;
			zz[c] = hv > 0.0 ? z_host[k][c] - hv : z_host[k][c];
		}
	}

	hipStream_t stream = nullptr;
	if (CALHIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking))) return -1;
	DevBuf dz, dh0, dh1, dinst, ditems, dtrash;
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
		for (int s = 0; s < steps; ++s)
			hipLaunchKernelGGL(batch_step, dim3(per_xcd * 8), dim3(64 * FUSED_WAVES), 0, stream,
			                   (const BatchInst*)dinst.p, (const int4*)ditems.p, n_items, per_xcd,
			                   (const double*)dz.p, (const double*)bufs[s & 1], bufs[(s + 1) & 1],
			                   (double*)dtrash.p);
		if (CALHIP_CHECK(hipGetLastError())) break;
		if (CALHIP_CHECK(hipEventRecord(e1, stream))) break;
		if (CALHIP_CHECK(hipMemcpyAsync(hh.data(), bufs[steps & 1], bytes, hipMemcpyDeviceToHost, stream))) break;
		if (CALHIP_CHECK(hipStreamSynchronize(stream))) break;
		float ms = 0.f;
		if (CALHIP_CHECK(hipEventElapsedTime(&ms, e0, e1))) break;
		if (timing_ms) *timing_ms = ms;
		for (int k = 0; k < n; ++k)
			memcpy(h_out[k], hh.data() + inst[k].off, sizeof(double) * (size_t)rows[k] * cols[k]);
		rc = 0;
	} while (0);
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipStreamSynchronize(stream);
	(void)hipStreamDestroy(stream);
	return rc;
}
