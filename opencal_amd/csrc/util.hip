// Whole-substate device operations: constant fills (full grid or active
// cells), copies, and the active-cell set rebuild.
//
// Active-set rebuild (OpenCAL's calUpdateActiveCells2D, cal2D.c:382-440):
// flags -> row-major list of linear indices + count, in two launches:
//   1. per-chunk counts of set flags (chunk = 4096 flags per 256-thread block,
//      16 flags per thread read as one 16-byte load);
//   2. each block sums the counts of the chunks before it, scans its threads'
//      counts in LDS and writes its indices in row-major order; the last
//      block also writes the total.
#include "engine.hpp"
#include "device_common.hpp"

namespace calhip {

namespace {

constexpr int FILL_BLOCK = 256;
constexpr int RB_THREADS = 256;
constexpr int RB_PER_THREAD = 16;
constexpr int RB_CHUNK = RB_THREADS * RB_PER_THREAD;

template <typename T>
__global__ void fill_kernel(T* __restrict__ buf, T v, size_t n, const int* __restrict__ cells,
                            const int* __restrict__ count)
{
	size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t)gridDim.x * blockDim.x;
	int cnt = count ? count[0] : 0;
	if (cnt > 0) {
		for (size_t k = tid; k < (size_t)cnt; k += stride) buf[cells[k]] = v;
	} else {
		for (size_t c = tid; c < n; c += stride) buf[c] = v;
	}
}

// OpenCAL's update of one substate in active mode (calUpdateSubstate2D*,
// cal2D.c:694-713): next -> current at the listed cells, or over the whole
// grid when the list is empty (A.cells == NULL).
template <typename T>
__global__ void copy_cells_kernel(T* __restrict__ dst, const T* __restrict__ src, size_t n,
                                  const int* __restrict__ cells, const int* __restrict__ count)
{
	size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t)gridDim.x * blockDim.x;
	int cnt = count[0];
	if (cnt > 0) {
		for (size_t k = tid; k < (size_t)cnt; k += stride) dst[cells[k]] = src[cells[k]];
	} else {
		for (size_t c = tid; c < n; c += stride) dst[c] = src[c];
	}
}

__device__ __forceinline__ int count_bytes16(const uint8_t* p, size_t base, size_t n)
{
	int c = 0;
	if (base + RB_PER_THREAD <= n) {
		uint4 v = *reinterpret_cast<const uint4*>(p + base);
		unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
		for (int q = 0; q < 4; ++q)
#pragma unroll
			for (int b = 0; b < 4; ++b) c += ((w[q] >> (8 * b)) & 0xffu) ? 1 : 0;
	} else {
		for (size_t k = base; k < n; ++k) c += p[k] ? 1 : 0;
	}
	return c;
}

__global__ __launch_bounds__(RB_THREADS) void rebuild_count(const uint8_t* __restrict__ flags, size_t n,
                                                            int* __restrict__ block_counts)
{
	__shared__ int red[RB_THREADS / 64];
	size_t base = (size_t)blockIdx.x * RB_CHUNK + (size_t)threadIdx.x * RB_PER_THREAD;
	int c = base < n ? count_bytes16(flags, base, n) : 0;
	for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0) {
		int s = 0;
		for (int w = 0; w < RB_THREADS / 64; ++w) s += red[w];
		block_counts[blockIdx.x] = s;
	}
}

__global__ __launch_bounds__(RB_THREADS) void rebuild_scatter(const uint8_t* __restrict__ flags, size_t n,
                                                              const int* __restrict__ block_counts,
                                                              int nblocks, int* __restrict__ cells,
                                                              int* __restrict__ count)
{
	__shared__ int scan[RB_THREADS];
	__shared__ int wsum[RB_THREADS / 64];
	// offset of this chunk: sum of the counts of all earlier chunks
	int part = 0;
	for (int b = threadIdx.x; b < (int)blockIdx.x; b += RB_THREADS) part += block_counts[b];
	for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
	if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = part;
	__syncthreads();
	int block_off = 0;
	for (int w = 0; w < RB_THREADS / 64; ++w) block_off += wsum[w];
	__syncthreads();

	size_t base = (size_t)blockIdx.x * RB_CHUNK + (size_t)threadIdx.x * RB_PER_THREAD;
	int mine = base < n ? count_bytes16(flags, base, n) : 0;
	// inclusive scan of per-thread counts (Hillis-Steele in LDS)
	scan[threadIdx.x] = mine;
	__syncthreads();
	for (int off = 1; off < RB_THREADS; off <<= 1) {
		int v = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
		__syncthreads();
		scan[threadIdx.x] += v;
		__syncthreads();
	}
	int pos = block_off + scan[threadIdx.x] - mine;
	size_t end = base + RB_PER_THREAD < n ? base + RB_PER_THREAD : n;
	for (size_t k = base; k < end; ++k)
		if (flags[k]) cells[pos++] = (int)k;
	if (blockIdx.x == (unsigned)nblocks - 1 && threadIdx.x == RB_THREADS - 1) count[0] = block_off + scan[threadIdx.x];
}

}  // namespace

namespace k {

static int grid_for(size_t n)
{
	size_t b = (n + FILL_BLOCK - 1) / FILL_BLOCK;
	if (b > 8192) b = 8192;
	if (b < 1) b = 1;
	return (int)b;
}

int fill(ModelX* m, void* buf, int elem, double v, bool active_only)
{
	size_t n = m->cells();
	const int* cells = active_only ? m->d_cells : nullptr;
	const int* count = active_only ? m->d_count : nullptr;
	int g = grid_for(n);
	switch (elem) {
	case 1:
		hipLaunchKernelGGL(fill_kernel<int8_t>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (int8_t*)buf,
		                   (int8_t)v, n, cells, count);
		break;
	case 4:
		hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (int32_t*)buf,
		                   (int32_t)v, n, cells, count);
		break;
	default:
		hipLaunchKernelGGL(fill_kernel<double>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (double*)buf, v, n,
		                   cells, count);
		break;
	}
	m->stats.launches++;
	return CALHIP_CHECK(hipGetLastError());
}

int copy(ModelX* m, void* dst, const void* src, size_t bytes)
{
	m->stats.launches++;
	return CALHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, m->stream));
}

int copy_active(ModelX* m, void* dst, const void* src, int elem)
{
	size_t n = m->cells();
	int g = grid_for(n);
	switch (elem) {
	case 1:
		hipLaunchKernelGGL(copy_cells_kernel<int8_t>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (int8_t*)dst,
		                   (const int8_t*)src, n, m->d_cells, m->d_count);
		break;
	case 4:
		hipLaunchKernelGGL(copy_cells_kernel<int32_t>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (int32_t*)dst,
		                   (const int32_t*)src, n, m->d_cells, m->d_count);
		break;
	default:
		hipLaunchKernelGGL(copy_cells_kernel<double>, dim3(g), dim3(FILL_BLOCK), 0, m->stream, (double*)dst,
		                   (const double*)src, n, m->d_cells, m->d_count);
		break;
	}
	m->stats.launches++;
	return CALHIP_CHECK(hipGetLastError());
}

int rebuild_active(ModelX* m)
{
	size_t n = m->cells();
	int nb = (int)((n + RB_CHUNK - 1) / RB_CHUNK);
	if (m->n_count_blocks < nb) {
		if (m->d_block_counts) hipFree(m->d_block_counts);
		m->d_block_counts = nullptr;
		if (CALHIP_CHECK(hipMalloc(&m->d_block_counts, sizeof(int) * (size_t)nb))) return -1;
		m->n_count_blocks = nb;
	}
	hipLaunchKernelGGL(rebuild_count, dim3(nb), dim3(RB_THREADS), 0, m->stream, m->d_flags, n,
	                   m->d_block_counts);
	hipLaunchKernelGGL(rebuild_scatter, dim3(nb), dim3(RB_THREADS), 0, m->stream, m->d_flags, n,
	                   m->d_block_counts, nb, m->d_cells, m->d_count);
	m->stats.launches += 2;
	return CALHIP_CHECK(hipGetLastError());
}

int active_count_host(ModelX* m, int* out)
{
	if (CALHIP_CHECK(hipMemcpyAsync(out, m->d_count, sizeof(int), hipMemcpyDeviceToHost, m->stream))) return -1;
	return CALHIP_CHECK(hipStreamSynchronize(m->stream));
}

}  // namespace k
}  // namespace calhip
