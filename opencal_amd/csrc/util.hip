// Whole-substate device operations: constant fills (full grid or active
// cells), copies, and the active-cell set rebuild.
//
// Active-set rebuild (OpenCAL's calUpdateActiveCells2D, cal2D.c:382-440):
// flags -> row-major list of linear indices + count, in one launch.  A block
// takes the next chunk of 4096 flags (its index from an atomic counter, so
// the chunks before it belong to blocks that have started), reads 16 flags
// per thread as one 16-byte load, ranks its set flags with one wavefront
// ballot per flag position and the waves' totals in LDS, and finds the
// number of set flags before its chunk by a decoupled look-back (Merrill
// and Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// 2016): each chunk publishes its own count and then its inclusive prefix in
// a status word; a wave reads the 64 chunks before it at a time and adds
// counts back to the first inclusive prefix.  The chunk that is last in the
// grid also writes the total.
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

// status word of a chunk: rebuild epoch (bits 63..32), kind (31..30: 1 =
// the chunk's own count, 2 = the count up to and including the chunk),
// value (29..0)
constexpr unsigned long long ST_AGG = 1ull << 30, ST_INC = 2ull << 30;
constexpr unsigned long long ST_VAL = (1ull << 30) - 1ull;
// re-reads of a look-back window before a block counts the flags before its
// chunk itself (never needed while the blocks before it make progress; it
// bounds the wait whatever happens)
constexpr int RB_SPIN_MAX = 1 << 16;

__device__ __forceinline__ unsigned long long st_load(const unsigned long long* p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_store(unsigned long long* p, unsigned long long v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// set flags in [0, end), counted by one wave (the look-back's last resort)
__device__ int count_before(const uint8_t* flags, size_t end)
{
	const int lane = threadIdx.x & 63;
	int c = 0;
	for (size_t k = (size_t)lane; k < end; k += 64) c += flags[k] ? 1 : 0;
	for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
	return __shfl(c, 0, 64);
}

__global__ __launch_bounds__(RB_THREADS) void rebuild_onepass(const uint8_t* __restrict__ flags, size_t n,
                                                              unsigned long long* status, int nb,
                                                              unsigned epoch, int* __restrict__ cells,
                                                              int* __restrict__ count)
{
	__shared__ int s_chunk, s_prefix;
	__shared__ int wtot[RB_THREADS / 64];
	const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
	// chunks in start order: the counter after the status words advances by
	// nb per rebuild
	if (threadIdx.x == 0) s_chunk = (int)(atomicAdd(status + nb, 1ull) % (unsigned long long)nb);
	__syncthreads();
	const int b = s_chunk;
	const size_t base = (size_t)b * RB_CHUNK + (size_t)threadIdx.x * RB_PER_THREAD;
	unsigned bits = 0u;   // bit q: flag base + q is set
	if (base + RB_PER_THREAD <= n) {
		const uint4 v = *reinterpret_cast<const uint4*>(flags + base);
		const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
		for (int q = 0; q < 16; ++q) bits |= ((w[q >> 2] >> (8 * (q & 3))) & 0xffu) ? 1u << q : 0u;
	} else {
		for (size_t k = base; k < n; ++k) bits |= flags[k] ? 1u << (k - base) : 0u;
	}
	// set flags of the lanes below this one: per flag position, a ballot
	const unsigned long long below = (1ull << lane) - 1ull;
	int before = 0;
#pragma unroll
	for (int q = 0; q < 16; ++q) before += __popcll(__ballot((bits >> q) & 1u) & below);
	const int mine = __popc(bits);
	if (lane == 63) wtot[wid] = before + mine;
	__syncthreads();
	int woff = 0, agg = 0;
#pragma unroll
	for (int w = 0; w < RB_THREADS / 64; ++w) {
		woff += w < wid ? wtot[w] : 0;
		agg += wtot[w];
	}
	if (wid == 0) {
		const unsigned long long ep = (unsigned long long)epoch << 32;
		int prefix = 0;
		if (b == 0) {
			if (lane == 0) st_store(status, ep | ST_INC | (unsigned long long)agg);
		} else {
			if (lane == 0) st_store(status + b, ep | ST_AGG | (unsigned long long)agg);
			int p = b - 1;   // newest chunk of the window
			int spins = 0;
			while (true) {
				const int q = p - lane;
				const unsigned long long s = q >= 0 ? st_load(status + q) : (ep | ST_INC);
				const bool ready = (s >> 32) == (unsigned long long)epoch;
				const bool inc = ready && (s & (3ull << 30)) == ST_INC;
				const int val = (ready && q >= 0) ? (int)(s & ST_VAL) : 0;
				const unsigned long long rmask = __ballot(ready), imask = __ballot(inc);
				int upto = -1;   // the window's lanes 0..upto are summed
				if (imask) {
					const int first = __ffsll((long long)imask) - 1;
					const unsigned long long need = first == 63 ? ~0ull : (2ull << first) - 1ull;
					if ((rmask & need) == need) upto = first;
				} else if (rmask == ~0ull) {
					upto = 63;
				}
				if (upto >= 0) {
					int sum = lane <= upto ? val : 0;
					for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
					prefix += __shfl(sum, 0, 64);
					if (upto < 63 || imask) break;
					p -= 64;
					spins = 0;
					continue;
				}
				if (++spins >= RB_SPIN_MAX) {
					// count the flags before the window's remaining chunks directly
					prefix += count_before(flags, (size_t)(p + 1) * RB_CHUNK);
					break;
				}
			}
			if (lane == 0) st_store(status + b, ep | ST_INC | (unsigned long long)(prefix + agg));
		}
		if (lane == 0) {
			s_prefix = prefix;
			if (b == nb - 1) count[0] = prefix + agg;
		}
	}
	__syncthreads();
	int pos = s_prefix + woff + before;
	for (int q = 0; q < 16; ++q)
		if ((bits >> q) & 1u) cells[pos++] = (int)(base + (size_t)q);
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
	if (n < ((size_t)1 << 30)) {
		// one launch: status words (nb) and the chunk counter, zeroed once
		if (m->n_count_blocks != nb || !m->d_block_counts) {
			if (m->d_block_counts) hipFree(m->d_block_counts);
			m->d_block_counts = nullptr;
			m->n_count_blocks = 0;
			const size_t bytes = sizeof(unsigned long long) * ((size_t)nb + 1);
			if (CALHIP_CHECK(hipMalloc(&m->d_block_counts, bytes))) return -1;
			if (CALHIP_CHECK(hipMemsetAsync(m->d_block_counts, 0, bytes, m->stream))) return -1;
			m->n_count_blocks = nb;
			m->rebuild_epoch = 0;
		}
		const unsigned epoch = ++m->rebuild_epoch;
		hipLaunchKernelGGL(rebuild_onepass, dim3(nb), dim3(RB_THREADS), 0, m->stream, m->d_flags, n,
		                   reinterpret_cast<unsigned long long*>(m->d_block_counts), nb, epoch, m->d_cells,
		                   m->d_count);
		m->stats.launches++;
		return CALHIP_CHECK(hipGetLastError());
	}
	// 2^30 cells and more (beyond the status word's count field): per-chunk
	// counts, then each chunk sums those before it
	if (m->n_count_blocks != -nb || !m->d_block_counts) {
		if (m->d_block_counts) hipFree(m->d_block_counts);
		m->d_block_counts = nullptr;
		m->n_count_blocks = 0;
		if (CALHIP_CHECK(hipMalloc(&m->d_block_counts, sizeof(int) * (size_t)nb))) return -1;
		m->n_count_blocks = -nb;
	}
	hipLaunchKernelGGL(rebuild_count, dim3(nb), dim3(RB_THREADS), 0, m->stream, m->d_flags, n,
	                   reinterpret_cast<int*>(m->d_block_counts));
	hipLaunchKernelGGL(rebuild_scatter, dim3(nb), dim3(RB_THREADS), 0, m->stream, m->d_flags, n,
	                   reinterpret_cast<const int*>(m->d_block_counts), nb, m->d_cells, m->d_count);
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
