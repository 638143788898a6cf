// Global reductions of a substate on the GPU: include/cal2DReduction.h and
// include/cal3DReduction.h (OpenCAL-OMP/source/cal2DReduction.c:18-641,
// cal3DReduction.c:18-642).
//
// The reference folds `op` over current[0..n) in index order from cell 0
// (its serial build; its chunk bounds end at `rows` and its OpenMP partials
// each start from cell 0, which the fold here does not repeat).  The GPU
// computes the same result in a fixed two-level tree: 256-thread blocks over
// a grid-stride loop, then one block over the block partials.
//  * Integer and byte SUM/PROD are sums/products modulo 2^64 truncated to
//    the type, i.e. the wrapped serial fold; the bitwise and logical folds
//    are order-free; MAX/MIN keep (value, index) and prefer the smaller index
//    among equal values, which is the cell the serial `acc < v` test keeps
//    (for reals NaN cells are never taken, and a NaN in cell 0 is the result).
//  * Real SUM is Neumaier-compensated at every combine (sum, correction),
//    real PROD a plain product tree; both are deterministic for a given n.
//  * Real BINARY_* convert to CALint as x86's truncating conversion does
//    (INT_MIN for NaN and out-of-range values).
// Bytes read per call: n * sizeof(T), plus the copy of a host array when
// the model is not on the GPU.
#include <math.h>
#include <string.h>

#include "engine.hpp"

extern "C" {
#include "cal2DReduction.h"
#include "cal3DReduction.h"
}

namespace calhip {
namespace {

constexpr int R_THREADS = 256;
constexpr int R_MAX_BLOCKS = 1024;

struct RAcc {
	double a;              // real value / compensated sum
	double b;              // compensation / "holds a value" flag
	long long i;           // integer value or index
	unsigned long long u;  // wrapped integer arithmetic, flags, bits
};

__host__ __device__ inline int to_cint(double v)
{
	return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000;
}

template <typename T>
struct IsReal {
	static constexpr bool value = false;
};
template <>
struct IsReal<double> {
	static constexpr bool value = true;
};

template <typename T>
__device__ inline long long ival(T v)
{
	if (sizeof(T) == 1) return (long long)(signed char)v;
	return (long long)v;
}

template <typename T>
__device__ inline RAcc r_ident(int op)
{
	RAcc r;
	r.a = 0.0;
	r.b = 0.0;
	r.i = 0;
	r.u = 0;
	if (IsReal<T>::value) {
		if (op == REDUCTION_PROD) r.a = 1.0;
	} else {
		if (op == REDUCTION_MAX) r.i = (long long)0x8000000000000000ull;
		if (op == REDUCTION_MIN) r.i = 0x7fffffffffffffffll;
	}
	if (op == REDUCTION_PROD) r.u = 1;
	if (op == REDUCTION_LOGICAL_AND) r.u = 1;
	if (op == REDUCTION_BINARY_AND) r.u = ~0ull;
	return r;
}

template <typename T>
__device__ inline RAcc r_load(int op, T v, long long idx)
{
	RAcc r = r_ident<T>(op);
	if (IsReal<T>::value) {
		const double x = (double)v;
		switch (op) {
		case REDUCTION_MAX:
		case REDUCTION_MIN:
			r.a = x;
			r.i = idx;
			r.b = isnan(x) ? 0.0 : 1.0;
			break;
		case REDUCTION_SUM: r.a = x; break;
		case REDUCTION_PROD: r.a = x; break;
		case REDUCTION_LOGICAL_AND:
		case REDUCTION_LOGICAL_OR:
		case REDUCTION_LOGICAL_XOR: r.u = (x != 0.0) ? 1u : 0u; break;
		default: r.u = (unsigned long long)(long long)to_cint(x); break;
		}
	} else {
		const long long x = ival<T>(v);
		switch (op) {
		case REDUCTION_MAX:
		case REDUCTION_MIN: r.i = x; break;
		case REDUCTION_LOGICAL_AND:
		case REDUCTION_LOGICAL_OR:
		case REDUCTION_LOGICAL_XOR: r.u = x != 0 ? 1u : 0u; break;
		default: r.u = (unsigned long long)x; break;
		}
	}
	return r;
}

// (s, c) + (t, d) with the Neumaier correction of s + t.
__device__ inline RAcc sum2(const RAcc& x, const RAcc& y)
{
	RAcc r = x;
	const double s = x.a + y.a;
	const double e = fabs(x.a) >= fabs(y.a) ? (x.a - s) + y.a : (y.a - s) + x.a;
	r.a = s;
	r.b = x.b + y.b + e;
	return r;
}

template <typename T>
__device__ inline RAcc r_comb(int op, const RAcc& x, const RAcc& y)
{
	RAcc r = x;
	switch (op) {
	case REDUCTION_MAX:
	case REDUCTION_MIN:
		if (IsReal<T>::value) {
			if (x.b == 0.0) return y;
			if (y.b == 0.0) return x;
			const bool y_better = op == REDUCTION_MAX ? y.a > x.a : y.a < x.a;
			const bool x_better = op == REDUCTION_MAX ? x.a > y.a : x.a < y.a;
			if (y_better) return y;
			if (x_better) return x;
			return y.i < x.i ? y : x;
		}
		r.i = op == REDUCTION_MAX ? (x.i > y.i ? x.i : y.i) : (x.i < y.i ? x.i : y.i);
		return r;
	case REDUCTION_SUM:
		if (IsReal<T>::value) return sum2(x, y);
		r.u = x.u + y.u;
		return r;
	case REDUCTION_PROD:
		if (IsReal<T>::value)
			r.a = x.a * y.a;
		else
			r.u = x.u * y.u;
		return r;
	case REDUCTION_LOGICAL_AND:
	case REDUCTION_BINARY_AND: r.u = x.u & y.u; return r;
	case REDUCTION_LOGICAL_OR:
	case REDUCTION_BINARY_OR: r.u = x.u | y.u; return r;
	default: r.u = x.u ^ y.u; return r;
	}
}

template <typename T>
__device__ inline RAcc block_reduce(int op, RAcc acc)
{
	__shared__ RAcc sh[R_THREADS];
	sh[threadIdx.x] = acc;
	__syncthreads();
	for (int s = R_THREADS / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s) sh[threadIdx.x] = r_comb<T>(op, sh[threadIdx.x], sh[threadIdx.x + s]);
		__syncthreads();
	}
	return sh[0];
}

template <typename T>
__global__ void __launch_bounds__(R_THREADS) reduce_blocks(const T* __restrict__ v, long long n, int op, RAcc* part)
{
	RAcc acc = r_ident<T>(op);
	const long long stride = (long long)gridDim.x * R_THREADS;
	for (long long t = (long long)blockIdx.x * R_THREADS + threadIdx.x; t < n; t += stride)
		acc = r_comb<T>(op, acc, r_load<T>(op, v[t], t));
	acc = block_reduce<T>(op, acc);
	if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
__global__ void __launch_bounds__(R_THREADS) reduce_final(const RAcc* part, int nparts, int op, RAcc* out)
{
	RAcc acc = r_ident<T>(op);
	for (int t = threadIdx.x; t < nparts; t += R_THREADS) acc = r_comb<T>(op, acc, part[t]);
	acc = block_reduce<T>(op, acc);
	if (threadIdx.x == 0) out[0] = acc;
}

template <typename T>
T finish(int op, const RAcc& r)
{
	if (IsReal<T>::value) {
		switch (op) {
		case REDUCTION_MAX:
		case REDUCTION_MIN:
		case REDUCTION_PROD: return (T)r.a;
		case REDUCTION_SUM: return (T)(isfinite(r.a) ? r.a + r.b : r.a);
		case REDUCTION_LOGICAL_AND:
		case REDUCTION_LOGICAL_OR:
		case REDUCTION_LOGICAL_XOR: return (T)(r.u ? 1 : 0);
		default: return (T)(int)(unsigned)(r.u & 0xffffffffull);
		}
	}
	switch (op) {
	case REDUCTION_MAX:
	case REDUCTION_MIN: return (T)r.i;
	case REDUCTION_LOGICAL_AND:
	case REDUCTION_LOGICAL_OR:
	case REDUCTION_LOGICAL_XOR: return (T)(r.u ? 1 : 0);
	default:
		if (sizeof(T) == 1) return (T)(signed char)(unsigned char)(r.u & 0xffu);
		return (T)(int)(unsigned)(r.u & 0xffffffffull);
	}
}

template <typename T>
T reduce_substate(ModelX* m, void* q, int op, const char* fn)
{
	SubX* s = S(q);
	if (!m) {
		set_error(70, std::string(fn) + ": the substate has no model");
		return (T)0;
	}
	// a fused program's transient substate has no buffer content: write it out
	if (m->phase == Phase::Device && s->tstate != Transient::None && ensure_host(m)) return (T)0;
	const size_t n = m->cells();
	const bool on_device = m->phase == Phase::Device && !s->host_only;
	if (on_device && m->prog && m->prog->before_read(m)) return (T)0;
	if (n == 1 || op <= (int)REDUCTION_NONE || op > (int)REDUCTION_BINARY_XOR) {
		// no fold: the value of cell 0 (cal2DReduction.c:30)
		if (!on_device) return ((T*)s->h_cur())[0];
		T v0 = (T)0;
		if (CALHIP_CHECK(hipMemcpy(&v0, s->d_cur, sizeof(T), hipMemcpyDeviceToHost))) return (T)0;
		return v0;
	}
	if (ensure_stream(m)) return (T)0;
	if (!m->d_red && CALHIP_CHECK(hipMalloc(&m->d_red, sizeof(RAcc) * (R_MAX_BLOCKS + 1)))) return (T)0;
	const T* src = (const T*)s->d_cur;
	if (!on_device) {
		const size_t bytes = n * sizeof(T);
		if (m->stage_bytes < bytes) {
			if (m->d_stage) (void)hipFree(m->d_stage);
			m->d_stage = nullptr;
			m->stage_bytes = 0;
			if (CALHIP_CHECK(hipMalloc(&m->d_stage, bytes))) return (T)0;
			m->stage_bytes = bytes;
		}
		if (CALHIP_CHECK(hipMemcpyAsync(m->d_stage, s->h_cur(), bytes, hipMemcpyHostToDevice, m->stream)))
			return (T)0;
		src = (const T*)m->d_stage;
	}
	const long long blocks_needed = ((long long)n + R_THREADS - 1) / R_THREADS;
	const int nb = (int)(blocks_needed < R_MAX_BLOCKS ? blocks_needed : R_MAX_BLOCKS);
	RAcc* part = (RAcc*)m->d_red;
	reduce_blocks<T><<<nb, R_THREADS, 0, m->stream>>>(src, (long long)n, op, part);
	reduce_final<T><<<1, R_THREADS, 0, m->stream>>>(part, nb, op, part + R_MAX_BLOCKS);
	if (CALHIP_CHECK(hipGetLastError())) return (T)0;
	RAcc r;
	T v0 = (T)0;
	if (CALHIP_CHECK(hipMemcpyAsync(&r, part + R_MAX_BLOCKS, sizeof r, hipMemcpyDeviceToHost, m->stream)) ||
	    CALHIP_CHECK(hipMemcpyAsync(&v0, src, sizeof(T), hipMemcpyDeviceToHost, m->stream)) ||
	    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
		return (T)0;
	// a NaN in cell 0 stays the result of the serial MAX/MIN fold
	if (IsReal<T>::value && (op == REDUCTION_MAX || op == REDUCTION_MIN) && v0 != v0) return v0;
	return finish<T>(op, r);
}

inline ModelX* owner_of(void* q) { return S(q)->owner; }

}  // namespace
}  // namespace calhip

using namespace calhip;

// calReductionComputeProd*: the reference's byte variants reduce with
// REDUCTION_PROD but its int and real variants with REDUCTION_SUM
// (OpenCAL-OMP/source/cal2DReduction.c:563-571, cal3DReduction.c:563-571); a
// drop-in model gets the same values here.
template <typename T>
constexpr REDUCTION_OPERATION compute_prod_op()
{
	return sizeof(T) == 1 ? REDUCTION_PROD : REDUCTION_SUM;
}

#define RED_FUNCS(D, SUF, T)                                                                                     \
	extern "C" T calReductionOperation##D##SUF(struct CALModel##D* model, struct CALSubstate##D##SUF* substate,   \
	                                            enum REDUCTION_OPERATION operation)                              \
	{                                                                                                            \
		(void)model;                                                                                             \
		return reduce_substate<T>(owner_of(substate), substate, (int)operation, "calReductionOperation" #D #SUF); \
	}                                                                                                            \
	extern "C" T getValue##D##SUF##AtIndex(struct CALSubstate##D##SUF* substate, CALint index)                   \
	{                                                                                                            \
		ModelX* m = owner_of(substate);                                                                          \
		if (m && host_read(m, substate)) return (T)0;                                                           \
		return substate->current[index];                                                                         \
	}                                                                                                            \
	extern "C" T calReductionComputeMax##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)            \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_MAX);                                               \
	}                                                                                                            \
	extern "C" T calReductionComputeMin##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)            \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_MIN);                                               \
	}                                                                                                            \
	extern "C" T calReductionComputeSum##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)            \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_SUM);                                               \
	}                                                                                                            \
	extern "C" T calReductionComputeProd##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)           \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, compute_prod_op<T>());                                        \
	}                                                                                                            \
	extern "C" T calReductionComputeLogicalAnd##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)     \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_LOGICAL_AND);                                       \
	}                                                                                                            \
	extern "C" T calReductionComputeBinaryAnd##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)      \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_BINARY_AND);                                        \
	}                                                                                                            \
	extern "C" T calReductionComputeLogicalOr##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)      \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_LOGICAL_OR);                                        \
	}                                                                                                            \
	extern "C" T calReductionComputeBinaryOr##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)       \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_BINARY_OR);                                         \
	}                                                                                                            \
	extern "C" T calReductionComputeLogicalXor##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)     \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_LOGICAL_XOR);                                       \
	}                                                                                                            \
	extern "C" T calReductionComputeBinaryXor##D##SUF(struct CALModel##D* m, struct CALSubstate##D##SUF* q)      \
	{                                                                                                            \
		return calReductionOperation##D##SUF(m, q, REDUCTION_BINARY_XOR);                                        \
	}

RED_FUNCS(2D, b, CALbyte)
RED_FUNCS(2D, i, CALint)
RED_FUNCS(2D, r, CALreal)
RED_FUNCS(3D, b, CALbyte)
RED_FUNCS(3D, i, CALint)
RED_FUNCS(3D, r, CALreal)
