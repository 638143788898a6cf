// Device-side helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace calhip {
namespace dev {

constexpr int WAVE = 64;

// Single toroidal wrap (calGetToroidalX, calCommon.h:62).
__device__ __forceinline__ int wrap1(int x, int size)
{
	return x < 0 ? x + size : (x >= size ? x - size : x);
}

// Lane l receives lane l-1 / l+1 of the same wave (lanes at the ends receive
// an unspecified value; callers never use them).
__device__ __forceinline__ double from_left(double v) { return __shfl_up(v, 1, WAVE); }
__device__ __forceinline__ double from_right(double v) { return __shfl_down(v, 1, WAVE); }

// The same shifts with DPP wave shifts (one VALU op per 32 bits, no LDS
// round trip): lane l receives lane l-1 (wave_shr:1) / l+1 (wave_shl:1); the
// end lanes receive 0.
__device__ __forceinline__ double dpp_shift(double v, bool to_higher)
{
	const long long b = __double_as_longlong(v);
	int lo, hi;
	if (to_higher) {
		lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, false);
		hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
	} else {
		lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xf, 0xf, false);
		hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
	}
	return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double dpp_from_left(double v) { return dpp_shift(v, true); }
__device__ __forceinline__ double dpp_from_right(double v) { return dpp_shift(v, false); }

// Workgroups are dispatched round-robin over the 8 XCDs; this remaps the
// launch index so that each XCD receives one contiguous range of logical
// tiles (neighbouring tiles then share that XCD's L2).
__device__ __forceinline__ int xcd_logical_block(int b, int per_xcd)
{
	return (b & 7) * per_xcd + (b >> 3);
}

}  // namespace dev
}  // namespace calhip
