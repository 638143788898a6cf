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

// Workgroups are dispatched round-robin over the 8 XCDs; this remaps the
// launch index so that each XCD receives one contiguous range of logical
// tiles (neighbouring tiles then share that XCD's L2).
__device__ __forceinline__ int xcd_logical_block(int b, int per_xcd)
{
	return (b & 7) * per_xcd + (b >> 3);
}

}  // namespace dev
}  // namespace calhip
