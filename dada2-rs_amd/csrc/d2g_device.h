// d2g_device.h — wavefront helpers shared by the DP kernels (device code).
#pragma once

#include <hip/hip_runtime.h>

namespace d2g {

// lane i <- lane i-1 (DPP wave_shr:1); lane 0 keeps `old`
__device__ __forceinline__ int from_lower(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}
// lane i <- lane i+1 (DPP wave_shl:1); lane 63 keeps `old`
__device__ __forceinline__ int from_upper(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}

// The same shifts where the lane past the wavefront's end reads 0 (for
// callers whose edge lanes discard the value)
__device__ __forceinline__ int shift_from_lower(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int shift_from_upper(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace d2g
