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

// Adds one per lane with `pred` to counter[key]; the lanes of a wavefront
// with equal keys share one atomic (device-scope atomics on one word
// serialise across the chip, and equal-length pairs all share one bin).
// Returns each such lane's slot: the counter's old value plus the lane's
// rank among its key's lanes.  The key groups are found with ballots first,
// then every group's leader issues its atomic in one instruction, so a call
// costs one memory round trip however many keys the wavefront holds.  Every
// lane of the wavefront must call this.
__device__ __forceinline__ uint32_t wave_add_by_key(uint32_t* counter, uint32_t key, bool pred) {
  const uint32_t lane = __lane_id();
  uint64_t pending = __builtin_amdgcn_ballot_w64(pred);
  bool leads = false;
  uint32_t count = 0, rank = 0;
  int leader_of_mine = (int)lane;
  while (pending) {
    const int leader = __builtin_ffsll((long long)pending) - 1;
    const uint32_t k = (uint32_t)__shfl((int)key, leader);
    const uint64_t grp = __builtin_amdgcn_ballot_w64(pred && key == k) & pending;
    if ((int)lane == leader) {
      leads = true;
      count = (uint32_t)__builtin_popcountll(grp);
    }
    if ((grp >> lane) & 1ull) {
      leader_of_mine = leader;
      rank = (uint32_t)__builtin_popcountll(grp & ((1ull << lane) - 1ull));
    }
    pending &= ~grp;
  }
  uint32_t base = 0;
  if (leads) base = atomicAdd(&counter[key], count);
  base = (uint32_t)__shfl((int)base, leader_of_mine);
  return pred ? base + rank : 0u;
}

}  // namespace d2g
