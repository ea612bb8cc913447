// bimera.hip — the per-pair part of the bimera parent search: from the
// (query, parent) alignment left in device memory by the vectorized DP
// (align_vectorized_with_buf on ASCII bytes, end_gap 0, band = max_shift;
// chimera.rs:234-240, 336-342), get_lr (chimera.rs:99-175) and
// get_ham_endsfree (chimera.rs:54-86), one lane per pair.
//
// out[5 i .. 5 i + 4] = left, right, left_oo, right_oo, ham.  left/right are
// the exact-match overlaps of the query (al0) with the parent (al1) from
// each end, after skipping end gaps in the query and crediting up to
// max_shift end gaps in the parent; the _oo variants continue across one
// mismatch when allow_one_off; ham counts mismatching columns between the
// first and last column where neither row is in an end gap.
#include "d2g_launch.h"

namespace d2g {

__global__ __launch_bounds__(256) void bimera_lr_kernel(uint32_t n, const uint8_t* al0,
                                                        const uint8_t* al1, const uint64_t* al_off,
                                                        const uint32_t* al_len, int allow_one_off,
                                                        uint32_t max_shift, uint32_t* out) {
  const uint8_t gap = (uint8_t)'-';
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const uint8_t* s0 = al0 + al_off[t];
    const uint8_t* s1 = al1 + al_off[t];
    const int len = (int)al_len[t];
    // left overlap
    int pos = 0;
    while (pos < len && s0[pos] == gap) ++pos;
    uint32_t left = 0;
    while (pos < len && s1[pos] == gap && left < max_shift) {
      ++pos;
      ++left;
    }
    while (pos < len && s0[pos] == s1[pos]) {
      ++pos;
      ++left;
    }
    uint32_t left_oo = 0;
    if (allow_one_off) {
      uint32_t loo = left;
      pos += 1;
      if (pos < len && s0[pos] != gap) ++loo;
      while (pos < len && s0[pos] == s1[pos]) {
        ++pos;
        ++loo;
      }
      left_oo = loo;
    }
    // right overlap
    int rp = len - 1;
    while (rp >= 0 && s0[rp] == gap) --rp;
    uint32_t right = 0;
    while (rp >= 0 && s1[rp] == gap && right < max_shift) {
      --rp;
      ++right;
    }
    while (rp >= 0 && s0[rp] == s1[rp]) {
      --rp;
      ++right;
    }
    uint32_t right_oo = 0;
    if (allow_one_off) {
      uint32_t roo = right;
      rp -= 1;
      if (rp >= 0 && s0[rp] != gap) ++roo;
      while (rp >= 0 && s0[rp] == s1[rp]) {
        --rp;
        ++roo;
      }
      right_oo = roo;
    }
    // hamming distance inside the end gaps of either row
    int i = 0;
    bool g0 = true, g1 = true;
    while ((g0 || g1) && i < len) {
      g0 = g0 && s0[i] == gap;
      g1 = g1 && s1[i] == gap;
      if (g0 || g1) ++i;
    }
    int j = len - 1;
    g0 = true;
    g1 = true;
    while ((g0 || g1) && j >= i) {
      g0 = g0 && s0[j] == gap;
      g1 = g1 && s1[j] == gap;
      if (g0 || g1) --j;
    }
    uint32_t ham = 0;
    for (int p = i; p <= j; ++p) ham += s0[p] != s1[p];
    uint32_t* o = out + 5 * (size_t)t;
    o[0] = left;
    o[1] = right;
    o[2] = left_oo;
    o[3] = right_oo;
    o[4] = ham;
  }
}

hipError_t launch_bimera_lr(uint32_t n, const uint8_t* al0, const uint8_t* al1,
                            const uint64_t* al_off, const uint32_t* al_len, int allow_one_off,
                            uint32_t max_shift, uint32_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(bimera_lr_kernel, dim3(blocks), dim3(256), 0, s, n, al0, al1, al_off, al_len,
                     allow_one_off, max_shift, out);
  return hipGetLastError();
}

}  // namespace d2g
