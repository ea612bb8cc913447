// prep.hip — device construction of every unique's resident k-mer screen.
//
// Restates raw_assign_kmers (kmers.rs:213-221): the dense 8-bit histogram
// of base-4 window codes (assign_kmer8, kmers.rs:96-101, u16 counts clamped
// to 255) and the positional kord vector (assign_kmer_order,
// kmers.rs:129-135).  One workgroup per unique; the u16 histogram lives in
// LDS as packed pairs.  Both layouts of each vector are written:
//   *_row[u][w]   read by the screen through scalar loads (centers)
//   *_il[w][u]    read by the screen lane-per-unique, coalesced (queries)
// Per-vector byte sums and a "has a 255 bucket" flag let the screen compute
// sum(min(a,b)) = (sum a + sum b - sum|a-b|) / 2 with v_sad_u8 and detect the
// reference's saturation case (kmers.rs:259) exactly.
#include "d2g_launch.h"

namespace d2g {

__global__ __launch_bounds__(256) void prep_kernel(UniqueDev u) {
  extern __shared__ uint32_t hist[];  // 4^k / 2 words, two u16 counts each
  __shared__ uint32_t red_sum[256];
  __shared__ uint32_t red_255[256];
  const uint32_t q = blockIdx.x;
  const int tid = threadIdx.x;
  const uint32_t k = u.k;
  const uint32_t nk = 1u << (2 * k);
  const uint32_t len = u.len[q];
  const uint8_t* seq = u.seq + u.off[q];
  for (uint32_t i = tid; i < nk / 2; i += 256) hist[i] = 0;
  __syncthreads();
  // windows: counts and kord (two windows per kord word)
  const uint32_t nwin = len >= k ? len - k + 1 : 0;
  uint16_t* kord_row16 = reinterpret_cast<uint16_t*>(u.kord_row + (size_t)q * u.owords);
  for (uint32_t s = tid; s < nwin; s += 256) {
    uint32_t idx = 0;
    bool ok = true;
    for (uint32_t t = 0; t < k; ++t) {
      uint32_t nt = seq[s + t];
      ok = ok && nt >= 1 && nt <= 4;
      idx = idx * 4 + (nt - 1);
    }
    if (ok) atomicAdd(&hist[idx >> 1], 1u << (16 * (idx & 1)));
    kord_row16[s] = ok ? (uint16_t)idx : (uint16_t)0;  // kmers.rs:132-134
  }
  __syncthreads();
  // interleaved kord copy
  for (uint32_t w = tid; w < u.owords; w += 256)
    u.kord_il[(size_t)w * u.npad + q] = u.kord_row[(size_t)q * u.owords + w];
  // kmer8 words: 4 buckets each, min(count, 255)
  uint32_t sum = 0, any255 = 0;
  for (uint32_t w = tid; w < u.kwords; w += 256) {
    uint32_t h01 = hist[2 * w], h23 = hist[2 * w + 1];
    uint32_t c0 = h01 & 0xFFFF, c1 = h01 >> 16, c2 = h23 & 0xFFFF, c3 = h23 >> 16;
    c0 = c0 > 255 ? 255 : c0;
    c1 = c1 > 255 ? 255 : c1;
    c2 = c2 > 255 ? 255 : c2;
    c3 = c3 > 255 ? 255 : c3;
    uint32_t word = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
    sum += c0 + c1 + c2 + c3;
    any255 |= (c0 == 255) | (c1 == 255) | (c2 == 255) | (c3 == 255);
    u.kmer_row[(size_t)q * u.kwords + w] = word;
    u.kmer_il[(size_t)w * u.npad + q] = word;
  }
  red_sum[tid] = sum;
  red_255[tid] = any255;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red_sum[tid] += red_sum[tid + st];
      red_255[tid] |= red_255[tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    u.ksum[q] = red_sum[0];
    u.has255[q] = (uint8_t)red_255[0];
  }
}

hipError_t launch_prep(const UniqueDev& u, hipStream_t s) {
  size_t lds = (size_t)(1u << (2 * u.k)) / 2 * sizeof(uint32_t);
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)prep_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(prep_kernel, dim3(u.n), dim3(256), lds, s, u);
  return hipGetLastError();
}

}  // namespace d2g
