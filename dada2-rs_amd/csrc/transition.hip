// transition.hip — the transition-by-quality counts of a finished
// partition: each counted raw aligned to its cluster center, every center
// position aligned to a query base adding the raw's weight to
// mat[4 (center base - 1) + (query base - 1)][quality of the query base].
// Two callers of the reference count this way:
//   * transition_counts (error.rs:194-236) over compute_final_subs
//     (dada.rs:646-673): qualities >= ncol are not counted (error.rs:229),
//     the u32 sums wrap;
//   * build_trans_mat (learn_errors.rs:618-690): qualities are clamped to
//     nq - 1 and the u32 sums saturate.
// The kernel keeps exact u64 sums (an LDS matrix per workgroup, added to the
// global one at the end); the host wraps or saturates them.  A gapless pair
// aligns position p with position p (align_gapless_with_buf,
// nwalign.rs:645-655); a DP pair's alignment strings (center first, '-' for
// gaps) give the map that al2subs builds from them (nwalign.rs:1169-1224).
#include <algorithm>

#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

struct TransDev {
  UniqueDev u;
  const GaplessTask* gl;
  const DpTask* dp;
  const Counters* ctr;
  const uint8_t* al0;
  const uint8_t* al1;
  const uint64_t* al_off;
  const uint32_t* al_len;
  const uint32_t* weight;  // per sorted unique: reads when counted, else 0
  int use_quals;
  uint32_t ncol;           // quality columns counted (1 without qualities)
  int clamp;               // qualities >= ncol: clamped to ncol - 1, else skipped
  unsigned long long* mat; // [16][ncol], exact sums
};

__device__ __forceinline__ void trans_add(unsigned long long* h, uint32_t ncol, int clamp,
                                          uint32_t c, uint32_t q, uint32_t qual, uint32_t w) {
  // nt - 1 with saturating_sub (error.rs:221-222)
  const uint32_t t = 4u * (c > 0 ? c - 1u : 0u) + (q > 0 ? q - 1u : 0u);
  const uint32_t col = clamp ? min(qual, ncol - 1u) : qual;
  if (t < 16u && col < ncol) atomicAdd(&h[t * ncol + col], (unsigned long long)w);
}

// One wavefront per pair, the lanes over 64 consecutive alignment columns
// (coalesced loads); a column's query position is the count of query bases
// before it, from a ballot.  Columns are taken kTrChunks x 64 at a time so
// their loads are in flight together.
constexpr int kTrChunks = 4;
constexpr uint32_t kTrThreads = 512;

__global__ __launch_bounds__(kTrThreads) void transition_kernel(TransDev a) {
  extern __shared__ unsigned long long tr_lds[];
  const uint32_t ne = 16u * a.ncol;
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) tr_lds[i] = 0;
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t n_gl = a.ctr->n_gapless, n_all = n_gl + a.ctr->n_dp;
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t t = (uint32_t)uni((int)(blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u));
       t < n_all; t += waves) {
    const bool gl = t < n_gl;
    const uint32_t uq = gl ? a.gl[t].u_idx : a.dp[t - n_gl].u_idx;
    const uint32_t cq = gl ? a.gl[t].c_idx : a.dp[t - n_gl].c_idx;
    const uint32_t w = a.weight[uq];
    if (w == 0) continue;
    const uint8_t* qs = a.u.seq + a.u.off[uq];
    const uint8_t* ql = a.use_quals ? a.u.qual + a.u.off[uq] : nullptr;
    if (gl) {  // position p with position p; the shorter one padded with gaps
      const uint8_t* cs = a.u.seq + a.u.off[cq];
      const uint32_t len = min(a.u.len[uq], a.u.len[cq]);
      for (uint32_t k0 = 0; k0 < len; k0 += 64u * kTrChunks) {
        uint32_t c0[kTrChunks], c1[kTrChunks], qv[kTrChunks];
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j) {
          const uint32_t k = k0 + 64u * j + lane;
          const bool ok = k < len;
          c0[j] = ok ? cs[k] : 0u;
          c1[j] = ok ? qs[k] : 0u;
          qv[j] = (ok && ql) ? ql[k] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j)
          if (c0[j] != 0u) trans_add(tr_lds, a.ncol, a.clamp, c0[j], c1[j], qv[j], w);
      }
    } else {
      const uint8_t* x0 = a.al0 + a.al_off[uq];
      const uint8_t* x1 = a.al1 + a.al_off[uq];
      const uint32_t L = a.al_len[uq];
      uint32_t qbase = 0;  // query bases before this batch of columns
      for (uint32_t k0 = 0; k0 < L; k0 += 64u * kTrChunks) {
        uint32_t c0[kTrChunks], c1[kTrChunks], qp[kTrChunks];
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j) {
          const uint32_t k = k0 + 64u * j + lane;
          const bool ok = k < L;
          c0[j] = ok ? x0[k] : (uint32_t)'-';
          c1[j] = ok ? x1[k] : (uint32_t)'-';
        }
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j) {
          const uint64_t m = __builtin_amdgcn_ballot_w64(c1[j] != (uint32_t)'-');
          qp[j] = qbase + (uint32_t)__builtin_popcountll(m & below);
          qbase += (uint32_t)__builtin_popcountll(m);
        }
        uint32_t qv[kTrChunks];
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j) {
          const bool base = c0[j] != (uint32_t)'-' && c1[j] != (uint32_t)'-';
          qv[j] = (base && ql) ? ql[qp[j]] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kTrChunks; ++j)
          if (c0[j] != (uint32_t)'-' && c1[j] != (uint32_t)'-')
            trans_add(tr_lds, a.ncol, a.clamp, c0[j], c1[j], qv[j], w);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x)
    if (tr_lds[i]) atomicAdd(&a.mat[i], tr_lds[i]);
}

// compute_aux's per-raw sums (error.rs:55-140, 251-305) over the same final
// alignments: each counted raw's substitution count (Sub.nsubs, for the n0 /
// n1 of cluster_stats) and cluster_quality's per center position sums
// (reads, and quality x reads when qualities are used).  The raws come in
// cluster order and each wavefront takes a contiguous range of them, adding
// into its own LDS columns (a raw's aligned center positions are distinct
// across lanes, so no atomics there) and flushing them to the cluster's row
// when the cluster changes.  Read sums are u32 (wrapping like the
// reference's), quality sums exact u64.
struct AuxDev {
  UniqueDev u;
  const uint8_t* al0;
  const uint8_t* al1;
  const uint64_t* al_off;
  const uint32_t* al_len;
  const uint32_t* order;    // counted sorted uniques, grouped by cluster
  uint32_t n_order;
  const uint32_t* cluster;  // per sorted unique: its cluster
  const uint32_t* want;     // per sorted unique: its center (sorted index)
  const uint32_t* weight;   // per sorted unique
  int gapless;              // band 0: every pair positional
  int use_quals;
  uint32_t maxlen;          // row length of the sums
  uint32_t lds_stride;      // positions per wave in LDS (maxlen rounded to 64)
  uint32_t* nsubs;          // per sorted unique
  uint32_t* qual_reads;     // [clusters][maxlen]
  unsigned long long* qual_sum;
};

__device__ __forceinline__ void aux_flush(const AuxDev& a, uint32_t cl, uint32_t clen, uint32_t* nr,
                                          unsigned long long* qs, uint32_t lane) {
  for (uint32_t p = lane; p < clen; p += 64u) {
    const uint32_t r = nr[p];
    const unsigned long long q = qs[p];
    if (r) atomicAdd(&a.qual_reads[(size_t)cl * a.maxlen + p], r);
    if (q) atomicAdd(&a.qual_sum[(size_t)cl * a.maxlen + p], q);
    nr[p] = 0;
    qs[p] = 0;
  }
}

__global__ __launch_bounds__(256) void aux_kernel(AuxDev a) {
  extern __shared__ unsigned long long aux_lds[];
  const uint32_t lane = __lane_id();
  const uint32_t wib = threadIdx.x / 64u;
  unsigned long long* qs = aux_lds + (size_t)wib * a.lds_stride;
  uint32_t* nr = reinterpret_cast<uint32_t*>(aux_lds + (size_t)(blockDim.x / 64u) * a.lds_stride) +
                 (size_t)wib * a.lds_stride;
  for (uint32_t p = lane; p < a.lds_stride; p += 64u) {
    qs[p] = 0;
    nr[p] = 0;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  const uint32_t wave = (uint32_t)uni((int)(blockIdx.x * (blockDim.x / 64u) + wib));
  const uint32_t per = (a.n_order + waves - 1) / waves;
  const uint32_t k0 = min(a.n_order, wave * per), k1 = min(a.n_order, k0 + per);
  uint32_t cur = 0xFFFFFFFFu, cur_len = 0;
  for (uint32_t k = k0; k < k1; ++k) {
    const uint32_t uq = a.order[k];
    const uint32_t cl = a.cluster[uq], cq = a.want[uq], w = a.weight[uq];
    if (cl != cur) {
      if (cur != 0xFFFFFFFFu) aux_flush(a, cur, cur_len, nr, qs, lane);
      cur = cl;
      cur_len = min(a.u.len[cq], a.maxlen);  // center_len.min(maxlen), error.rs:268
    }
    const uint8_t* ql = a.use_quals ? a.u.qual + a.u.off[uq] : nullptr;
    uint32_t nsub = 0;
    if (a.gapless) {  // position p with position p up to the shorter length
      const uint8_t* cs = a.u.seq + a.u.off[cq];
      const uint8_t* xs = a.u.seq + a.u.off[uq];
      const uint32_t len = min(a.u.len[uq], a.u.len[cq]);
      for (uint32_t p = lane; p < len; p += 64u) {
        nsub += cs[p] != xs[p] ? 1u : 0u;
        if (p < cur_len) {
          nr[p] += w;
          if (ql) qs[p] += (unsigned long long)ql[p] * w;
        }
      }
    } else {
      const uint8_t* x0 = a.al0 + a.al_off[uq];
      const uint8_t* x1 = a.al1 + a.al_off[uq];
      const uint32_t L = a.al_len[uq];
      uint32_t cbase = 0, qbase = 0;  // center / query bases before the chunk
      for (uint32_t c0i = 0; c0i < L; c0i += 64u) {
        const uint32_t k2 = c0i + lane;
        const bool ok = k2 < L;
        const uint32_t c0 = ok ? x0[k2] : (uint32_t)'-';
        const uint32_t c1 = ok ? x1[k2] : (uint32_t)'-';
        const bool b0 = c0 != (uint32_t)'-', b1 = c1 != (uint32_t)'-';
        const uint64_t m0 = __builtin_amdgcn_ballot_w64(b0);
        const uint64_t m1 = __builtin_amdgcn_ballot_w64(b1);
        const uint32_t pos0 = cbase + (uint32_t)__builtin_popcountll(m0 & below);
        const uint32_t pos1 = qbase + (uint32_t)__builtin_popcountll(m1 & below);
        cbase += (uint32_t)__builtin_popcountll(m0);
        qbase += (uint32_t)__builtin_popcountll(m1);
        if (b0 && b1) {
          // al2subs (nwalign.rs:1169-1224): bases that differ, neither N (5)
          nsub += (c0 != c1 && c0 != 5u && c1 != 5u) ? 1u : 0u;
          if (pos0 < cur_len) {  // cluster_quality reads center positions below its length
            nr[pos0] += w;
            if (ql) qs[pos0] += (unsigned long long)ql[pos1] * w;
          }
        }
      }
    }
    for (int off = 32; off > 0; off >>= 1) nsub += __shfl_xor(nsub, off);
    if (lane == 0) a.nsubs[uq] = nsub;
  }
  if (cur != 0xFFFFFFFFu) aux_flush(a, cur, cur_len, nr, qs, lane);
}

hipError_t launch_aux(const AuxArgs& s, hipStream_t st) {
  AuxDev d;
  d.u = *s.u;
  d.al0 = s.al0;
  d.al1 = s.al1;
  d.al_off = s.al_off;
  d.al_len = s.al_len;
  d.order = s.order;
  d.n_order = s.n_order;
  d.cluster = s.cluster;
  d.want = s.want;
  d.weight = s.weight;
  d.gapless = s.gapless;
  d.use_quals = s.use_quals;
  d.maxlen = s.maxlen;
  d.lds_stride = (s.maxlen + 63u) & ~63u;
  d.nsubs = s.nsubs;
  d.qual_reads = s.qual_reads;
  d.qual_sum = s.qual_sum;
  // 12 bytes of LDS per position and wave; up to 4 waves per block
  const uint32_t per_wave = d.lds_stride * 12u;
  const uint32_t wpb = std::max(1u, std::min(4u, 61440u / per_wave));
  if (per_wave > 61440u) return hipErrorInvalidValue;
  if (s.n_order == 0) return hipSuccess;
  const uint32_t blocks = std::max(1u, std::min(s.grid, (s.n_order + wpb - 1) / wpb));
  hipLaunchKernelGGL(aux_kernel, dim3(blocks), dim3(64 * wpb), (size_t)per_wave * wpb, st, d);
  return hipGetLastError();
}

hipError_t launch_transition(const TransArgs& s, hipStream_t st) {
  TransDev d;
  d.u = *s.u;
  d.gl = s.gapless_tasks;
  d.dp = s.dp_tasks;
  d.ctr = s.ctr;
  d.al0 = s.al0;
  d.al1 = s.al1;
  d.al_off = s.al_off;
  d.al_len = s.al_len;
  d.weight = s.weight;
  d.use_quals = s.use_quals;
  d.ncol = s.ncol;
  d.clamp = s.clamp;
  d.mat = s.mat;
  const size_t lds = (size_t)16 * s.ncol * sizeof(unsigned long long);
  hipLaunchKernelGGL(transition_kernel, dim3(s.grid), dim3(kTrThreads), lds, st, d);
  return hipGetLastError();
}

}  // namespace d2g
