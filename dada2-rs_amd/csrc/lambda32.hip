// lambda32.hip — mixed-precision lambda and the FP64 abundance p-values over
// a resident set of aligned pairs.
//
// A pair is a query (bases 1..4, Phred qualities) with its al2subs list
// against the center: records (query position << 16 | center base), in
// ascending query position.  compute_lambda (pval.rs:253-320) is the ordered
// product over query positions of err[t][q], t = 5 (nt-1) on the diagonal and
// 4 (nt0-1) + (nt-1) at a substitution.
//
// Pass 1 (pairs_loglam_kernel): log lambda as an FP32 compensated sum of a
// table log_err[t][q] = (float)log(err[t][q]) computed in FP64, the terms
// split over all 64 lanes and the lanes' (sum, error) accumulators joined by
// TwoSum in a fixed tree.  Each table entry carries one FP32 rounding
// (<= 2^-24 relative) and every term has the same sign, so the sum's
// relative error stays near 2^-24 plus the compensated-summation bound.
// pairs_pval_kernel
// (one pair per lane) then takes get_pA / calc_pA (pval.rs:376-398, 208-241)
// in FP64 from lambda =
// exp(log lambda): the sentinels (singleton, hamming 0, lambda 0), the
// Poisson upper tail P(X >= reads) = P(reads, E) (regularised lower
// incomplete gamma, E = lambda * center reads), the reference's direct
// series where that returns 0 (pval.rs:329-353), and the 1 - e^{-E}
// conditioning with its Taylor replacement below 1e-7.  Pairs whose p-value
// lands within a factor 10 of omega_a are appended to a list instead.
//
// Pass 2 (pairs_recompute_kernel): for the listed pairs only, lambda as the
// exact ordered FP64 product and the p-value again, so the p < omega_a call
// is taken on FP64 lambda wherever it is close.
//
// The recompute feeds its per-pair chains from LDS, as the compare's lambda
// kernel does: a workgroup of four wavefronts takes 64 pairs, each wavefront
// stages its 16 pairs' (transition row << 2 | nt - 1, quality) codes with
// coalesced loads (lane l: positions 4l .. 4l + 3 of a pair's rows), the
// owning lanes (0..15) patch their substitution positions, and each owning
// lane runs its chain from its row.
//
// The incomplete gamma is statrs 0.18's gamma_lr (d2g_statrs.h) with the
// device exp/log/sin.
#include <float.h>

#include "d2g_launch.h"
#include "d2g_statrs.h"

namespace d2g {

constexpr int kPSeg = 256;              // query positions per staged segment (4 per lane)
constexpr int kPRow = 2 * kPSeg + 4;    // bytes of one pair's LDS row (odd word stride)
constexpr int kPWaves = 4;              // wavefronts per workgroup
constexpr int kPPairsW = 16;            // pairs per wavefront
constexpr int kPTile = kPWaves * kPPairsW;
constexpr int kPUnroll = 8;             // pairs staged per round of loads
constexpr int kPSubBatch = 8;           // substitution records requested together per lane
constexpr uint32_t kPOneRow = 16;       // transition row of the past-the-end entry

struct PairDev {
  uint32_t n;
  const uint8_t* seq;       // rows padded to 16 bytes (the pool stays under 2^32 bytes)
  const uint8_t* qual;      // nullable, same layout
  const uint64_t* off;      // n + 1
  const uint32_t* len;      // n
  const uint32_t* subs;     // (query pos << 16 | center base), under 2^32 records
  const uint64_t* sub_off;  // n + 1
  const uint32_t* reads;
  const uint32_t* center_reads;
};

template <class T>
__device__ __forceinline__ T pld(const void* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(base) + byte_off);
}
__device__ __forceinline__ uint32_t prl(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
__device__ __forceinline__ void pwave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One owning lane's pair (lanes 0..15 of each wavefront; others hold none).
struct PairRow {
  bool live;
  uint32_t t;        // pair index
  uint32_t lq;       // query length
  uint32_t qoff;     // row offset in the pools
  uint32_t s0, ns;   // substitution records
};

__device__ __forceinline__ PairRow pair_row(const PairDev& a, bool live, uint32_t t) {
  PairRow r;
  r.live = live;
  r.t = live ? t : 0u;
  r.lq = live ? a.len[r.t] : 0u;
  r.qoff = (uint32_t)a.off[r.t];
  const uint64_t so = a.sub_off[r.t];
  r.s0 = (uint32_t)so;
  r.ns = live ? (uint32_t)(a.sub_off[r.t + 1] - so) : 0u;
  return r;
}

// Stages the segment at s0 of the wavefront's 16 pairs into their rows
// (rows + p * kPRow): two rounds of kPUnroll pairs, the second round's words
// requested before the first round's bytes are converted.  Positions past a
// pair's end get transition row kPOneRow.
__device__ __forceinline__ void stage_pairs(const PairDev& a, const PairRow& me, uint32_t s0,
                                            uint32_t lane, bool quals, uint8_t* rows) {
  const uint8_t* qual = quals ? a.qual : a.seq;
  uint32_t qw[2][kPUnroll], uw[2][kPUnroll], lq[2][kPUnroll];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < kPUnroll; ++r) {
      const uint32_t pr = (uint32_t)(h * kPUnroll + r);
      lq[h][r] = prl(me.lq, pr);
      const uint32_t pq = prl(me.qoff, pr);
      const uint32_t pos = s0 + 4u * lane;
      const uint32_t o = pos < lq[h][r] ? pq + pos : pq;  // past the end: the row's first word
      qw[h][r] = pld<uint32_t>(a.seq, o);
      uw[h][r] = pld<uint32_t>(qual, o);
    }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < kPUnroll; ++r) {
      const uint32_t pr = (uint32_t)(h * kPUnroll + r);
      uint32_t packed[2] = {0u, 0u};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t pos = s0 + 4u * lane + (uint32_t)b;
        const bool in = pos < lq[h][r];
        const uint32_t nt1 = ((qw[h][r] >> (8 * b)) & 0xFFu) - 1u;
        const uint32_t tr = in ? 5u * nt1 : kPOneRow;
        const uint32_t qi = (quals && in) ? (uw[h][r] >> (8 * b)) & 0xFFu : 0u;
        const uint32_t code = (((tr << 2) | (nt1 & 3u)) & 0xFFu) | (qi << 8);
        packed[b >> 1] |= code << (16 * (b & 1));
      }
      uint32_t* row = reinterpret_cast<uint32_t*>(rows + pr * kPRow) + 2 * lane;
      row[0] = packed[0];
      row[1] = packed[1];
    }
  }
}

// The owning lanes rewrite the transition row at their substitution
// positions of the segment (records carry the center base).
__device__ __forceinline__ void patch_pairs(const PairDev& a, const PairRow& me, uint32_t s0,
                                            uint8_t* myb) {
  for (uint32_t k0 = 0; __builtin_amdgcn_ballot_w64(k0 < me.ns); k0 += kPSubBatch) {
    uint32_t rec[kPSubBatch];
#pragma unroll
    for (int j = 0; j < kPSubBatch; ++j) {
      const uint32_t k = k0 + (uint32_t)j;
      rec[j] = pld<uint32_t>(a.subs, (k < me.ns ? me.s0 + k : 0u) << 2);
    }
#pragma unroll
    for (int j = 0; j < kPSubBatch; ++j) {
      const uint32_t k = k0 + (uint32_t)j;
      const uint32_t qp = rec[j] >> 16;
      const bool here = k < me.ns && qp >= s0 && qp - s0 < (uint32_t)kPSeg;
      uint8_t* c = myb + (here ? 2 * (qp - s0) : 2 * kPSeg);  // else the row's pad byte
      const uint32_t nt1 = (uint32_t)c[0] & 3u;
      c[0] = (uint8_t)(((4u * ((rec[j] & 0xFFFFu) - 1u) + nt1) << 2) | nt1);
    }
  }
}

// poisson_upper_tail_direct, pval.rs:329-353.  log_kfact[r] = sum_{i=1..r}
// log(i), the reference's running sum, tabulated on the host in the same
// order.
__device__ double poisson_tail_direct(uint32_t reads, double lambda, const double* log_kfact) {
  const double k = (double)reads;
  const double log_lambda = log(lambda);
  const double log_leading = -lambda + k * log_lambda - log_kfact[reads];
  double series = 1.0, term = 1.0;
  for (uint32_t j = 1; j < 10000u; ++j) {
    term *= lambda / (k + (double)j);
    if (!isfinite(term) || term <= 0.0) break;
    const double ns = series + term;
    if (ns == series) break;
    series = ns;
  }
  return exp(log_leading + log(series));
}

// get_pA with prior = detect_singletons = false (b_p_update's per-raw call)
__device__ double get_pA_dev(uint32_t reads, double lambda, uint32_t hamming, uint32_t bi_reads,
                             const double* log_kfact) {
  if (reads == 1) return 1.0;
  if (hamming == 0) return 1.0;
  if (lambda == 0.0) return 0.0;
  const double e = lambda * (double)bi_reads;
  if (!(e > 0.0)) return 1.0;
  double p = statrs_gamma_lr((double)reads, e);
  if (p == 0.0) p = poisson_tail_direct(reads, e, log_kfact);
  double norm = 1.0 - exp(-e);
  if (norm < 1e-7) norm = e - 0.5 * e * e;
  return p / norm;
}

// Pass 1: FP32 compensated log-lambda of every pair, all 64 lanes.
// A wavefront takes 8 pairs at a time; lane l reads positions 4l .. 4l + 3
// of each pair's segment (coalesced words of the sequence and quality rows)
// and looks its four terms up in the log table, with the segment's
// substitutions patched in through a per-pair byte row in LDS (the record's
// center base; 0 = diagonal).  Each lane adds its four terms in FP32 and
// keeps, per pair, a sum of those partials and its compensation (each
// addition's exact rounding error, by TwoSum), and one
// transposing reduction combines the lanes' accumulators pairwise over lane
// bits 3..5, then 0..2, so pair j's total lands on lanes 8j .. 8j + 7.  Every
// pair is summed by the same tree whatever its slot or wavefront, so the
// result repeats run to run and does not depend on the grid.  An err entry
// of 0 (log -inf) makes lambda 0: it is flagged, not summed.
struct F2 {
  float s, c;
};
// a + t: the rounded sum, the rounding error added to the compensation
__device__ __forceinline__ F2 f2_add(const F2 a, const float t) {
  const float s = a.s + t;
  const float bb = s - a.s;
  const float e = (a.s - (s - bb)) + (t - bb);
  return F2{s, a.c + e};
}
// a + b of two accumulators; symmetric in a and b (TwoSum's error is exact)
__device__ __forceinline__ F2 f2_join(const F2 a, const F2 b) {
  const float s = a.s + b.s;
  const float bb = s - a.s;
  const float e = (a.s - (s - bb)) + (b.s - bb);
  return F2{s, (a.c + b.c) + e};
}
__device__ __forceinline__ F2 f2_xor(const F2 v, const int m) {
  return F2{__shfl_xor(v.s, m), __shfl_xor(v.c, m)};
}

constexpr int kLPairs = 8;   // pairs per wavefront group (lane bits 3..5 in the reduction)
constexpr int kLShift = 3;   // log2(64 / kLPairs): pair j's total lands on lanes 8j .. 8j + 7

__global__ __launch_bounds__(64 * kPWaves) void pairs_loglam_kernel(PairDev a, const double* err,
                                                                     uint32_t ncol, int use_quals,
                                                                     float* out_loglam) {
  extern __shared__ float logerr[];  // 16 ncol entries and a 0 after them, then the patch rows
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ne = 16u * ncol;
  for (uint32_t i = threadIdx.x; i <= ne; i += 64u * kPWaves)
    logerr[i] = i < ne ? (float)log(err[i]) : 0.0f;
  // per wavefront: one kPSeg-byte patch row per pair of the group, all zero
  // between uses
  uint8_t* patch = reinterpret_cast<uint8_t*>(logerr + ((ne + 4) & ~3u)) +
                   (size_t)wave * kLPairs * kPSeg;
  for (uint32_t i = lane; i < (uint32_t)(kLPairs * kPSeg / 4); i += 64u)
    reinterpret_cast<uint32_t*>(patch)[i] = 0u;
  const bool quals = use_quals && a.qual != nullptr;
  __syncthreads();
  const uint32_t gw = blockIdx.x * kPWaves + wave, nwv = gridDim.x * kPWaves;
  for (uint32_t base = gw * kLPairs; base < a.n; base += nwv * kLPairs) {
    // lanes 0..15 fetch their pair's fields; the pair loops read them by readlane
    const uint32_t tl = base + lane;
    const bool livel = lane < (uint32_t)kLPairs && tl < a.n;
    const uint32_t my_lq = livel ? a.len[tl] : 0u;
    const uint32_t my_qoff = livel ? (uint32_t)a.off[tl] : 0u;
    const uint64_t my_so = livel ? a.sub_off[tl] : 0ull;
    const uint32_t my_ns = livel ? (uint32_t)(a.sub_off[tl + 1] - my_so) : 0u;
    uint32_t lqmax = 0;
#pragma unroll
    for (int j = 0; j < kLPairs; ++j) lqmax = max(lqmax, prl(my_lq, (uint32_t)j));
    F2 acc[kLPairs];
#pragma unroll
    for (int j = 0; j < kLPairs; ++j) acc[j] = F2{0.0f, 0.0f};
    uint32_t zmask = 0;
    for (uint32_t s0 = 0; s0 < lqmax; s0 += kPSeg) {
      // every load of the segment in flight together: the sequence and
      // quality words, and record `lane` of each pair's substitution list
      const uint32_t pos = s0 + 4u * lane;
      uint32_t qw[kLPairs], uw[kLPairs], rq[kLPairs];
#pragma unroll
      for (int j = 0; j < kLPairs; ++j) {
        const uint32_t lq = prl(my_lq, (uint32_t)j);
        const uint32_t o = prl(my_qoff, (uint32_t)j) + pos;
        const bool in = pos < lq;
        qw[j] = in ? pld<uint32_t>(a.seq, o) : 0u;
        uw[j] = (in && quals) ? pld<uint32_t>(a.qual, o) : 0u;
        const uint32_t ns = prl(my_ns, (uint32_t)j);
        const uint32_t sb = prl((uint32_t)my_so, (uint32_t)j);
        rq[j] = lane < ns ? a.subs[(size_t)sb + lane] : 0xFFFFFFFFu;
      }
      // the segment's substitutions into the pairs' patch rows (records
      // past the first 64 of a pair, rare, in a loop)
#pragma unroll
      for (int j = 0; j < kLPairs; ++j) {
        const uint32_t qp = rq[j] >> 16;
        if (rq[j] != 0xFFFFFFFFu && qp >= s0 && qp - s0 < (uint32_t)kPSeg)
          patch[j * kPSeg + (qp - s0)] = (uint8_t)(rq[j] & 0xFFu);
        const uint32_t ns = prl(my_ns, (uint32_t)j);
        if (ns > 64u) {
          const uint32_t sb = prl((uint32_t)my_so, (uint32_t)j);
          for (uint32_t k = lane + 64u; k < ns; k += 64u) {
            const uint32_t rec = a.subs[(size_t)sb + k];
            const uint32_t q2 = rec >> 16;
            if (q2 >= s0 && q2 - s0 < (uint32_t)kPSeg) patch[j * kPSeg + (q2 - s0)] = (uint8_t)(rec & 0xFFu);
          }
        }
      }
      pwave_sync();
#pragma unroll
      for (int j = 0; j < kLPairs; ++j) {
        const uint32_t lq = prl(my_lq, (uint32_t)j);
        const uint32_t pw = *reinterpret_cast<const uint32_t*>(patch + j * kPSeg + 4u * lane);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          // positions past the end read the table's 0 entry
          const bool in = pos + (uint32_t)b < lq;
          // bases are 1..4 (checked when the pairs are loaded); the masks
          // keep every table index inside the 16 rows
          const uint32_t nt1 = (((qw[j] >> (8 * b)) & 0xFFu) - 1u) & 3u;
          const uint32_t cb = (pw >> (8 * b)) & 0xFFu;
          const uint32_t tr = cb ? 4u * ((cb - 1u) & 3u) + nt1 : 5u * nt1;
          const uint32_t qi = (uw[j] >> (8 * b)) & 0xFFu;
          const float term = logerr[in ? tr * ncol + qi : ne];
          if (term == -INFINITY)
            zmask |= 1u << j;
          else
            acc[j] = f2_add(acc[j], term);
        }
      }
      pwave_sync();
      // the patched bytes back to zero (LDS operations of one wavefront
      // complete in order, so the next segment's writes come after)
#pragma unroll
      for (int j = 0; j < kLPairs; ++j) {
        const uint32_t qp = rq[j] >> 16;
        if (rq[j] != 0xFFFFFFFFu && qp >= s0 && qp - s0 < (uint32_t)kPSeg)
          patch[j * kPSeg + (qp - s0)] = 0u;
        const uint32_t ns = prl(my_ns, (uint32_t)j);
        if (ns > 64u) {
          const uint32_t sb = prl((uint32_t)my_so, (uint32_t)j);
          for (uint32_t k = lane + 64u; k < ns; k += 64u) {
            const uint32_t q2 = a.subs[(size_t)sb + k] >> 16;
            if (q2 >= s0 && q2 - s0 < (uint32_t)kPSeg) patch[j * kPSeg + (q2 - s0)] = 0u;
          }
        }
      }
      pwave_sync();
    }
    // reduce-scatter over lane bits 3..5: at bit b, pairs (2i, 2i + 1) become
    // slot i, lanes with bit b clear holding pair 2i and the others 2i + 1,
    // each the join of its own and its partner lane's accumulators
#pragma unroll
    for (int step = 0; step < 6 - kLShift; ++step) {
      const int bit = (1 << kLShift) << step;
      const bool hi = (lane & (uint32_t)bit) != 0;
#pragma unroll
      for (int i = 0; i < kLPairs / 2; ++i) {
        if (i >= (kLPairs >> (step + 1))) break;
        const F2 give = hi ? acc[2 * i] : acc[2 * i + 1];
        const F2 keep = hi ? acc[2 * i + 1] : acc[2 * i];
        acc[i] = f2_join(keep, f2_xor(give, bit));
      }
    }
    F2 tot = acc[0];
#pragma unroll
    for (int bit = 1; bit < (1 << kLShift); bit <<= 1) tot = f2_join(tot, f2_xor(tot, bit));
    const uint32_t j_me = lane >> kLShift;
    bool zero = false;
#pragma unroll
    for (int j = 0; j < kLPairs; ++j)
      if (__builtin_amdgcn_ballot_w64((zmask >> j) & 1u) && j_me == (uint32_t)j) zero = true;
    const uint32_t t = base + j_me;
    if ((lane & ((1u << kLShift) - 1u)) == 0 && t < a.n)
      out_loglam[t] = zero ? -INFINITY : tot.s + tot.c;
  }
}

// The first p-values, one pair per lane: lambda = exp(log lambda); pairs
// whose p-value lands within a factor 10 of omega go to `list` instead.
__global__ __launch_bounds__(256) void pairs_pval_kernel(PairDev a, const float* loglam,
                                                         const double* log_kfact, double omega,
                                                         double* out_lambda, double* out_p,
                                                         uint8_t* out_flag, uint32_t* list,
                                                         uint32_t* n_list) {
  const uint32_t lane = __lane_id();
  for (uint32_t b = blockIdx.x * blockDim.x; b < a.n; b += gridDim.x * blockDim.x) {
    const uint32_t t = b + threadIdx.x;
    bool near = false;
    if (t < a.n) {
      const float ll = loglam[t];
      const double lam = ll == -INFINITY ? 0.0 : exp((double)ll);
      const uint32_t ham = (uint32_t)(a.sub_off[t + 1] - a.sub_off[t]);
      const double p = get_pA_dev(a.reads[t], lam, ham, a.center_reads[t], log_kfact);
      near = p >= omega * 0.1 && p <= omega * 10.0;
      if (!near) {
        out_lambda[t] = lam;
        out_p[t] = p;
        out_flag[t] = 0;
      }
    }
    // append the near pairs: one atomic per wavefront, ranks from the ballot
    const uint64_t m = __builtin_amdgcn_ballot_w64(near);
    if (m) {
      uint32_t b0 = 0;
      if (lane == 0) b0 = atomicAdd(n_list, (uint32_t)__builtin_popcountll(m));
      b0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)b0);
      if (near) list[b0 + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull))] = t;
    }
  }
}

// Pass 2: the listed pairs' exact ordered FP64 product and p-value.
__global__ __launch_bounds__(64 * kPWaves) void pairs_recompute_kernel(
    PairDev a, const double* err, uint32_t ncol, int use_quals, const uint32_t* list,
    const uint32_t* n_list, const double* log_kfact, double* out_lambda, double* out_p,
    uint8_t* out_flag) {
  extern __shared__ double errd[];  // 16 ncol entries and 1.0 after them
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
  const uint32_t ne = 16u * ncol;
  const uint32_t cnt = *n_list;
  if (blockIdx.x * (uint32_t)kPTile >= cnt) return;  // uniform over the workgroup
  for (uint32_t i = threadIdx.x; i <= ne; i += 64u * kPWaves) errd[i] = i < ne ? err[i] : 1.0;
  uint8_t* rows = reinterpret_cast<uint8_t*>(errd + ne + 1) + wave * kPPairsW * kPRow;
  uint8_t* myb = rows + min(lane, (uint32_t)kPPairsW - 1u) * kPRow;
  const bool quals = use_quals && a.qual != nullptr;
  __syncthreads();
  for (uint32_t base = blockIdx.x * (uint32_t)kPTile; base < cnt; base += gridDim.x * (uint32_t)kPTile) {
    const uint32_t k = base + wave * kPPairsW + lane;
    const bool live = lane < (uint32_t)kPPairsW && k < cnt;
    const PairRow me = pair_row(a, live, live ? list[k] : 0u);
    double lam = 1.0;
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < me.lq); s0 += kPSeg) {
      stage_pairs(a, me, s0, lane, quals, rows);
      patch_pairs(a, me, s0, myb);
      pwave_sync();
      const uint32_t seg_n = me.lq > s0 ? min(me.lq - s0, (uint32_t)kPSeg) : 0u;
      if (lane < (uint32_t)kPPairsW) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(myb);
        for (uint32_t i = 0; __builtin_amdgcn_ballot_w64(i < seg_n); i += 32) {
          uint32_t w16[16];
#pragma unroll
          for (int h = 0; h < 16; ++h) w16[h] = w[(i >> 1) + (uint32_t)h];
          double f[32];
#pragma unroll
          for (int e = 0; e < 32; ++e) {
            const uint32_t code = (w16[e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
            f[e] = errd[((code & 0xFFu) >> 2) * ncol + (code >> 8)];  // past the end: 1.0
          }
#pragma unroll
          for (int e = 0; e < 32; ++e) lam *= f[e];
        }
      }
      pwave_sync();
    }
    if (me.live) {
      out_lambda[me.t] = lam;
      out_p[me.t] = get_pA_dev(a.reads[me.t], lam, me.ns, a.center_reads[me.t], log_kfact);
      out_flag[me.t] = 1;
    }
  }
}

hipError_t launch_pairs_lambda(const PairArgs& s, hipStream_t st) {
  PairDev d;
  d.n = s.n;
  d.seq = s.seq;
  d.qual = s.qual;
  d.off = s.off;
  d.len = s.len;
  d.subs = s.subs;
  d.sub_off = s.sub_off;
  d.reads = s.reads;
  d.center_reads = s.center_reads;
  if (s.n == 0) return hipSuccess;
  const uint32_t tiles = (s.n + kPTile - 1) / kPTile;
  const uint32_t blocks = min(tiles, s.max_blocks);
  const size_t rows = (size_t)kPTile * kPRow;
  const size_t lds1 = (((size_t)16 * s.ncol + 4) & ~(size_t)3) * sizeof(float) +
                      (size_t)kPWaves * kLPairs * kPSeg;
  const size_t lds2 = ((size_t)16 * s.ncol + 1) * sizeof(double) + rows;
  if (lds1 > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)pairs_loglam_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
    if (e != hipSuccess) return e;
  }
  if (lds2 > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)pairs_recompute_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(pairs_loglam_kernel, dim3(blocks), dim3(64 * kPWaves), lds1, st, d, s.err,
                     s.ncol, s.use_quals, s.out_loglam);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (s.ev_mid) {
    e = hipEventRecord(s.ev_mid, st);
    if (e != hipSuccess) return e;
  }
  const uint32_t pblocks = min((s.n + 255) / 256, s.max_blocks);
  hipLaunchKernelGGL(pairs_pval_kernel, dim3(pblocks), dim3(256), 0, st, d,
                     (const float*)s.out_loglam, s.log_kfact, s.omega, s.out_lambda, s.out_p,
                     s.out_flag, s.list, s.n_recomputed);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pairs_recompute_kernel, dim3(blocks), dim3(64 * kPWaves), lds2, st, d, s.err,
                     s.ncol, s.use_quals, (const uint32_t*)s.list, (const uint32_t*)s.n_recomputed,
                     s.log_kfact, s.out_lambda, s.out_p, s.out_flag);
  return hipGetLastError();
}

}  // namespace d2g
