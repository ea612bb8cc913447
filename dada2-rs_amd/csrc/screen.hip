// screen.hip — the k-mer screen and method dispatch of the compare.
//
// Restates, per (center, unique) pair, cluster.rs:294-300 (greedy skip),
// raw_align_with_buf (nwalign.rs:1261-1301: kmer_dist8 with the u16
// overflow fallback, kord_dist when gapless, the strict `kdist > cutoff`
// shroud) and the first dispatch test of raw_align_dp (nwalign.rs:1324:
// band == 0 or |kodist - kdist| < f64::EPSILON -> gapless).
//
// Layout: a workgroup owns QB = 64, 128 or 256 length-sorted uniques and 16
// centers, split over SPLIT groups of QB/(64 QL) wavefronts: wavefront group
// h compares the QB uniques with centers 16/SPLIT * h .. 16/SPLIT * (h+1) - 1,
// each lane taking QL = 1 or 2 uniques (64 apart).  Query
// vectors are read coalesced from the interleaved layout (the groups of one
// workgroup read the same lines, so all but the first come from L2); center
// vectors come through scalar loads from the row layout.  A scalar load is
// waited for before its words are used (scalar loads return out of order),
// so a wavefront stalls once per (center, 16-word chunk); splitting the
// centers over more wavefronts gives the SIMD more of them to interleave,
// and a second unique per lane gives each stall twice the work.
// sum(min(a, b)) over the 4^k buckets is (sum a + sum b - sum |a - b|) / 2,
// with sum |a - b| from v_sad_u8 (four buckets per instruction).
//
// The shroud test kdist > cutoff, kdist = 1 - dot / n in f64 (n = min(len)
// - k + 1), is non-increasing in dot, so it is dot < dstar[n] with dstar
// tabulated on the host from the same f64 expression; kdist itself is
// formed only for the pairs that pass (for the gapless dispatch).
//
// Pairs that pass are appended to the block's own gapless or DP list (LDS
// counters; each block owns QB x 16 slots of each list) with their
// (len_center, len_query) bin; per-block counts go to `partials` and the
// scatter kernel compacts the lists.
#include <float.h>

#include "d2g_device.h"
#include <cstring>

#include "d2g_launch.h"

namespace d2g {

constexpr int kScreenChunk = 16;  // k-mer words per scalar load

struct ScreenDev {
  UniqueDev u;
  const uint32_t* dstar;  // [n]: smallest dot with kdist <= cutoff
  const uint32_t* centers;
  uint32_t n_centers;
  const uint8_t* lock;
  int greedy, use_kmers, gapless, band;
  double kdist_cutoff;
  BinSpace bins;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
  GaplessTask* gapless_tasks;
  DpTask* dp_unsorted;
  uint32_t* bin_count;
  uint32_t* partials;
  Counters* ctr;
  // the pairs of this pass: with a center of length <= long_len, uniques
  // [sr_lo, sr_hi); with a longer center, uniques [lr_lo, lr_hi)
  uint32_t long_len, sr_lo, sr_hi, lr_lo, lr_hi;
  uint32_t nx, ny;  // screen_kernel's grid: tiles of uniques x rows of 16 centers
};

__device__ __forceinline__ bool pair_in_pass(const ScreenDev& a, uint32_t center_len, uint32_t q) {
  return center_len <= a.long_len ? (q >= a.sr_lo && q < a.sr_hi) : (q >= a.lr_lo && q < a.lr_hi);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends to a workgroup-local list: one LDS atomic per wavefront, ranks
// from the ballot.  All lanes of the wavefront must call this.
__device__ __forceinline__ uint32_t wave_append(uint32_t* lds_counter, bool pred) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(pred);
  uint32_t base = 0;
  if (mask) {
    const int leader = __builtin_ffsll((long long)mask) - 1;
    const uint32_t cnt = (uint32_t)__builtin_popcountll(mask);
    if ((int)lane_id() == leader) base = atomicAdd(lds_counter, cnt);
    base = __shfl(base, leader);
  }
  const uint64_t below = mask & ((1ull << lane_id()) - 1ull);
  return base + (uint32_t)__builtin_popcountll(below);
}

__device__ __forceinline__ uint16_t kord_il_at(const UniqueDev& u, uint32_t q, uint32_t s) {
  uint32_t w = u.kord_il[(size_t)(s >> 1) * u.npad + q];
  return (uint16_t)((s & 1) ? (w >> 16) : (w & 0xFFFF));
}
__device__ __forceinline__ uint16_t kord_row_at(const UniqueDev& u, uint32_t q, uint32_t s) {
  return reinterpret_cast<const uint16_t*>(u.kord_row + (size_t)q * u.owords)[s];
}

// kmer_dist (kmers.rs:233-241) numerator sum_x min(cnt_a[x], cnt_b[x]) from
// the positional k-mer lists (all windows valid: bases are checked at load).
// Quadratic and lane-serial: reached only when kmer_dist8 saturates
// (a k-mer seen >= 255 times in both sequences).
__device__ uint32_t u16_min_dot(const UniqueDev& u, uint32_t qa, uint32_t qb) {
  uint32_t ka = u.len[qa] - u.k + 1, kb = u.len[qb] - u.k + 1;
  uint32_t dot = 0;
  for (uint32_t i = 0; i < ka; ++i) {
    uint16_t x = kord_il_at(u, qa, i);
    bool seen = false;
    for (uint32_t j = 0; j < i && !seen; ++j) seen = kord_il_at(u, qa, j) == x;
    if (seen) continue;
    uint32_t ca = 0, cb = 0;
    for (uint32_t j = i; j < ka; ++j) ca += kord_il_at(u, qa, j) == x;
    for (uint32_t j = 0; j < kb; ++j) cb += kord_row_at(u, qb, j) == x;
    dot += ca < cb ? ca : cb;
  }
  return dot;
}

// Per-unique fields a lane needs for each of its uniques.
struct QueryFields {
  uint32_t qq;   // the unique (sorted index; idle lanes repeat the last)
  bool valid;
  uint32_t len, reads, sum, thr;  // thr = dstar[len - k + 1]
  bool lock, has255;
};
// Per-center fields, wavefront-uniform.
struct CenterFields {
  uint32_t ci, len, reads, sum, thr;
  bool has255;
};

// Indices of a wavefront's CW centers from c0 (centers past the end repeat
// the last one; their results are unused) and of its lanes' QL uniques
// wave_base + 64 j + lane (lanes past the end repeat the last unique).
template <int QL, int CW>
__device__ __forceinline__ void load_ids(const ScreenDev& a, const uint32_t c0,
                                         const uint32_t wave_base, const uint32_t lane,
                                         CenterFields (&cf)[CW], QueryFields (&qf)[QL]) {
  const UniqueDev& u = a.u;
#pragma unroll
  for (int c = 0; c < CW; ++c) cf[c].ci = a.centers[min(c0 + (uint32_t)c, a.n_centers - 1)];
#pragma unroll
  for (int j = 0; j < QL; ++j) {
    const uint32_t q = wave_base + 64u * (uint32_t)j + lane;
    qf[j].valid = q < u.n;
    qf[j].qq = qf[j].valid ? q : u.n - 1;
  }
}

// The other fields of those centers and uniques.
template <int QL, int CW>
__device__ __forceinline__ void load_rest(const ScreenDev& a, CenterFields (&cf)[CW],
                                          QueryFields (&qf)[QL]) {
  const UniqueDev& u = a.u;
  const uint32_t k = u.k;
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    cf[c].len = u.len[cf[c].ci];
    cf[c].reads = u.reads[cf[c].ci];
    cf[c].sum = u.ksum[cf[c].ci];
    cf[c].has255 = u.has255[cf[c].ci] != 0;
  }
#pragma unroll
  for (int j = 0; j < QL; ++j) {
    qf[j].len = u.len[qf[j].qq];
    qf[j].reads = u.reads[qf[j].qq];
    qf[j].lock = a.lock ? a.lock[qf[j].qq] != 0 : false;
    qf[j].sum = u.ksum[qf[j].qq];
    qf[j].has255 = u.has255[qf[j].qq] != 0;
  }
  if (a.use_kmers) {
#pragma unroll
    for (int c = 0; c < CW; ++c) cf[c].thr = a.dstar[cf[c].len - k + 1];
#pragma unroll
    for (int j = 0; j < QL; ++j) qf[j].thr = a.dstar[qf[j].len - k + 1];
  }
}

// The decision for each (center, unique) pair of a lane, once its sums of
// |a - b| are in: greedy skip, shroud (with the exact kmer_dist8 where both
// vectors hold a 255 bucket), gapless or DP route, and the append to the
// block's task lists.  All lanes of the wavefront call it (the appends use
// ballots); nc centers of cf are live.
template <int QL, int CW, int KC>
__device__ __forceinline__ void screen_pairs(const ScreenDev& a, const QueryFields (&qf)[QL],
                                             const CenterFields (&cf)[CW],
                                             const uint32_t (&sad)[QL][CW],
                                             const int32_t (&keq)[QL][CW], const uint32_t nc,
                                             const uint32_t c0, uint32_t* s_ngl, uint32_t* s_ndp,
                                             GaplessTask* gl_region, DpTask* dp_region,
                                             uint32_t& n_skip, uint32_t& n_shroud) {
  const UniqueDev& u = a.u;
  const uint32_t k = u.k;
  // unrolled over the centers: the per-center arrays are read at constant
  // indices (a runtime index would put them in scratch memory)
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    if ((uint32_t)c >= nc) continue;
    const CenterFields& C = cf[c];
#pragma unroll
    for (int j = 0; j < QL; ++j) {
      const QueryFields& Q = qf[j];
      const bool in = Q.valid && pair_in_pass(a, C.len, Q.qq);
      const uint32_t sd = sad[j][c];
      // results stay in the length-sorted order (coalesced); d2g_fetch_results
      // restores the caller's unique order
      const uint32_t out_idx = (c0 + c) * u.n + Q.qq;
      // cluster.rs:295: greedy skip
      const bool skip = a.greedy && (Q.reads > C.reads || Q.lock);
      uint32_t dot = 0;
      bool shroud = false;
      if (a.use_kmers) {
        dot = (Q.sum + C.sum - sd) >> 1;
        if (Q.has255 && C.has255 && !skip) {
          // exact kmer_dist8 (kmers.rs:254-269): overflow when min == 255;
          // the words are loaded KC at a time (kwords is a multiple of 16),
          // so one round trip serves KC words
          bool ovf = false;
          uint32_t d2 = 0;
          for (uint32_t w0 = 0; w0 < u.kwords; w0 += KC) {
            uint32_t xs[KC], ys[KC];
#pragma unroll
            for (int t = 0; t < KC; ++t) {
              xs[t] = u.kmer_il[(size_t)(w0 + t) * u.npad + Q.qq];
              ys[t] = u.kmer_row[(size_t)C.ci * u.kwords + w0 + t];
            }
#pragma unroll
            for (int t = 0; t < KC; ++t)
#pragma unroll
              for (int b = 0; b < 4; ++b) {
                const uint32_t m = min((xs[t] >> (8 * b)) & 255u, (ys[t] >> (8 * b)) & 255u);
                ovf |= m == 255;
                d2 += m;
              }
          }
          dot = ovf ? u16_min_dot(u, Q.qq, C.ci) : d2;  // nwalign.rs:1270-1278
          if (ovf && in) atomicAdd(&a.ctr->n_exact_fallback, 1u);
        }
        // kdist > cutoff (raw_align_with_buf, nwalign.rs:1296)
        shroud = !skip && dot < (Q.len <= C.len ? Q.thr : C.thr);
      }
      // kord_dist (kmers.rs:314-328) needs the count of equal positions of an
      // equal-length pair that passes: precomputed (keq >= 0) or counted
      // here by the whole wavefront, one such pair at a time, from
      // coalesced reads of the two order rows and one wavefront sum (a
      // lane's own loop would wait on one load per word)
      int32_t eqc = keq[j][c];
      const bool needk = in && !skip && !shroud && a.use_kmers && a.gapless && Q.len == C.len &&
                         eqc < 0;
      for (uint64_t need = __builtin_amdgcn_ballot_w64(needk); need; need &= need - 1ull) {
        const int l = (int)__builtin_ctzll(need);
        const uint32_t qq = (uint32_t)__builtin_amdgcn_readlane((int)Q.qq, l);
        const uint32_t klen = (uint32_t)__builtin_amdgcn_readlane((int)Q.len, l) - (k - 1);
        const uint32_t nw = (klen + 1) / 2;
        const uint32_t* qrow = u.kord_row + (size_t)qq * u.owords;
        const uint32_t* crow = u.kord_row + (size_t)C.ci * u.owords;
        uint32_t cnt = 0;
        for (uint32_t w = lane_id(); w < nw; w += 64u) {
          const uint32_t x = qrow[w] ^ crow[w];
          cnt += (x & 0xFFFFu) == 0;
          if (2 * w + 1 < klen) cnt += (x >> 16) == 0;
        }
        for (int off = 32; off > 0; off >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, off);
        if ((int)lane_id() == l) eqc = (int32_t)cnt;
      }
      bool route_gapless = false;
      if (!skip && !shroud) {
        const uint32_t mn = Q.len < C.len ? Q.len : C.len;
        const double kdist = a.use_kmers ? 1.0 - (double)dot / (double)(mn - k + 1) : 0.0;
        double kodist = -1.0;
        if (a.use_kmers && a.gapless && Q.len == C.len) {
          const uint32_t klen = Q.len - (k - 1);
          const uint32_t eq = eqc >= 0 ? (uint32_t)eqc : 0u;
          kodist = 1.0 - (double)eq / (double)klen;
        }
        route_gapless = a.band == 0 || fabs(kodist - kdist) < DBL_EPSILON;
      }
      const bool route_dp = !skip && !shroud && !route_gapless;
      if (in && (skip || shroud)) {
        a.out_lambda[out_idx] = 0.0;
        a.out_hamming[out_idx] = kHammingNone;
        a.out_skipped[out_idx] = skip ? 1 : 0;
      }
      n_skip += in && skip;
      n_shroud += in && shroud;
      const bool to_gl = in && route_gapless;
      const uint32_t gpos = wave_append(s_ngl, to_gl);
      if (to_gl) {
        GaplessTask t;
        t.c_idx = C.ci;
        t.u_idx = Q.qq;
        t.out_idx = out_idx;
        t._pad = 0;
        gl_region[gpos] = t;
      }
      const bool to_dp = in && route_dp;
      const uint32_t dpos = wave_append(s_ndp, to_dp);
      if (to_dp) {
        DpTask t;
        t.s1_off = u.off[C.ci];
        t.s2_off = u.off[Q.qq];
        t.out_idx = out_idx;
        t.bin = (C.len - (uint32_t)a.bins.lmin) * (uint32_t)a.bins.range + (Q.len - (uint32_t)a.bins.lmin);
        t.u_idx = Q.qq;
        t.c_idx = C.ci;
        dp_region[dpos] = t;
        atomicAdd(&a.bin_count[t.bin], 1u);
      }
    }
  }
}

// Block totals of skipped / shrouded pairs and the block's list lengths,
// summed by the plan kernel (no device-scope atomics on shared counter
// words).  All threads of the block call it.
template <int NW>
__device__ __forceinline__ void screen_block_totals(const ScreenDev& a, const uint32_t blk,
                                                    const uint32_t n_skip, const uint32_t n_shroud,
                                                    const uint32_t& s_ngl, const uint32_t& s_ndp) {
  __shared__ uint32_t s_cnt[2][NW];
  uint32_t s1 = n_skip, s2 = n_shroud;
  for (int off = 32; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off);
    s2 += __shfl_xor(s2, off);
  }
  if (lane_id() == 0) {
    s_cnt[0][threadIdx.x >> 6] = s1;
    s_cnt[1][threadIdx.x >> 6] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t1 = 0, t2 = 0;
    for (int w = 0; w < NW; ++w) {
      t1 += s_cnt[0][w];
      t2 += s_cnt[1][w];
    }
    a.partials[4 * blk] = t1;
    a.partials[4 * blk + 1] = t2;
    a.partials[4 * blk + 2] = s_ngl;
    a.partials[4 * blk + 3] = s_ndp;
  }
}

template <int QB, int SPLIT, int QL>
__global__ __launch_bounds__(QB / QL * SPLIT) void screen_kernel(ScreenDev a) {
  constexpr int CW = (int)kScreenCenters / SPLIT;    // centers per wavefront
  constexpr int GT = QB / QL;                          // threads per wavefront group
  constexpr uint32_t kRegion = QB * kScreenCenters;  // one task slot per pair of the block
  const UniqueDev& u = a.u;
  // this block's task lists: kRegion slots each, counted in LDS
  // Workgroup order: workgroup L runs on XCD L mod 8, and XCD x takes the
  // logical blocks [x per, (x + 1) per) in unique-tile-major order, so the
  // center rows of one tile of uniques run on one XCD close in time and
  // re-read the tile's k-mer vectors from that XCD's L2.
  const uint32_t total = a.nx * a.ny;
  const uint32_t per = (total + 7u) / 8u;
  const uint32_t blk = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (blk >= total) return;  // the whole workgroup: the grid is 8 per
  const uint32_t tile = blk / a.ny, row = blk % a.ny;
  __shared__ uint32_t s_ngl, s_ndp;
  if (threadIdx.x == 0) {
    s_ngl = 0;
    s_ndp = 0;
  }
  __syncthreads();
  GaplessTask* gl_region = a.gapless_tasks + (size_t)blk * kRegion;
  DpTask* dp_region = a.dp_unsorted + (size_t)blk * kRegion;
  const uint32_t part = uni(threadIdx.x / GT);
  const uint32_t in_group = threadIdx.x % GT;
  const uint32_t wave_base = tile * QB + (in_group / 64u) * 64u * QL;
  const uint32_t lane = lane_id();
  const uint32_t c0 = row * kScreenCenters + part * CW;
  const uint32_t nc = c0 < a.n_centers ? min((uint32_t)CW, a.n_centers - c0) : 0u;

  CenterFields cf[CW];
  QueryFields qf[QL];
  load_ids<QL, CW>(a, c0, wave_base, lane, cf, qf);
  load_rest<QL, CW>(a, cf, qf);

  // sum |a - b| over the dense vectors, four buckets per v_sad_u8
  uint32_t sad[QL][CW];
#pragma unroll
  for (int j = 0; j < QL; ++j)
#pragma unroll
    for (int c = 0; c < CW; ++c) sad[j][c] = 0;
  if (a.use_kmers && nc > 0) {
    for (uint32_t w0 = 0; w0 < u.kwords; w0 += kScreenChunk) {
      uint32_t qw[QL][kScreenChunk];
#pragma unroll
      for (int j = 0; j < QL; ++j)
#pragma unroll
        for (int t = 0; t < kScreenChunk; ++t)
          qw[j][t] = u.kmer_il[(size_t)(w0 + t) * u.npad + qf[j].qq];
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const uint32_t* cw = u.kmer_row + (size_t)cf[c].ci * u.kwords + w0;
#pragma unroll
        for (int t = 0; t < kScreenChunk; ++t)
#pragma unroll
          for (int j = 0; j < QL; ++j) sad[j][c] = __builtin_amdgcn_sad_u8(qw[j][t], cw[t], sad[j][c]);
      }
    }
  }

  int32_t keq[QL][CW];
#pragma unroll
  for (int j = 0; j < QL; ++j)
#pragma unroll
    for (int c = 0; c < CW; ++c) keq[j][c] = -1;
  uint32_t n_skip = 0, n_shroud = 0;
  screen_pairs<QL, CW, 1>(a, qf, cf, sad, keq, nc, c0, &s_ngl, &s_ndp, gl_region, dp_region, n_skip,
                          n_shroud);
  screen_block_totals<GT * SPLIT / 64>(a, blk, n_skip, n_shroud, s_ngl, s_ndp);
}

// One center against every unique (the plain path and each compare of the
// partition loop).  A wavefront owns kS1Uniques consecutive uniques, its
// task region.  It reads each unique's whole k-mer vector in one coalesced
// row load (lane l holds words [l KL, l KL + KL)) with all of the region's
// loads in flight at once, and forms the per-lane sums of |a - b| against
// the center's words, held in registers.  A transposing reduction
// (reduce-scatter over lane bits 2, 3, 4, 5, then sums over bits 0 and 1)
// leaves unique j's total on lanes 4j .. 4j + 3, and lane 4j makes unique
// j's decision through screen_pairs: the region's task lists come out in
// unique order, like the multi-center screen's, for the same plan and
// scatter.
template <int KL>
__global__ __launch_bounds__(64 * kS1Waves) void screen1_kernel(ScreenDev a) {
  static_assert(kS1Uniques == 16, "the reduction maps 16 uniques onto lane bits 2..5");
  const UniqueDev& u = a.u;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / 64u;
  const uint32_t region = blockIdx.x * kS1Waves + wave;
  const uint32_t q0 = region * kS1Uniques;
  __shared__ uint32_t s_ngl[kS1Waves], s_ndp[kS1Waves];
  if (q0 >= u.n) return;  // wavefront-uniform; no workgroup barrier below
  if (lane == 0) {
    s_ngl[wave] = 0;
    s_ndp[wave] = 0;
  }
  // the region's k-mer rows first: they depend on nothing loaded
  const uint32_t w0 = lane * (uint32_t)KL;
  const bool mine = a.use_kmers && w0 < u.kwords;  // k < 5: lanes past the vector hold zeros
  uint32_t qw[kS1Uniques][KL];
#pragma unroll
  for (uint32_t j = 0; j < kS1Uniques; ++j) {
    const uint32_t q = min(q0 + j, u.n - 1);
    const uint32_t* qrow = u.kmer_row + (size_t)q * u.kwords + w0;
#pragma unroll
    for (int t = 0; t < KL; ++t) qw[j][t] = mine ? qrow[t] : 0u;
  }
  CenterFields cf[1];
  QueryFields qf[1];
  cf[0].ci = a.centers[0];
  const uint32_t j_me = lane >> 2;  // the unique this lane decides for (lanes 4j)
  qf[0].valid = (lane & 3u) == 0 && q0 + j_me < u.n;
  qf[0].qq = min(q0 + j_me, u.n - 1);
  load_rest<1, 1>(a, cf, qf);

  uint32_t v[kS1Uniques];
#pragma unroll
  for (uint32_t j = 0; j < kS1Uniques; ++j) v[j] = 0;
  if (a.use_kmers) {
    uint32_t cw[KL];
    const uint32_t* crow = u.kmer_row + (size_t)cf[0].ci * u.kwords + w0;
#pragma unroll
    for (int t = 0; t < KL; ++t) cw[t] = mine ? crow[t] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < kS1Uniques; ++j)
#pragma unroll
      for (int t = 0; t < KL; ++t) v[j] = __builtin_amdgcn_sad_u8(qw[j][t], cw[t], v[j]);
  }
  // reduce-scatter: at lane bit b, pairs (v[2i], v[2i+1]) become one vector
  // whose lanes with bit b clear sum v[2i] over lanes l and l ^ (1 << b)
  // and whose lanes with bit b set sum v[2i+1]; after bits 2..5, vector 0
  // holds unique (l >> 2) on lane l
#pragma unroll
  for (uint32_t step = 0; step < 4; ++step) {
    const uint32_t bit = 4u << step;
    const bool hi = (lane & bit) != 0;
    const uint32_t pairs = kS1Uniques >> (step + 1);
#pragma unroll
    for (uint32_t i = 0; i < kS1Uniques / 2; ++i) {
      if (i >= pairs) break;
      const uint32_t give = hi ? v[2 * i] : v[2 * i + 1];
      const uint32_t keep = hi ? v[2 * i + 1] : v[2 * i];
      v[i] = keep + (uint32_t)__shfl_xor((int)give, (int)bit);
    }
  }
  uint32_t tot = v[0];
  tot += (uint32_t)__shfl_xor((int)tot, 1);
  tot += (uint32_t)__shfl_xor((int)tot, 2);

  // The gapless dispatch's kord_dist (kmers.rs:314-328) for the pairs
  // that pass the screen with equal lengths: the whole wavefront reads their
  // k-mer order rows (one coalesced pass per 64 words, every needing pair's
  // loads in flight together) and the same transposing reduction puts pair
  // j's count of equal positions on lanes 4j.  Lanes that decide no pair
  // skip it (keq 0, unused); a pair whose vectors both hold a 255 count
  // takes its exact dot first and counts its own (keq -1).
  int32_t keq[1][1] = {{qf[0].valid ? -1 : 0}};
  if (a.use_kmers && a.gapless) {
    const QueryFields& Q = qf[0];
    const CenterFields& C = cf[0];
    const bool skip = a.greedy && (Q.reads > C.reads || Q.lock);
    const bool exact = Q.has255 && C.has255;
    const uint32_t dot = (Q.sum + C.sum - tot) >> 1;
    const bool shroud = !skip && dot < (Q.len <= C.len ? Q.thr : C.thr);
    const bool need = Q.valid && pair_in_pass(a, C.len, Q.qq) && !skip && !exact && !shroud &&
                      Q.len == C.len;
    const uint64_t nm = __builtin_amdgcn_ballot_w64(need);
    if (nm) {
      // every needing pair has the center's length
      const uint32_t klen = C.len - (u.k - 1);
      const uint32_t nw = (klen + 1) / 2;
      const uint32_t* crow = u.kord_row + (size_t)C.ci * u.owords;
      uint32_t part[kS1Uniques];
#pragma unroll
      for (uint32_t j = 0; j < kS1Uniques; ++j) part[j] = 0;
      for (uint32_t w0 = 0; w0 < nw; w0 += 64u) {
        const uint32_t w = w0 + lane;
        const bool inw = w < nw;
        const uint32_t cw = inw ? crow[w] : 0u;
        uint32_t x[kS1Uniques];
#pragma unroll
        for (uint32_t j = 0; j < kS1Uniques; ++j) {
          const uint32_t q = min(q0 + j, u.n - 1);
          x[j] = (inw && ((nm >> (4 * j)) & 1ull)) ? u.kord_row[(size_t)q * u.owords + w] ^ cw
                                                     : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t j = 0; j < kS1Uniques; ++j)
          part[j] += ((x[j] & 0xFFFFu) == 0 ? 1u : 0u) +
                     ((2 * w + 1 < klen && (x[j] >> 16) == 0) ? 1u : 0u);
      }
#pragma unroll
      for (uint32_t step = 0; step < 4; ++step) {
        const uint32_t bit = 4u << step;
        const bool hi = (lane & bit) != 0;
#pragma unroll
        for (uint32_t i = 0; i < kS1Uniques / 2; ++i) {
          if (i >= (kS1Uniques >> (step + 1))) break;
          const uint32_t give = hi ? part[2 * i] : part[2 * i + 1];
          const uint32_t keep = hi ? part[2 * i + 1] : part[2 * i];
          part[i] = keep + (uint32_t)__shfl_xor((int)give, (int)bit);
        }
      }
      uint32_t eq = part[0];
      eq += (uint32_t)__shfl_xor((int)eq, 1);
      eq += (uint32_t)__shfl_xor((int)eq, 2);
      if (need) keq[0][0] = (int32_t)eq;
    }
  }
  const uint32_t sad[1][1] = {{tot}};
  uint32_t n_skip = 0, n_shroud = 0;
  GaplessTask* gl_region = a.gapless_tasks + (size_t)region * kS1Uniques;
  DpTask* dp_region = a.dp_unsorted + (size_t)region * kS1Uniques;
  screen_pairs<1, 1, 16>(a, qf, cf, sad, keq, 1u, 0u, &s_ngl[wave], &s_ndp[wave], gl_region,
                         dp_region, n_skip, n_shroud);
  const uint32_t t_skip = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(n_skip != 0));
  const uint32_t t_shroud =
      (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(n_shroud != 0));
  if (lane == 0) {
    a.partials[4 * region] = t_skip;
    a.partials[4 * region + 1] = t_shroud;
    a.partials[4 * region + 2] = s_ngl[wave];
    a.partials[4 * region + 3] = s_ndp[wave];
  }
}

void screen_thresholds(double cutoff, uint32_t nmax, uint32_t* out) {
  out[0] = 0;
  for (uint32_t n = 1; n <= nmax; ++n) {
    // the predicate 1 - d/n <= cutoff is monotone in d: bisect for its
    // first true d in 0 .. n + 1 (n + 1: never true)
    uint32_t lo = 0, hi = n + 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      const double kdist = 1.0 - (double)mid / (double)n;
      if (!(kdist > cutoff)) hi = mid; else lo = mid + 1;
    }
    out[n] = lo;
  }
}

uint32_t screen_blocks(uint32_t n_uniques, uint32_t n_centers, uint32_t qblock) {
  return ((n_uniques + qblock - 1) / qblock) * ((n_centers + kScreenCenters - 1) / kScreenCenters);
}

__global__ __launch_bounds__(kPairBlock) void pair_kernel(ScreenDev a, const uint32_t* want) {
  __shared__ uint32_t s_ngl, s_ndp;
  if (threadIdx.x == 0) {
    s_ngl = 0;
    s_ndp = 0;
  }
  __syncthreads();
  const UniqueDev& u = a.u;
  const uint32_t q = blockIdx.x * kPairBlock + threadIdx.x;
  const uint32_t ci = q < u.n ? want[q] : 0xFFFFFFFFu;
  bool live = ci < u.n && pair_in_pass(a, u.len[ci], q);
  // raw_align_with_buf (nwalign.rs:1244-1310): without the k-mer screen
  // kdist is 0.0 and kodist -1.0, so only band 0 takes the gapless
  // alignment (nwalign.rs:1324); with it, the pair's own distances decide
  double kdist = 0.0, kodist = -1.0;
  if (a.use_kmers && live) {
    const uint32_t k = u.k, lq = u.len[q], lc = u.len[ci];
    bool ovf = false;
    uint32_t dot = 0;
    for (uint32_t w = 0; w < u.kwords; ++w) {  // kmer_dist8, kmers.rs:254-269
      const uint32_t x = u.kmer_il[(size_t)w * u.npad + q];
      const uint32_t y = u.kmer_row[(size_t)ci * u.kwords + w];
      for (int b = 0; b < 4; ++b) {
        const uint32_t m = min((x >> (8 * b)) & 255u, (y >> (8 * b)) & 255u);
        ovf |= m == 255;
        dot += m;
      }
    }
    if (ovf) dot = u16_min_dot(u, q, ci);  // nwalign.rs:1270-1278
    const uint32_t mn = lq < lc ? lq : lc;
    kdist = 1.0 - (double)dot / (double)(mn - k + 1);
    if (a.gapless && lq == lc) {  // kord_dist, kmers.rs:314-328
      const uint32_t klen = lq - (k - 1);
      uint32_t eq = 0;
      for (uint32_t s2 = 0; s2 < klen; ++s2) eq += kord_il_at(u, q, s2) == kord_row_at(u, ci, s2);
      kodist = 1.0 - (double)eq / (double)klen;
    }
    live = !(kdist > a.kdist_cutoff);  // shrouded: no alignment
  }
  const bool gapless_route = a.band == 0 || (a.gapless && fabs(kodist - kdist) < DBL_EPSILON);
  const bool to_gl = live && gapless_route;
  const bool to_dp = live && !gapless_route;
  GaplessTask* gl_region = a.gapless_tasks + (size_t)blockIdx.x * kPairBlock;
  DpTask* dp_region = a.dp_unsorted + (size_t)blockIdx.x * kPairBlock;
  const uint32_t gpos = wave_append(&s_ngl, to_gl);
  if (to_gl) {
    GaplessTask t;
    t.c_idx = ci;
    t.u_idx = q;
    t.out_idx = q;
    t._pad = 0;
    gl_region[gpos] = t;
  }
  const uint32_t dpos = wave_append(&s_ndp, to_dp);
  uint32_t bin = 0;
  if (to_dp) {
    const uint32_t lc = u.len[ci], lq = u.len[q];
    bin = (lc - (uint32_t)a.bins.lmin) * (uint32_t)a.bins.range + (lq - (uint32_t)a.bins.lmin);
    DpTask t;
    t.s1_off = u.off[ci];
    t.s2_off = u.off[q];
    t.out_idx = q;
    t.bin = bin;
    t.u_idx = q;
    t.c_idx = ci;
    dp_region[dpos] = t;
  }
  (void)wave_add_by_key(a.bin_count, bin, to_dp);
  __syncthreads();
  if (threadIdx.x == 0) {
    a.partials[4 * blockIdx.x] = 0;
    a.partials[4 * blockIdx.x + 1] = 0;
    a.partials[4 * blockIdx.x + 2] = s_ngl;
    a.partials[4 * blockIdx.x + 3] = s_ndp;
  }
}

hipError_t launch_pair_screen(const ScreenArgs& s, const uint32_t* want, hipStream_t st) {
  if (want == nullptr) return hipErrorInvalidValue;
  ScreenDev d;
  std::memset(&d, 0, sizeof(d));
  d.u = *s.u;
  d.use_kmers = s.use_kmers;
  d.gapless = s.gapless;
  d.kdist_cutoff = s.kdist_cutoff;
  d.band = s.band;
  d.bins = s.bins;
  d.gapless_tasks = s.gapless_tasks;
  d.dp_unsorted = s.dp_unsorted;
  d.bin_count = s.bin_count;
  d.partials = s.partials;
  d.ctr = s.ctr;
  d.long_len = s.long_len;
  d.sr_lo = s.sr_lo;
  d.sr_hi = s.sr_hi;
  d.lr_lo = s.lr_lo;
  d.lr_hi = s.lr_hi;
  hipLaunchKernelGGL(pair_kernel, dim3(pair_blocks(d.u.n)), dim3(kPairBlock), 0, st, d, want);
  return hipGetLastError();
}

template <int QB, int QL>
static void launch_qb(const ScreenDev& d, dim3 grid, int split, hipStream_t st) {
  switch (split) {
    case 1: hipLaunchKernelGGL((screen_kernel<QB, 1, QL>), grid, dim3(QB / QL), 0, st, d); break;
    case 2: hipLaunchKernelGGL((screen_kernel<QB, 2, QL>), grid, dim3(QB / QL * 2), 0, st, d); break;
    default: hipLaunchKernelGGL((screen_kernel<QB, 4, QL>), grid, dim3(QB / QL * 4), 0, st, d); break;
  }
}

bool screen1_fits(uint32_t n_centers, int use_kmers, uint32_t kwords) {
  return n_centers == 1 && (!use_kmers || kwords <= 256);
}
uint32_t screen1_regions(uint32_t n_uniques) { return (n_uniques + kS1Uniques - 1) / kS1Uniques; }

hipError_t launch_screen(const ScreenArgs& s, hipStream_t st) {
  ScreenDev d;
  d.u = *s.u;
  d.centers = s.centers;
  d.n_centers = s.n_centers;
  d.lock = s.lock;
  d.greedy = s.greedy;
  d.use_kmers = s.use_kmers;
  d.gapless = s.gapless;
  d.band = s.band;
  d.kdist_cutoff = s.kdist_cutoff;
  d.bins = s.bins;
  d.out_lambda = s.out_lambda;
  d.out_hamming = s.out_hamming;
  d.out_skipped = s.out_skipped;
  d.gapless_tasks = s.gapless_tasks;
  d.dp_unsorted = s.dp_unsorted;
  d.bin_count = s.bin_count;
  d.partials = s.partials;
  d.ctr = s.ctr;
  d.dstar = s.dstar;
  d.long_len = s.long_len;
  d.sr_lo = s.sr_lo;
  d.sr_hi = s.sr_hi;
  d.lr_lo = s.lr_lo;
  d.lr_hi = s.lr_hi;
  if (s.split != 1 && s.split != 2 && s.split != 4) return hipErrorInvalidValue;
  if (s.use_kmers && s.dstar == nullptr) return hipErrorInvalidValue;
  if (s.single && screen1_fits(s.n_centers, s.use_kmers, d.u.kwords)) {
    const uint32_t blocks = (screen1_regions(d.u.n) + kS1Waves - 1) / kS1Waves;
    if (d.u.kwords > 64)
      hipLaunchKernelGGL(screen1_kernel<4>, dim3(blocks), dim3(64 * kS1Waves), 0, st, d);
    else
      hipLaunchKernelGGL(screen1_kernel<1>, dim3(blocks), dim3(64 * kS1Waves), 0, st, d);
    return hipGetLastError();
  }
  d.nx = (d.u.n + s.qblock - 1) / s.qblock;
  d.ny = (s.n_centers + kScreenCenters - 1) / kScreenCenters;
  const dim3 grid(8u * ((d.nx * d.ny + 7u) / 8u));
  const int key = (int)s.qblock * 4 + s.per_lane;
  switch (key) {
    case 64 * 4 + 1: launch_qb<64, 1>(d, grid, s.split, st); break;
    case 128 * 4 + 1: launch_qb<128, 1>(d, grid, s.split, st); break;
    case 256 * 4 + 1: launch_qb<256, 1>(d, grid, s.split, st); break;
    case 128 * 4 + 2: launch_qb<128, 2>(d, grid, s.split, st); break;
    case 256 * 4 + 2: launch_qb<256, 2>(d, grid, s.split, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace d2g
