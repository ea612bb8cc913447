// nw_band.hip — the scalar ends-free aligners of dada2-rs on gfx950,
// bit-exact: align_endsfree_with_buf (nwalign.rs:336-417) and
// align_endsfree_homo_with_buf (nwalign.rs:435-528), i32 arithmetic.
//
// Reference semantics restated (s1 = the center, s2 = the query; no swap):
//   * matrix (len1+1) x (len2+1), d zeroed, row 0 pointers 2 (left), column 0
//     pointers 3 (up) (nwalign.rs:456-466);
//   * row i computes j in [max(1, i-lband), min(len2, i+rband)], the band
//     widened by the length difference (band_adjust, nwalign.rs:276-284);
//     band < 0 computes every j;
//   * when band < len1 or band < len2, d = -9999 at j = i-lband-1 and
//     j = i+rband+1 for i = 0..len1 (fill_band_sentinels, nwalign.rs:287-308),
//     pointers there keep their initial value;
//   * left = d[i][j-1] + (i == len1 ? 0 : gap), up = d[i-1][j] + (j == len2 ?
//     0 : gap), where the homopolymer variant uses homo_gap instead of gap when
//     the base opposite the gap (s2[j-1] for left, s1[i-1] for up) lies in a
//     run of >= 3 identical bytes (homopolymer_mask_into, nwalign.rs:211-227);
//   * up if up >= diag && up >= left, else left if left >= diag, else diag;
//   * traceback from (len1, len2) (traceback_into, nwalign.rs:234-273).
//
// Mapping.  Cells of one anti-diagonal t = i + j are independent.  Band
// diagonals d = j - i run over [-lband-1, rband+1] (sentinels included),
// indexed D = d + lband + 1 in [0, dmax].  On anti-diagonal t only diagonals
// with D = t + lband + 1 (mod 2) hold cells, so lane c of a pair's group
// owns the 2K diagonals D in [2Kc, 2Kc + 2K) and computes K cells per step
// (K = 1, 2 or 4: the smallest that fits the band into one wavefront).  The
// up neighbour (D+1) and left neighbour (D-1) of the previous step are the
// lane's own cells except at the lane's ends, where one DPP shift brings the
// neighbour lane's; the diagonal neighbour is the cell's own value two steps
// back.  Every lane classifies its cells (interior, row 0 / column 0,
// sentinel, outside the matrix) from (i, j).
//
// Pointers go to a per-workgroup history slot (2 bits per cell, K words per
// 16 steps); a diagonal step over a mismatch is stored as 0 and told apart
// from an unset pointer by the cell's class.  The traceback sweeps the
// steps downwards, reading the path cell's word from its owner lane with
// ds_bpermute, and writes al2subs records (compare) or alignment strings.
#include <algorithm>

#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

constexpr int kBandSentinel = -9999;  // nwalign.rs:287-308

struct BandDev {
  BinSpace bins;
  const uint32_t* bin_start;
  const uint32_t* wave_start;
  const DpTask* tasks;
  const uint8_t* pool;
  int match, mismatch, gap, homo_gap;
  Counters* ctr;
  int emit;
  uint32_t* sub_pool;
  size_t sub_capacity;
  uint32_t sub_stride;
  uint32_t* sub_cnt;
  uint32_t* hist;
  size_t hist_stride;
  uint8_t* al0;
  uint8_t* al1;
  const uint64_t* al_off;
  uint32_t* al_len;
};

// Per-bin cell counts for the plan's GCUPS total (BinInfo.cells).
__global__ void band_info_kernel(BinSpace bs, BinInfo* info) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bs.nbins) return;
  int l1, l2;
  bin_lengths(bs, (uint32_t)b, &l1, &l2);
  const BandGeom g = band_geom(l1, l2, bs.band, bs.tile);
  BinInfo bi;
  bi.steady_lo = 0;
  bi.steady_hi = 0;
  bi.cells = band_cells(g, bs.band);
  bi.masked = 0;
  info[b] = bi;
}

hipError_t launch_band_info(const BinSpace& bs, BinInfo* info, hipStream_t s) {
  const int threads = 64;
  const int blocks = (bs.nbins + threads - 1) / threads;
  hipLaunchKernelGGL(band_info_kernel, dim3(blocks), dim3(threads), 0, s, bs, info);
  return hipGetLastError();
}

// Per-wavefront constants.
struct BandK {
  int len1, len2, lband, dmax;
  int match, mismatch, gap, homo;
};
// Per-lane state for K cells per step: cell m of the lane is diagonal
// D = 2Kc + 2m + parity; bit m of the masks flags D <= dmax (ok) and the
// sentinel diagonals (sent); prev / prev2 hold the scores of the last two
// steps.
template <int K>
struct BandLane {
  int d0;  // d of D = 2Kc
  uint32_t okE, okO, sentE, sentO;
  int prev2[K], prev[K];
};

// One anti-diagonal t for every pair of the wavefront; PAR = parity of the
// lane's diagonals on this step.  Writes the K cells' pointer codes.
template <bool HOMO, int PAR, int K>
__device__ __forceinline__ void band_step(const int t, BandLane<K>& L, const BandK Kc,
                                          const uint16_t* q1, const uint16_t* q2, uint32_t* pout) {
  // the one neighbour value from another lane: D+1 of cell K-1 (odd steps)
  // or D-1 of cell 0 (even steps)
  const int ext = PAR ? from_upper(L.prev[0], 0) : from_lower(L.prev[K - 1], 0);
  int nv[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const int d = L.d0 + 2 * m + PAR;
    const int i = (t - d) >> 1;  // t - d is even on this step
    const int j = i + d;
    const bool ok = (((PAR ? L.okO : L.okE) >> m) & 1u) != 0;
    const bool inm = ok && (unsigned)i <= (unsigned)Kc.len1 && (unsigned)j <= (unsigned)Kc.len2;
    const int ii = min(max(i - 1, 0), Kc.len1 - 1), jj = min(max(j - 1, 0), Kc.len2 - 1);
    const uint32_t va = q1[ii], vb = q2[jj];
    const bool differ = (va & 0xFFu) != (vb & 0xFFu);
    const int diag = L.prev2[m] + (differ ? Kc.mismatch : Kc.match);
    const int lraw = PAR ? L.prev[m] : (m >= 1 ? L.prev[m > 0 ? m - 1 : 0] : ext);
    const int uraw = PAR ? (m + 1 < K ? L.prev[m + 1 < K ? m + 1 : 0] : ext) : L.prev[m];
    int gl = Kc.gap, gu = Kc.gap;
    if (HOMO) {
      gl = (vb >> 8) != 0 ? Kc.homo : Kc.gap;
      gu = (va >> 8) != 0 ? Kc.homo : Kc.gap;
    }
    gl = (i == Kc.len1) ? 0 : gl;  // free end gaps on the last row / column
    gu = (j == Kc.len2) ? 0 : gu;
    const int left = lraw + gl, up = uraw + gu;
    const bool upw = up >= diag && up >= left;
    const bool lw = left >= diag;
    int e = upw ? up : (lw ? left : diag);
    int p = upw ? 3 : (lw ? 2 : (differ ? 0 : 1));
    const bool sent = (((PAR ? L.sentO : L.sentE) >> m) & 1u) != 0;
    const bool interior = inm && i >= 1 && j >= 1 && !sent;
    const bool edge = inm && (i == 0 || j == 0);
    e = interior ? e : (sent ? kBandSentinel : 0);
    p = interior ? p : (edge ? (i == 0 ? 2 : 3) : 0);
    nv[m] = e;
    pout[m] = (uint32_t)p;
  }
#pragma unroll
  for (int m = 0; m < K; ++m) {
    L.prev2[m] = L.prev[m];
    L.prev[m] = nv[m];
  }
}

// Steps [16 blk, 16 blk + 16) ∩ [0, last]; P0 = parity of D at even steps.
// History word m of the block holds cell m's pointers.
template <bool HOMO, int P0, int K>
__device__ __forceinline__ void band_block(const int blk, const int last, BandLane<K>& L,
                                           const BandK Kc, const uint16_t* q1, const uint16_t* q2,
                                           uint32_t* hist_p, const int lane) {
  const int t0 = blk * 16;
  uint32_t hp[K];
#pragma unroll
  for (int m = 0; m < K; ++m) hp[m] = 0;
  uint32_t p[K];
  if (t0 + 15 <= last) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & 1)
        band_step<HOMO, 1 - P0, K>(t0 + k, L, Kc, q1, q2, p);
      else
        band_step<HOMO, P0, K>(t0 + k, L, Kc, q1, q2, p);
#pragma unroll
      for (int m = 0; m < K; ++m) hp[m] = hp[m] * 4u + p[m];
    }
  } else {
    int k = 0;
    for (; t0 + k <= last; ++k) {
      if (k & 1)
        band_step<HOMO, 1 - P0, K>(t0 + k, L, Kc, q1, q2, p);
      else
        band_step<HOMO, P0, K>(t0 + k, L, Kc, q1, q2, p);
#pragma unroll
      for (int m = 0; m < K; ++m) hp[m] = hp[m] * 4u + p[m];
    }
#pragma unroll
    for (int m = 0; m < K; ++m) hp[m] <<= 2 * (16 - k);  // step t0 + k' at bits 2 (15 - k')
  }
#pragma unroll
  for (int m = 0; m < K; ++m) hist_p[(blk * K + m) * 64 + lane] = hp[m];
}

// hp = 2 hp + (1 on the lanes of mask): one add with the mask as carry-in.
__device__ __forceinline__ uint32_t band_shift_in(uint32_t hp, uint64_t mask) {
  uint32_t r;
  uint64_t carry_out;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(hp), "s"(mask));
  return r;
}

// Per-lane constants of the steady steps: for each step parity e (of k) and
// cell m, the lanes whose cell is a sentinel or past dmax (special), and the
// value such a cell holds (-9999 or 0).
template <int K>
struct BandSpecial {
  uint64_t mask[2][K];
  int cval[2][K];
};

// A steady step (band_step where every cell that is neither a sentinel nor
// past dmax is interior and off the last row and column): the packed rows
// give base and homopolymer mask in one read at a fixed offset from the
// block's base, and the special cells are sent to their constant.  The
// pointer's two bits are shifted into hp[m].
// Scores of the steady steps, held in VGPRs (the selects take one
// scalar operand at most).
struct BandV {
  int match, mismatch, gap, homo;
};
__device__ __forceinline__ int in_vgpr(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <bool HOMO, int PAR, int E, int K>
__device__ __forceinline__ void band_steady_step(const int k, BandLane<K>& L, const BandV Kc,
                                                 const uint16_t* const (&q1b)[2][K],
                                                 const uint16_t* const (&q2b)[2][K],
                                                 const BandSpecial<K>& S, uint32_t (&hp)[K]) {
  const int ext = PAR ? from_upper(L.prev[0], 0) : from_lower(L.prev[K - 1], 0);
  int nv[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const uint32_t va = q1b[E][m][k >> 1], vb = q2b[E][m][k >> 1];
    const bool eq = (va & 0xFFu) == (vb & 0xFFu);
    const int diag = L.prev2[m] + (eq ? Kc.match : Kc.mismatch);
    const int lraw = PAR ? L.prev[m] : (m >= 1 ? L.prev[m > 0 ? m - 1 : 0] : ext);
    const int uraw = PAR ? (m + 1 < K ? L.prev[m + 1 < K ? m + 1 : 0] : ext) : L.prev[m];
    const int gl = (HOMO && (vb >> 8) != 0) ? Kc.homo : Kc.gap;
    const int gu = (HOMO && (va >> 8) != 0) ? Kc.homo : Kc.gap;
    const int left = lraw + gl, up = uraw + gu;
    const bool upw = up >= diag && up >= left;
    const bool lw = left >= diag;
    const int e = upw ? up : (lw ? left : diag);
    const bool sp = __builtin_amdgcn_inverse_ballot_w64(S.mask[E][m]);
    nv[m] = sp ? S.cval[E][m] : e;
    const uint64_t mup = __builtin_amdgcn_ballot_w64(upw);
    const uint64_t mlw = __builtin_amdgcn_ballot_w64(lw);
    const uint64_t meq = __builtin_amdgcn_ballot_w64(eq);
    hp[m] = band_shift_in(hp[m], (mup | mlw) & ~S.mask[E][m]);
    hp[m] = band_shift_in(hp[m], (mup | (~mlw & meq)) & ~S.mask[E][m]);
  }
#pragma unroll
  for (int m = 0; m < K; ++m) {
    L.prev2[m] = L.prev[m];
    L.prev[m] = nv[m];
  }
}

// Steps [16 blk, 16 blk + 16), all steady.  Cell m's row index on the
// block's steps of parity e advances by one every two steps from the base
// taken at the block's first step of that parity; bases of special cells
// are clamped to the padded rows (their reads are discarded).
// Cell m's bases are cell 0's shifted by (-m, +m) (diagonal d + 2m), so
// the reads of neighbouring cells coincide; cell 0's bases are clamped to
// the padded rows far enough that every cell's reads stay inside them.
template <bool HOMO, int P0, int K>
__device__ __forceinline__ void band_block_steady(const int blk, BandLane<K>& L, const BandV Kc,
                                                  const int len1, const int len2,
                                                  const uint16_t* q1, const uint16_t* q2,
                                                  const BandSpecial<K>& S, uint32_t* hist_p,
                                                  const int lane) {
  static_assert(K <= 2, "steady steps hold K <= 2 cells per lane");
  const int t0 = blk * 16;
  const uint16_t* q1b[2][K];
  const uint16_t* q2b[2][K];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int d = L.d0 + (e ? 1 - P0 : P0);
    const int ii = ((t0 + e - d) >> 1) - 1, jj = ii + d;
    const uint16_t* b1 = q1 + min(max(ii, -7), len1 - 1);
    const uint16_t* b2 = q2 + min(max(jj, -8), len2 - 2);
#pragma unroll
    for (int m = 0; m < K; ++m) {
      q1b[e][m] = b1 - m;
      q2b[e][m] = b2 + m;
    }
  }
  uint32_t hp[K];
#pragma unroll
  for (int m = 0; m < K; ++m) hp[m] = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (k & 1)
      band_steady_step<HOMO, 1 - P0, 1, K>(k, L, Kc, q1b, q2b, S, hp);
    else
      band_steady_step<HOMO, P0, 0, K>(k, L, Kc, q1b, q2b, S, hp);
  }
#pragma unroll
  for (int m = 0; m < K; ++m) hist_p[(blk * K + m) * 64 + lane] = hp[m];
}

// Per-wave constants of the traceback.
struct BandBw {
  int gpos, c, lband, len1, len2, dmax;
  bool sentinels, real, sub_ok;
  size_t my_sub;
  const uint16_t* q1;
  const uint16_t* q2;
  uint64_t ao;
};
struct BandPath {
  int rc;  // step of the path cell; -1 once ended
  int D;   // its diagonal index
  int ns, nt;
  bool bad;
};

// One step t of the traceback (traceback_into, nwalign.rs:234-273) for
// every pair of the wavefront.
template <bool EMIT, int K>
__device__ __forceinline__ void band_bw_step(const int t, const int k, const uint32_t* hp,
                                             BandPath& ps, const BandBw& B, const BandDev& a) {
  const int owner = (int)((unsigned)ps.D / (unsigned)(2 * K));
  const int mi = (ps.D - owner * 2 * K) >> 1;
  uint32_t w = 0;
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute((B.gpos + owner) * 4, (int)hp[m]);
    w = mi == m ? x : w;
  }
  int pp = (int)__builtin_amdgcn_ubfe(w, (uint32_t)(30 - 2 * k), 2u);
  const bool on = ps.rc == t;
  const int d = ps.D - B.lband - 1;
  const int i = (t - d) >> 1, j = i + d;
  // pointer 0: a mismatching diagonal on an interior cell, else unset
  if (__builtin_amdgcn_ballot_w64(on && pp == 0)) {
    const bool sentd = B.sentinels && (ps.D == 0 || ps.D == B.dmax);
    const bool interior = i >= 1 && j >= 1 && i <= B.len1 && j <= B.len2 && !sentd;
    const bool sub = on && pp == 0 && interior;
    if (!EMIT && sub && B.c == owner && B.real && B.sub_ok)
      a.sub_pool[B.my_sub + ps.ns] = ((uint32_t)(j - 1) << 16) | (uint32_t)(i - 1);
    ps.ns += sub ? 1 : 0;
    pp = sub ? 1 : pp;
  }
  if (EMIT) {
    if (on && pp != 0 && B.c == owner && B.real) {
      a.al0[B.ao + ps.nt] = pp == 2 ? (uint8_t)'-' : (uint8_t)B.q1[max(i - 1, 0)];
      a.al1[B.ao + ps.nt] = pp == 3 ? (uint8_t)'-' : (uint8_t)B.q2[max(j - 1, 0)];
    }
    ps.nt += (on && pp != 0) ? 1 : 0;
  }
  ps.bad = ps.bad || (on && pp == 0);
  // diag: (i-1, j-1), same diagonal two steps back; left: (i, j-1), D-1;
  // up: (i-1, j), D+1
  ps.D += on ? (pp == 2 ? -1 : (pp == 3 ? 1 : 0)) : 0;
  ps.rc = on ? (pp == 1 ? t - 2 : (pp == 0 ? -1 : t - 1)) : ps.rc;
}

template <bool EMIT, int K>
__device__ __forceinline__ void band_backward(const int last, const uint32_t* hist_p,
                                              const int lane, BandPath& ps, const BandBw& B,
                                              const BandDev& a) {
  const int top = last >> 4;
  uint32_t hp[K], hn[K];
#pragma unroll
  for (int m = 0; m < K; ++m) {
    hp[m] = hist_p[(top * K + m) * 64 + lane];
    hn[m] = top >= 1 ? hist_p[((top - 1) * K + m) * 64 + lane] : 0u;
  }
  const int tlo = top * 16 > 1 ? top * 16 : 1;
  for (int t = last; t >= tlo; --t) band_bw_step<EMIT, K>(t, t & 15, hp, ps, B, a);
  for (int b = top - 1; b >= 1; --b) {
#pragma unroll
    for (int m = 0; m < K; ++m) {
      hp[m] = hn[m];
      hn[m] = hist_p[((b - 1) * K + m) * 64 + lane];
    }
#pragma unroll
    for (int k = 15; k >= 0; --k) band_bw_step<EMIT, K>(b * 16 + k, k, hp, ps, B, a);
  }
  if (top >= 1) {
#pragma unroll
    for (int m = 0; m < K; ++m) hp[m] = hn[m];
    for (int t = 15; t >= 1; --t) band_bw_step<EMIT, K>(t, t, hp, ps, B, a);
  }
}

// Path state of the one-pair traceback (wavefront-uniform).  Substitution
// records wait in two 64-entry registers (record n in lane n & 63 of
// sbuf[(n >> 6) & 1]) and are written a full register at a time between
// blocks; `flushed` counts the registers written.
struct BandSolo {
  int rc, D, ns, nt, flushed;
  bool bad;
  uint32_t sbuf[2];
};

// s_waitcnt vmcnt(0): a store written here is complete before the next
// history block is used, so the loads ahead never wait behind stores.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0xF70); }

// The path's moves inside history block blk (steps [16 blk, 16 blk + 16)),
// whose words are hb.  Same moves, records and end conditions as
// band_bw_step; returns false once the path has ended.
//
// A diagonal move keeps the diagonal D and goes two steps back, so within
// one block the pointers of steps t, t-2, t-4, ... of D sit 4 bits apart in
// the owner's word: shifted to step t, a run of diagonal-over-match
// pointers (1) is the count of trailing zero fields of
// (x ^ 0x11111111) & 0x33333333.  Each iteration takes the whole run and
// then at most one other move, so a block costs one iteration per
// mismatch or gap on the path.  Steps below the block read as zero fields,
// which ends a run at the block's edge; row 0 and column 0 hold pointers 2
// and 3, so no run reaches step 1 or 0.  In alignment-output mode a run's n
// columns (the bases q1[i-1-l], q2[j-1-l], l < n) are written by lanes
// 0 .. n-1.
template <bool EMIT, int K>
__device__ __forceinline__ bool band_solo_block(const int blk, const uint32_t (&hb)[K],
                                                BandSolo& st, const bool writer, const int lane,
                                                const BandBw& B, const BandDev& a) {
  const int lo = imax(blk * 16, 1);
  while (st.rc >= lo) {
    const int owner = (int)((unsigned)st.D / (unsigned)(2 * K));
    const int mi = (st.D - owner * 2 * K) >> 1;
    uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)hb[0], owner);
#pragma unroll
    for (int m = 1; m < K; ++m) {
      const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)hb[m], owner);
      w = mi == m ? x : w;
    }
    const uint32_t x = w >> ((30 - 2 * (st.rc & 15)) & 31);
    const uint32_t rm = (x ^ 0x11111111u) & 0x33333333u;
    const int n = (int)((__builtin_ctz(rm | 0x80000000u) + 1u) >> 2);
    const int d = st.D - B.lband - 1;
    if (EMIT && n > 0) {
      const int i0 = (st.rc - d) >> 1, j0 = i0 + d;
      if (writer && lane < n) {
        a.al0[B.ao + st.nt + lane] = (uint8_t)B.q1[i0 - 1 - lane];
        a.al1[B.ao + st.nt + lane] = (uint8_t)B.q2[j0 - 1 - lane];
      }
      st.nt += n;
    }
    st.rc -= 2 * n;
    if (st.rc < lo) break;
    int pp = (int)((x >> ((4 * n) & 31)) & 3u);  // 0, 2 or 3: a run stops at non-1
    const int i = (st.rc - d) >> 1, j = i + d;
    if (pp == 0) {  // a mismatching diagonal on an interior cell, else unset
      const bool sentd = B.sentinels && (st.D == 0 || st.D == B.dmax);
      const bool interior = i >= 1 && j >= 1 && i <= B.len1 && j <= B.len2 && !sentd;
      if (!EMIT && interior) {
        const uint32_t rec = ((uint32_t)(j - 1) << 16) | (uint32_t)(i - 1);
        const bool mine = lane == (st.ns & 63);
        if ((st.ns >> 6) & 1)
          st.sbuf[1] = mine ? rec : st.sbuf[1];
        else
          st.sbuf[0] = mine ? rec : st.sbuf[0];
      }
      st.ns += interior ? 1 : 0;
      st.bad = !interior;
      pp = interior ? 1 : 0;
    }
    if (EMIT && pp != 0) {
      if (writer && lane == 0) {
        a.al0[B.ao + st.nt] = pp == 2 ? (uint8_t)'-' : (uint8_t)B.q1[max(i - 1, 0)];
        a.al1[B.ao + st.nt] = pp == 3 ? (uint8_t)'-' : (uint8_t)B.q2[max(j - 1, 0)];
      }
      ++st.nt;
    }
    // diag: D, two steps back; left (2): D - 1; up (3): D + 1; unset: end
    st.D += (pp >> 1) * (2 * (pp & 1) - 1);
    st.rc = pp == 0 ? -1 : st.rc - 1 - (pp == 1 ? 1 : 0);
  }
  return st.rc >= 1;
}

// Writes the register of 64 substitution records filled by the last block
// (a block adds at most 16 records, so at most one).
template <bool EMIT>
__device__ __forceinline__ void band_solo_flush(BandSolo& st, const bool store, const int lane,
                                                const BandBw& B, const BandDev& a) {
  if (!EMIT && (st.ns >> 6) > st.flushed) {
    if (store)
      a.sub_pool[B.my_sub + st.flushed * 64 + lane] = (st.flushed & 1) ? st.sbuf[1] : st.sbuf[0];
    ++st.flushed;
    wait_vm_all();
  }
}

// The traceback of a wavefront that holds one pair: the path is
// wavefront-uniform, so it advances one cell per iteration (a diagonal move
// skips a step) with its state in scalar registers, and the path cell's
// pointer word comes from its owner lane by v_readlane.  A move takes at
// most two steps, so the path visits every block from the top one down.
// History words come in two groups of kG blocks: group B's load is issued
// after group A's first block and A's next load after A's last block, so
// each group's load is issued kG - 1 or more blocks before its first use and
// no load is younger than that where the loop starts over.
template <bool EMIT, int K>
__device__ __forceinline__ void band_backward_solo(const int last, const uint32_t* hist_p,
                                                   const int lane, BandPath& ps, const BandBw& B,
                                                   const BandDev& a) {
  constexpr int kG = K == 1 ? 8 : (K == 2 ? 4 : 2);
  int b = last >> 4;
  uint32_t ga[kG][K], gb[kG][K];
#pragma unroll
  for (int s = 0; s < kG; ++s)
#pragma unroll
    for (int m = 0; m < K; ++m) ga[s][m] = hist_p[(imax(b - s, 0) * K + m) * 64 + lane];
  BandSolo st;
  st.rc = uni(ps.rc);
  st.D = uni(ps.D);
  st.ns = 0;
  st.nt = 0;
  st.flushed = 0;
  st.bad = false;
  st.sbuf[0] = st.sbuf[1] = 0;
  // the wavefront holds one pair, on lanes 0 .. lanes-1; record n sits in
  // lane n & 63 whatever the pair's width, so every lane writes for it
  const bool writer = __builtin_amdgcn_readfirstlane(B.real ? 1 : 0) != 0;
  const bool store = !EMIT && B.sub_ok && writer;
  bool more = st.rc >= 1;
  while (more) {
#pragma unroll
    for (int s = 0; s < kG; ++s) {
      more = band_solo_block<EMIT, K>(b - s, ga[s], st, writer, lane, B, a);
      if (s == 0) {
#pragma unroll
        for (int u = 0; u < kG; ++u)
#pragma unroll
          for (int m = 0; m < K; ++m)
            gb[u][m] = hist_p[(imax(b - kG - u, 0) * K + m) * 64 + lane];
      }
      band_solo_flush<EMIT>(st, store, lane, B, a);
      if (!more) break;
    }
    if (!more) break;
#pragma unroll
    for (int u = 0; u < kG; ++u)
#pragma unroll
      for (int m = 0; m < K; ++m)
        ga[u][m] = hist_p[(imax(b - 2 * kG - u, 0) * K + m) * 64 + lane];
#pragma unroll
    for (int s = 0; s < kG; ++s) {
      more = band_solo_block<EMIT, K>(b - kG - s, gb[s], st, writer, lane, B, a);
      band_solo_flush<EMIT>(st, store, lane, B, a);
      if (!more) break;
    }
    b -= 2 * kG;
  }
  if (!EMIT) {
    for (; st.flushed * 64 < st.ns; ++st.flushed)
      if (store && lane < st.ns - st.flushed * 64)
        a.sub_pool[B.my_sub + st.flushed * 64 + lane] = (st.flushed & 1) ? st.sbuf[1] : st.sbuf[0];
  }
  ps.rc = st.rc;
  ps.D = st.D;
  ps.ns = st.ns;
  ps.nt = st.nt;
  ps.bad = st.bad;
}

// Forward, traceback and outputs of one wavefront whose bin has K cells
// per lane.
template <bool HOMO, int K>
__device__ __forceinline__ void band_pairs(const BandDev& a, const BandGeom& geo, const int c,
                                           const int lane, const int g_raw, const bool real,
                                           const DpTask& task, const uint32_t ti,
                                           const uint16_t* q1, const uint16_t* q2,
                                           uint32_t* hist_p, const bool solo) {
  const int len1 = geo.len1, len2 = geo.len2, last = len1 + len2, lanes = geo.lanes;
  BandK Kc;
  Kc.len1 = len1;
  Kc.len2 = len2;
  Kc.lband = geo.lband;
  Kc.dmax = geo.dmax;
  Kc.match = a.match;
  Kc.mismatch = a.mismatch;
  Kc.gap = a.gap;
  Kc.homo = a.homo_gap;
  BandLane<K> L;
  L.d0 = 2 * K * c - geo.lband - 1;
  L.okE = L.okO = L.sentE = L.sentO = 0;
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const int De = 2 * K * c + 2 * m, Do = De + 1;
    L.okE |= (De <= geo.dmax ? 1u : 0u) << m;
    L.okO |= (Do <= geo.dmax ? 1u : 0u) << m;
    L.sentE |= (geo.sentinels && (De == 0 || De == geo.dmax) ? 1u : 0u) << m;
    L.sentO |= (geo.sentinels && Do == geo.dmax ? 1u : 0u) << m;
    L.prev2[m] = 0;
    L.prev[m] = 0;
  }

  // ---- forward: anti-diagonals 0..last ----
  const int P0 = (geo.lband + 1) & 1;  // parity of D on even steps
  // steady steps: every band diagonal's cell has 1 <= i < len1, 1 <= j < len2
  const int t_lo = imax(geo.lband, geo.rband) + 2;
  const int t_hi = imin(2 * len1 - 2 - geo.lband, 2 * len2 - 2 - geo.rband);
  BandSpecial<K> S;
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const int par = e ? 1 - P0 : P0;
      const bool ok = (((par ? L.okO : L.okE) >> m) & 1u) != 0;
      const bool sent = (((par ? L.sentO : L.sentE) >> m) & 1u) != 0;
      S.mask[e][m] = __builtin_amdgcn_ballot_w64(sent || !ok);
      S.cval[e][m] = sent ? kBandSentinel : 0;
    }
  BandV V;
  V.match = in_vgpr(a.match);
  V.mismatch = in_vgpr(a.mismatch);
  V.gap = in_vgpr(a.gap);
  V.homo = in_vgpr(a.homo_gap);
  for (int blk = 0; blk * 16 <= last; ++blk) {
    const bool steady = geo.sentinels && blk * 16 >= t_lo && blk * 16 + 15 <= t_hi;
    if constexpr (K <= 2) {
      if (steady) {
        if (P0)
          band_block_steady<HOMO, 1, K>(blk, L, V, len1, len2, q1, q2, S, hist_p, lane);
        else
          band_block_steady<HOMO, 0, K>(blk, L, V, len1, len2, q1, q2, S, hist_p, lane);
        continue;
      }
    }
    if (P0)
      band_block<HOMO, 1, K>(blk, last, L, Kc, q1, q2, hist_p, lane);
    else
      band_block<HOMO, 0, K>(blk, last, L, Kc, q1, q2, hist_p, lane);
  }
  __syncthreads();

  // ---- traceback ----
  const size_t sub0 = (size_t)ti * a.sub_stride;
  const bool sub_ok = a.emit || sub0 + a.sub_stride <= a.sub_capacity;
  if (!a.emit && __builtin_amdgcn_ballot_w64(!sub_ok && real) && lane == 0)
    atomicOr(&a.ctr->err_flags, kErrSubOverflow);
  BandBw B;
  B.gpos = g_raw * lanes;
  B.c = c;
  B.lband = geo.lband;
  B.len1 = len1;
  B.len2 = len2;
  B.dmax = geo.dmax;
  B.sentinels = geo.sentinels;
  B.real = real;
  B.sub_ok = sub_ok;
  B.my_sub = sub_ok ? sub0 : 0;
  B.q1 = q1;
  B.q2 = q2;
  B.ao = (a.emit && real) ? a.al_off[task.out_idx] : 0;
  BandPath ps;
  ps.rc = real ? last : -1;
  ps.D = len2 - len1 + geo.lband + 1;
  ps.ns = 0;
  ps.nt = 0;
  ps.bad = false;
  if (solo) {
    if (a.emit)
      band_backward_solo<true, K>(last, hist_p, lane, ps, B, a);
    else
      band_backward_solo<false, K>(last, hist_p, lane, ps, B, a);
  } else if (a.emit) {
    band_backward<true, K>(last, hist_p, lane, ps, B, a);
  } else {
    band_backward<false, K>(last, hist_p, lane, ps, B, a);
  }
  if (__builtin_amdgcn_ballot_w64(ps.bad && real) && lane == 0)
    atomicOr(&a.ctr->err_flags, kErrTraceback);
  __syncthreads();

  // ---- outputs ----
  if (a.emit) {
    if (real) {
      uint8_t* o0 = a.al0 + B.ao;
      uint8_t* o1 = a.al1 + B.ao;
      const int nt = ps.nt;
      for (int x = c; x < nt / 2; x += lanes) {
        const int y = nt - 1 - x;
        const uint8_t t0 = o0[x], t1 = o1[x];
        o0[x] = o0[y];
        o1[x] = o1[y];
        o0[y] = t0;
        o1[y] = t1;
      }
      if (c == 0) a.al_len[task.out_idx] = ps.bad ? 0u : (uint32_t)nt;
    }
  } else if (real && c == 0) {
    a.sub_cnt[ti] = sub_ok ? ((uint32_t)ps.ns | (ps.bad ? 0x80000000u : 0u)) : 0u;
  }
  __syncthreads();
}

template <bool HOMO>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void nw_band_kernel(
    BandDev a) {
  extern __shared__ uint8_t band_lds[];
  const int lane = (int)__lane_id();
  const BinSpace bs = a.bins;
  const uint32_t total = (uint32_t)uni((int)a.ctr->total_waves);
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    // bin of wavefront w: last b with wave_start[b] <= w
    int lo = 0, hi = bs.nbins;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.wave_start[mid] <= w) lo = mid; else hi = mid;
    }
    const int bin = uni(lo);
    int l1, l2;
    bin_lengths(bs, (uint32_t)bin, &l1, &l2);
    const BandGeom geo = band_geom(l1, l2, bs.band, bs.tile);
    const int G = band_groups(geo, HOMO);
    const uint32_t first = a.bin_start[bin] + (w - a.wave_start[bin]) * (uint32_t)G;
    const int count = uni((int)min((uint32_t)G, a.bin_start[bin + 1] - first));
    const int lanes = geo.lanes;
    const int g_raw = lane / lanes;
    const int c = lane - g_raw * lanes;
    const bool real = g_raw < count;
    const int g = real ? g_raw : 0;  // idle lanes shadow group 0
    const DpTask task = a.tasks[first + (uint32_t)g];
    const int len1 = geo.len1, len2 = geo.len2;

    // ---- stage s1 (center) and s2 (query) as packed rows ----
    uint8_t* gb = band_lds + (size_t)g * band_group_bytes(geo, HOMO);
    uint16_t* q1 = reinterpret_cast<uint16_t*>(gb + band_q1_offset(geo));
    uint16_t* q2 = reinterpret_cast<uint16_t*>(gb + band_q2_offset(geo));
    if (real) {
      const uint8_t* src1 = a.pool + task.s1_off;
      const uint8_t* src2 = a.pool + task.s2_off;
      for (int x = c; x < len1; x += lanes) q1[x] = src1[x];
      for (int x = c; x < len2; x += lanes) q2[x] = src2[x];
    }
    __syncthreads();
    if (HOMO && real) {
      // bit 8: position y lies in a run of >= 3 identical bytes; written as
      // the row entry's high byte, so the low bytes read here never change
      for (int x = c; x < len1 + len2; x += lanes) {
        uint16_t* q = x < len1 ? q1 : q2;
        const int n = x < len1 ? len1 : len2;
        const int y = x < len1 ? x : x - len1;
        const uint8_t* qb = reinterpret_cast<const uint8_t*>(q);
        const uint8_t v = qb[2 * y];
        const bool m1 = y >= 1 && qb[2 * y - 2] == v;
        const bool m2 = y >= 2 && m1 && qb[2 * y - 4] == v;
        const bool p1 = y + 1 < n && qb[2 * y + 2] == v;
        const bool p2 = y + 2 < n && p1 && qb[2 * y + 4] == v;
        reinterpret_cast<uint8_t*>(q)[2 * y + 1] = (m2 || (m1 && p1) || p2) ? 1 : 0;
      }
    }
    __syncthreads();

    uint32_t* hist_p = a.hist + (size_t)blockIdx.x * a.hist_stride;
    const uint32_t ti = first + (uint32_t)g;
    if (geo.k == 1)
      band_pairs<HOMO, 1>(a, geo, c, lane, g_raw, real, task, ti, q1, q2, hist_p, G == 1);
    else if (geo.k == 2)
      band_pairs<HOMO, 2>(a, geo, c, lane, g_raw, real, task, ti, q1, q2, hist_p, G == 1);
    else
      band_pairs<HOMO, 4>(a, geo, c, lane, g_raw, real, task, ti, q1, q2, hist_p, G == 1);
  }
}

hipError_t launch_nw_band(const NwArgs& s, hipStream_t st) {
  if (s.grid == 0 || s.hist == nullptr) return hipErrorInvalidValue;
  if (!s.emit_alignment && s.sub_stride == 0) return hipErrorInvalidValue;
  BandDev d;
  d.bins = s.bins;
  d.bin_start = s.bin_start;
  d.wave_start = s.wave_start;
  d.tasks = s.tasks;
  d.pool = s.pool;
  d.match = s.match;
  d.mismatch = s.mismatch;
  d.gap = s.gap_p;
  d.homo_gap = s.homo_gap_p;
  d.ctr = s.ctr;
  d.emit = s.emit_alignment;
  d.sub_pool = s.sub_pool;
  d.sub_capacity = s.sub_capacity;
  d.sub_stride = s.sub_stride;
  d.sub_cnt = s.sub_cnt;
  d.hist = s.hist;
  d.hist_stride = s.hist_stride;
  d.al0 = s.al0;
  d.al1 = s.al1;
  d.al_off = s.al_off;
  d.al_len = s.al_len;
  const size_t lds = s.lds_bytes;
  const void* fn = s.bins.homo ? (const void*)nw_band_kernel<true> : (const void*)nw_band_kernel<false>;
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (s.bins.homo)
    hipLaunchKernelGGL(nw_band_kernel<true>, dim3(s.grid), dim3(64), lds, st, d);
  else
    hipLaunchKernelGGL(nw_band_kernel<false>, dim3(s.grid), dim3(64), lds, st, d);
  return hipGetLastError();
}

size_t nw_band_lds_for(const BinSpace& bs, const uint8_t* present) {
  size_t lds = 0;
  for (int b = 0; b < bs.nbins; ++b) {
    int l1, l2;
    bin_lengths(bs, (uint32_t)b, &l1, &l2);
    if (!bin_present(bs, present, l1, l2)) continue;
    const BandGeom g = band_geom(l1, l2, bs.band, bs.tile);
    const int G = band_groups(g, bs.homo != 0);
    if (G > 0) lds = std::max(lds, band_wave_bytes(g, G, bs.homo != 0));
  }
  lds = (lds + 255) & ~(size_t)255;
  return lds < 256 ? 256 : lds;
}

}  // namespace d2g
