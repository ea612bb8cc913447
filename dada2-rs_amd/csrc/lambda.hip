// lambda.hip — compute_lambda (pval.rs:253-320) for every aligned pair.
//
// lambda = prod over query positions pos = 0..len-1, left to right from 1.0,
// of err[t(pos) * ncol + q(pos)], where t(pos) = 5*(nt-1) (the query base on
// the diagonal) unless an al2subs substitution maps the query position to a
// center base (pval.rs:294-309).  The f64 multiplications keep that exact
// order, so results are bit-identical to the reference.
//
// Gapless pairs (align_gapless_with_buf, nwalign.rs:645-655) pair positions
// directly: substitutions are the positions < min(len) where bases differ.
// DP pairs read the substitution list the traceback wrote: (query position,
// center position) records.
//
// One lane per pair (the product is one ordered chain).  A workgroup of
// four wavefronts takes 64 pairs, 16 per wavefront (lanes 0..15 own them),
// and feeds the chains from LDS so that no global load sits on them.  Per
// segment of 256 query positions each wavefront stages its 16 pairs in turn
// into their LDS rows: lane l loads positions 4l .. 4l + 3 of the pair's
// query, quality and center rows (coalesced 256 bytes per row) and writes
// (transition row << 2 | nt - 1, quality) byte pairs, positions past the
// pair's end getting the code of an extra 1.0 entry after the error matrix
// (x * 1.0 = x for every f64).  The owning lanes then patch their pairs'
// substitution positions and run the multiplications without branches.
#include "d2g_launch.h"

namespace d2g {

constexpr int kLamSeg = 256;              // query positions per staged segment (4 per lane)
constexpr int kLamRow = 2 * kLamSeg + 4;  // bytes of one pair's LDS row (odd word stride:
                                          // lanes reading the same position hit distinct banks)
constexpr int kLamWaves = 4;              // wavefronts per workgroup
constexpr int kLamPairsW = 16;            // pairs per wavefront
constexpr int kLamTile = kLamWaves * kLamPairsW;
constexpr int kLamUnroll = 8;             // pairs staged per round of loads
constexpr int kLamSubBatch = 8;           // substitution records requested together per lane
constexpr uint32_t kLamOneRow = 16;       // transition row of the 1.0 entry (index 16 ncol)

struct LambdaDev {
  UniqueDev u;
  const GaplessTask* gl;
  const DpTask* dp;
  const uint32_t* sub_pool;
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
  const double* err;
  uint32_t ncol;
  int use_quals;
  Counters* ctr;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
};

// A load at a 32-bit byte offset from a base (the offset stays 32-bit, so
// a wavefront-uniform base takes the scalar-base addressing form).
template <class T>
__device__ __forceinline__ T ld_off(const void* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(base) + byte_off);
}

// Orders this wavefront's LDS writes before its later LDS reads by other
// lanes (LDS serves one wavefront's accesses in order; the fence keeps the
// compiler from moving them and waits for the writes).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// One round of staging: kLamUnroll pairs' words at positions 4 lane .. 4
// lane + 3 of the segment.
struct StageRound {
  uint32_t qw[kLamUnroll], uw[kLamUnroll], cw[kLamUnroll];
  uint32_t lq[kLamUnroll], lc[kLamUnroll];
};

// Requests the words of the wavefront's pairs p0 .. p0 + kLamUnroll - 1
// (pair p's fields sit in lane p; positions past a row's end read the row's
// first word, which stage_round ignores).
__device__ __forceinline__ void request_round(StageRound& R, uint32_t p0, uint32_t s0,
                                              uint32_t lane, uint32_t lq, uint32_t lc,
                                              uint32_t qoff, uint32_t coff, const uint8_t* seq,
                                              const uint8_t* qual) {
#pragma unroll
  for (int r = 0; r < kLamUnroll; ++r) {
    const uint32_t pr = p0 + (uint32_t)r;
    R.lq[r] = rl(lq, pr);
    R.lc[r] = rl(lc, pr);
    const uint32_t pq = rl(qoff, pr), pc = rl(coff, pr);
    const uint32_t pos = s0 + 4u * lane;
    const uint32_t qo = pos < R.lq[r] ? pq + pos : pq;
    const uint32_t co = pos < R.lc[r] ? pc + pos : pc;
    R.qw[r] = ld_off<uint32_t>(seq, qo);
    R.uw[r] = ld_off<uint32_t>(qual, qo);
    R.cw[r] = ld_off<uint32_t>(seq, co);
  }
}

// Writes the round's codes into the pairs' LDS rows; a gapless pair (lc >
// 0) substitutes where its bases differ below min(len), and its owning lane
// collects the count.
__device__ __forceinline__ void stage_round(const StageRound& R, uint32_t p0, uint32_t s0,
                                            uint32_t lane, bool quals, bool any_gl, uint8_t* rows,
                                            uint32_t& ham) {
#pragma unroll
  for (int r = 0; r < kLamUnroll; ++r) {
    const uint32_t pr = p0 + (uint32_t)r;
    uint32_t packed[2] = {0u, 0u};
    uint32_t h = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint32_t pos = s0 + 4u * lane + (uint32_t)b;
      const uint32_t nt1 = ((R.qw[r] >> (8 * b)) & 0xFFu) - 1u;
      const uint32_t cn = (R.cw[r] >> (8 * b)) & 0xFFu;
      const bool in = pos < R.lq[r];
      const bool sub = in && pos < R.lc[r] && cn != nt1 + 1u;
      h += sub ? 1u : 0u;
      const uint32_t tr = in ? (sub ? 4u * (cn - 1u) + nt1 : 5u * nt1) : kLamOneRow;
      const uint32_t qi = (quals && in) ? (R.uw[r] >> (8 * b)) & 0xFFu : 0u;
      const uint32_t code = (((tr << 2) | (nt1 & 3u)) & 0xFFu) | (qi << 8);
      packed[b >> 1] |= code << (16 * (b & 1));
    }
    if (any_gl) {
      for (int m = 1; m < 64; m <<= 1) h += __shfl_xor(h, m);
    }
    uint32_t* row = reinterpret_cast<uint32_t*>(rows + pr * kLamRow) + 2 * lane;
    row[0] = packed[0];
    row[1] = packed[1];
    if (lane == pr) ham += h;
  }
}

__global__ __launch_bounds__(64 * kLamWaves) void lambda_kernel(LambdaDev a) {
  extern __shared__ double err_lds[];
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x / 64u;
  const uint32_t ncol = a.ncol;
  const uint32_t ne = 16u * ncol;
  const uint32_t n_g = a.ctr->n_gapless;
  const uint32_t n = n_g + a.ctr->n_dp;
  // the error matrix and, after it, the 1.0 entry
  for (uint32_t i0 = 0; i0 <= ne; i0 += 4 * 64 * kLamWaves) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = a.err[min(i0 + (uint32_t)k * 64u * kLamWaves + threadIdx.x, ne - 1u)];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + (uint32_t)k * 64u * kLamWaves + threadIdx.x;
      if (i < ne) err_lds[i] = v[k];
      if (i == ne) err_lds[i] = 1.0;
    }
  }
  uint8_t* rows = reinterpret_cast<uint8_t*>(err_lds + ne + 1) + wave * kLamPairsW * kLamRow;
  const bool quals = a.use_quals && a.u.qual != nullptr;
  const uint8_t* seq = a.u.seq;
  const uint8_t* qual = quals ? a.u.qual : a.u.seq;
  __syncthreads();
  for (uint32_t base = blockIdx.x * (uint32_t)kLamTile; base < n; base += gridDim.x * (uint32_t)kLamTile) {
    // ---- this lane's pair (lanes 0..15; the sequence pool and the
    // substitution pool stay under 2^32 bytes / words: the host checks) ----
    const uint32_t t = base + wave * kLamPairsW + lane;
    const bool live = lane < (uint32_t)kLamPairsW && t < n;
    const bool gl = live && t < n_g;
    uint32_t u_idx = 0, c_idx = 0, out = 0, cnt = 0, sub0 = 0;
    if (gl) {
      const GaplessTask task = a.gl[t];
      u_idx = task.u_idx;
      c_idx = task.c_idx;
      out = task.out_idx;
    } else if (live) {
      const uint32_t di = t - n_g;
      const DpTask task = a.dp[di];
      u_idx = task.u_idx;
      c_idx = task.c_idx;
      out = task.out_idx;
      const uint32_t cnt_raw = a.sub_cnt[di];
      if (cnt_raw & 0x80000000u) atomicOr(&a.ctr->err_flags, kErrTraceback);
      cnt = cnt_raw & 0x7FFFFFFFu;
      sub0 = di * a.sub_stride;
    }
    const uint32_t lq = live ? a.u.len[u_idx] : 0u;
    // gapless pairs compare positions with the center; others see lc = 0
    const uint32_t lc = gl ? a.u.len[c_idx] : 0u;
    const uint32_t qoff = (uint32_t)a.u.off[u_idx], coff = (uint32_t)a.u.off[c_idx];
    const bool any_gl = __builtin_amdgcn_ballot_w64(gl) != 0;
    double lam = 1.0;
    uint32_t ham = 0;
    for (uint32_t s0 = 0; __builtin_amdgcn_ballot_w64(s0 < lq); s0 += kLamSeg) {
      // ---- stage the wavefront's 16 pairs (the second round's loads are
      // requested before the first round's bytes are converted) ----
      StageRound ra, rb;
      request_round(ra, 0, s0, lane, lq, lc, qoff, coff, seq, qual);
      request_round(rb, kLamUnroll, s0, lane, lq, lc, qoff, coff, seq, qual);
      __builtin_amdgcn_sched_barrier(0);
      stage_round(ra, 0, s0, lane, quals, any_gl, rows, ham);
      __builtin_amdgcn_sched_barrier(0);
      stage_round(rb, kLamUnroll, s0, lane, quals, any_gl, rows, ham);
      // ---- DP substitutions: each owning lane patches its pair's row
      // (lanes 16..63 own none: cnt = 0, and they point at row 15's pad
      // byte, which nothing reads) ----
      uint8_t* myb = rows + min(lane, (uint32_t)kLamPairsW - 1u) * kLamRow;
      for (uint32_t k0 = 0; __builtin_amdgcn_ballot_w64(k0 < cnt); k0 += kLamSubBatch) {
        uint32_t rec[kLamSubBatch], cb[kLamSubBatch];
#pragma unroll
        for (int j = 0; j < kLamSubBatch; ++j) {
          const uint32_t k = k0 + (uint32_t)j;
          // empty slots read the pool's first word: a pair whose list
          // overflowed the pool has cnt = 0 and a slot past the pool's end
          rec[j] = ld_off<uint32_t>(a.sub_pool, (k < cnt ? sub0 + k : 0u) << 2);
        }
#pragma unroll
        for (int j = 0; j < kLamSubBatch; ++j) {
          const uint32_t k = k0 + (uint32_t)j;
          cb[j] = ld_off<uint8_t>(seq, coff + (k < cnt ? rec[j] & 0xFFFFu : 0u));
        }
#pragma unroll
        for (int j = 0; j < kLamSubBatch; ++j) {
          const uint32_t k = k0 + (uint32_t)j;
          const uint32_t qp = rec[j] >> 16;
          // records outside the segment rewrite the row's pad byte instead
          const bool here = k < cnt && qp >= s0 && qp - s0 < (uint32_t)kLamSeg;
          uint8_t* c = myb + (here ? 2 * (qp - s0) : 2 * kLamSeg);
          const uint32_t nt1 = (uint32_t)c[0] & 3u;
          c[0] = (uint8_t)(((4u * (cb[j] - 1u) + nt1) << 2) | nt1);
        }
      }
      // the wavefront's staged rows are read by other lanes of the same
      // wavefront (segment counts differ between wavefronts, so no
      // workgroup barrier here)
      wave_lds_sync();
      // ---- the ordered product over the segment (owning lanes) ----
      const uint32_t seg_n = lq > s0 ? min(lq - s0, (uint32_t)kLamSeg) : 0u;
      if (lane < (uint32_t)kLamPairsW) {
        const uint32_t* my = reinterpret_cast<const uint32_t*>(myb);
        for (uint32_t i = 0; __builtin_amdgcn_ballot_w64(i < seg_n); i += 32) {
          uint32_t w16[16];
#pragma unroll
          for (int h = 0; h < 16; ++h) w16[h] = my[(i >> 1) + (uint32_t)h];
          double f[32];
#pragma unroll
          for (int e = 0; e < 32; ++e) {
            const uint32_t code = (w16[e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
            f[e] = err_lds[((code & 0xFFu) >> 2) * ncol + (code >> 8)];
          }
#pragma unroll
          for (int e = 0; e < 32; ++e) lam *= f[e];
        }
      }
      wave_lds_sync();  // rows are restaged for the next segment
    }
    if (live) {
      a.out_lambda[out] = lam;
      a.out_hamming[out] = gl ? ham : cnt;
      a.out_skipped[out] = 0;
    }
  }
}

hipError_t launch_lambda(const LambdaArgs& s, hipStream_t st) {
  LambdaDev d;
  d.u = *s.u;
  d.gl = s.gapless_tasks;
  d.dp = s.dp_tasks;
  d.sub_pool = s.sub_pool;
  d.sub_stride = s.sub_stride;
  d.sub_cnt = s.sub_cnt;
  d.err = s.err;
  d.ncol = s.ncol;
  d.use_quals = s.use_quals;
  d.ctr = s.ctr;
  d.out_lambda = s.out_lambda;
  d.out_hamming = s.out_hamming;
  d.out_skipped = s.out_skipped;
  const size_t lds = ((size_t)16 * s.ncol + 1) * sizeof(double) + (size_t)kLamTile * kLamRow;
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)lambda_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(lambda_kernel, dim3(s.grid), dim3(64 * kLamWaves), lds, st, d);
  return hipGetLastError();
}

}  // namespace d2g
