// d2g_stage.h — staging of a compare's aligned pairs into LDS rows, shared
// by the kernels that walk them position by position (lambda, likelihood).
//
// A workgroup of kStWaves wavefronts takes kStTile aligned pairs (the
// compare's gapless tasks, then its DP tasks), kStPairsW per wavefront;
// lanes 0 .. kStPairsW - 1 own them.  Per segment of kStSeg query
// positions each wavefront stages its pairs in turn into their LDS rows:
// lane l loads positions 4l .. 4l + 3 of the pair's query, quality and
// center rows (coalesced 256 bytes per row) and writes one 16-bit code per
// position, (transition row << 2 | nt - 1) | quality << 8, with the
// transitions of compute_lambda (pval.rs:294-309): t = 5 (nt-1) on the
// diagonal, 4 (nt0-1) + (nt-1) at a substitution.  Gapless pairs
// (align_gapless_with_buf, nwalign.rs:645-655) substitute where the bases
// differ below min(len); the owning lanes of DP pairs then patch the
// positions of their al2subs records (query position << 16 | center
// position).  Positions past a pair's end get transition row kStOneRow,
// which callers point at a neutral table entry.
#pragma once

#include "d2g_launch.h"

namespace d2g {

constexpr int kStSeg = 256;              // query positions per staged segment (4 per lane)
constexpr int kStRow = 2 * kStSeg + 4;   // bytes of one pair's LDS row (odd word stride:
                                         // lanes reading the same position hit distinct banks)
constexpr int kStWaves = 4;              // wavefronts per workgroup
constexpr int kStPairsW = 16;            // pairs per wavefront
constexpr int kStTile = kStWaves * kStPairsW;
constexpr int kStUnroll = 8;             // pairs staged per round of loads
constexpr int kStSubBatch = 8;           // substitution records requested together per lane
constexpr uint32_t kStOneRow = 16;       // transition row past a pair's end (index 16 ncol)

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

// The compare's aligned pairs: tasks [0, n_g) gapless, [n_g, n) DP.
struct TaskSet {
  UniqueDev u;
  const GaplessTask* gl;
  const DpTask* dp;
  const uint32_t* sub_pool;  // DP task di's records at di * sub_stride (pool < 2^32 words)
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
};

// The pair a lane owns (lanes 0 .. kStPairsW - 1; others and tasks past n
// own none).  The sequence pool stays under 2^32 bytes (the host checks).
struct TaskRow {
  bool live, gl;
  uint32_t t, out, u_idx, c_idx;
  uint32_t cnt, sub0;       // DP: substitution count and first record
  uint32_t lq, lc;          // query length; gapless: center length, else 0
  uint32_t qoff, coff;      // row offsets
  bool bad;                 // the traceback flagged this pair
};

__device__ __forceinline__ TaskRow task_row(const TaskSet& a, uint32_t n_g, uint32_t n, uint32_t t,
                                            bool owner) {
  TaskRow r;
  r.t = t;
  r.live = owner && t < n;
  r.gl = r.live && t < n_g;
  r.out = r.u_idx = r.c_idx = r.cnt = r.sub0 = 0;
  r.bad = false;
  if (r.gl) {
    const GaplessTask task = a.gl[t];
    r.u_idx = task.u_idx;
    r.c_idx = task.c_idx;
    r.out = task.out_idx;
  } else if (r.live) {
    const uint32_t di = t - n_g;
    const DpTask task = a.dp[di];
    r.u_idx = task.u_idx;
    r.c_idx = task.c_idx;
    r.out = task.out_idx;
    const uint32_t cnt_raw = a.sub_cnt[di];
    r.bad = (cnt_raw & 0x80000000u) != 0;
    r.cnt = cnt_raw & 0x7FFFFFFFu;
    r.sub0 = di * a.sub_stride;
  }
  r.lq = r.live ? a.u.len[r.u_idx] : 0u;
  r.lc = r.gl ? a.u.len[r.c_idx] : 0u;
  r.qoff = (uint32_t)a.u.off[r.u_idx];
  r.coff = (uint32_t)a.u.off[r.c_idx];
  return r;
}

// One round of staging: kStUnroll pairs' words at positions 4 lane .. 4
// lane + 3 of the segment.
struct StageRound {
  uint32_t qw[kStUnroll], uw[kStUnroll], cw[kStUnroll];
  uint32_t lq[kStUnroll], lc[kStUnroll];
};

// Requests the words of the wavefront's pairs p0 .. p0 + kStUnroll - 1
// (pair p's fields sit in lane p; positions past a row's end read the row's
// first word, which stage_round ignores).
__device__ __forceinline__ void request_round(StageRound& R, uint32_t p0, uint32_t s0,
                                              uint32_t lane, const TaskRow& me, const uint8_t* seq,
                                              const uint8_t* qual) {
#pragma unroll
  for (int r = 0; r < kStUnroll; ++r) {
    const uint32_t pr = p0 + (uint32_t)r;
    R.lq[r] = rl(me.lq, pr);
    R.lc[r] = rl(me.lc, pr);
    const uint32_t pq = rl(me.qoff, pr), pc = rl(me.coff, pr);
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
// collects the count in ham.
__device__ __forceinline__ void stage_round(const StageRound& R, uint32_t p0, uint32_t s0,
                                            uint32_t lane, bool quals, bool any_gl, uint8_t* rows,
                                            uint32_t& ham) {
#pragma unroll
  for (int r = 0; r < kStUnroll; ++r) {
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
      const uint32_t tr = in ? (sub ? 4u * (cn - 1u) + nt1 : 5u * nt1) : kStOneRow;
      const uint32_t qi = (quals && in) ? (R.uw[r] >> (8 * b)) & 0xFFu : 0u;
      const uint32_t code = (((tr << 2) | (nt1 & 3u)) & 0xFFu) | (qi << 8);
      packed[b >> 1] |= code << (16 * (b & 1));
    }
    if (any_gl) {
      for (int m = 1; m < 64; m <<= 1) h += __shfl_xor(h, m);
    }
    uint32_t* row = reinterpret_cast<uint32_t*>(rows + pr * kStRow) + 2 * lane;
    row[0] = packed[0];
    row[1] = packed[1];
    if (lane == pr) ham += h;
  }
}

// Substitution cursor of one DP pair.  A traceback writes a pair's records
// in descending query position (the path runs from the matrix's end, and a
// substitution is a diagonal move), so the segments, staged in ascending
// position, consume them from the end of the list: records [0, left) are
// still to patch, and nqp is the query position of record left - 1 once a
// batch has read it beyond the current segment (kQpUnknown before).
constexpr uint32_t kQpUnknown = 0xFFFFFFFFu;
struct SubCursor {
  uint32_t left, nqp;
};
__device__ __forceinline__ SubCursor sub_cursor(const TaskRow& me) {
  SubCursor c;
  c.left = me.cnt;
  c.nqp = kQpUnknown;
  return c;
}

// Stages the segment at s0 of the wavefront's pairs (rows: the wavefront's
// kStPairsW rows; myb: the lane's own row, row kStPairsW - 1 for lanes that
// own none): two rounds of codes, the second round's words requested before
// the first round's bytes are converted, then the DP substitution patches of
// the segment, kStSubBatch records per round from the cursor.  The caller
// runs wave_lds_sync() before reading other lanes' writes.
__device__ __forceinline__ void stage_segment(const TaskSet& a, const TaskRow& me, uint32_t s0,
                                              uint32_t lane, bool quals, bool any_gl,
                                              uint8_t* rows, uint8_t* myb, uint32_t& ham,
                                              SubCursor& cur) {
  const uint8_t* seq = a.u.seq;
  const uint8_t* qual = quals ? a.u.qual : a.u.seq;
  StageRound ra, rb;
  request_round(ra, 0, s0, lane, me, seq, qual);
  request_round(rb, kStUnroll, s0, lane, me, seq, qual);
  __builtin_amdgcn_sched_barrier(0);
  stage_round(ra, 0, s0, lane, quals, any_gl, rows, ham);
  __builtin_amdgcn_sched_barrier(0);
  stage_round(rb, kStUnroll, s0, lane, quals, any_gl, rows, ham);
  // DP substitutions of this segment (lanes that own none have left = 0;
  // lanes without a record here write their row's pad byte, which nothing
  // reads)
  const uint32_t end = s0 + (uint32_t)kStSeg;
  bool want = cur.left > 0 && (cur.nqp == kQpUnknown || cur.nqp < end);
  while (__builtin_amdgcn_ballot_w64(want)) {
    uint32_t rec[kStSubBatch], cb[kStSubBatch];
#pragma unroll
    for (int j = 0; j < kStSubBatch; ++j) {
      // slots past the list read the pool's first word: a pair whose list
      // overflowed the pool has cnt = 0 and a slot past the pool's end
      const bool v = want && (uint32_t)j < cur.left;
      rec[j] = ld_off<uint32_t>(a.sub_pool, (v ? me.sub0 + cur.left - 1u - (uint32_t)j : 0u) << 2);
    }
#pragma unroll
    for (int j = 0; j < kStSubBatch; ++j) {
      const bool v = want && (uint32_t)j < cur.left;
      cb[j] = ld_off<uint8_t>(seq, me.coff + (v ? rec[j] & 0xFFFFu : 0u));
    }
    // records ascend in query position along j: the ones in this segment
    // are a prefix
    uint32_t take = 0, nq = kQpUnknown;
    bool beyond = false;
#pragma unroll
    for (int j = 0; j < kStSubBatch; ++j) {
      const bool v = want && (uint32_t)j < cur.left;
      const uint32_t qp = rec[j] >> 16;
      const bool in = v && !beyond && qp < end;
      nq = (v && !beyond && !in) ? qp : nq;
      beyond = beyond || (v && !in);
      take += in ? 1u : 0u;
      const bool here = in && qp >= s0;
      uint8_t* c = myb + (here ? 2 * (qp - s0) : 2 * kStSeg);
      const uint32_t nt1 = (uint32_t)c[0] & 3u;
      c[0] = (uint8_t)(((4u * (cb[j] - 1u) + nt1) << 2) | nt1);
    }
    const uint32_t avail = want ? min(cur.left, (uint32_t)kStSubBatch) : 0u;
    cur.left -= take;
    cur.nqp = beyond ? nq : kQpUnknown;
    // a full batch in this segment may be followed by more of it
    want = want && !beyond && take == avail && cur.left > 0;
  }
}

// Table index of a staged code: transition row * ncol + quality.
__device__ __forceinline__ uint32_t code_index(uint32_t code, uint32_t ncol) {
  return ((code & 0xFFu) >> 2) * ncol + (code >> 8);
}

}  // namespace d2g
