// nw_vec.hip — the compressed anti-diagonal banded Needleman-Wunsch of
// align_vectorized_with_buf (nwalign.rs:761-1140), bit-exact, on gfx950.
//
// Mapping.  The reference stores cell (i, j) of anti-diagonal row = i + j in
// compressed column (2*start_col + j - i) / 2 of an ncol-wide row
// (nwalign.rs:783-787).  A wavefront holds G <= 64 / ncol pairs of one
// (len1, len2) geometry; lane c of group g owns column c of pair g for every
// row, so a column's scores stay in registers: prev2 (row-2, the diagonal
// source) and prev (row-1).  The left/up operands (nwalign.rs:957-964) are
// columns c-1/c on even rows and c/c+1 on odd rows: a one-lane DPP wavefront
// shift.  Columns 0 and ncol-1 never become active, so a shift across a
// group boundary is never consumed.
//
// Schedule.  The row walk (col_min, col_max, i_max, j_min, band wedges,
// ends-free recalculation; nwalign.rs:865-1088) depends only on the
// geometry.  sched_kernel replays it once per bin into a table of RowRec
// records; the DP reads one record per row through scalar loads, so the
// per-row scalar work is a load instead of the walk.  The cell update is
// branch-free and specialised on the swap flag (the tie precedence of
// dploop, nwalign.rs:700-737).  Saturating i16 arithmetic is emulated on
// i32 with clamps.
//
// Steady rows (the bulk of the anti-diagonals) take a short update: their
// sequence bytes are read at fixed LDS offsets from a per-block base, the
// active lanes come as a 64-bit mask from the schedule, and where the scores
// provably stay inside i16 the saturation clamps are dropped (a clamp only
// changes a sentinel-derived candidate, which never wins there).
//
// Traceback.  Each lane shifts its 2-bit pointer per row into a private
// history word (word = 4 word + pointer, so row r sits at bits 2 (15 - r mod
// 16)) flushed to global memory every 16 rows.  A diagonal step over a
// mismatch is stored as pointer 0; only active cells carry diagonal
// pointers, so the traceback tells it from the reference's invalid pointer
// by the row's active range.  The traceback (nwalign.rs:1101-1139) walks
// every pair's path down the history blocks a run of diagonal matches at a
// time (the path state is replicated across the pair's lanes).  Mismatching
// diagonal steps become al2subs substitutions (nwalign.rs:1169-1224),
// written straight to global memory as (query position << 16 | center
// position) for the lambda kernel; in alignment-output mode the alignment
// strings are written instead.
#include <algorithm>
#include <type_traits>

#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

struct NwDev {
  BinSpace bins;
  const uint32_t* bin_start;
  const uint32_t* wave_start;
  const DpTask* tasks;
  const uint8_t* pool;
  const RowRec* sched;
  const uint64_t* sched_off;
  const BinInfo* info;
  int match, mismatch, gap, end_gap;
  int steady_clamp;  // keep the i16 clamps in steady rows
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

// ---------------------------------------------------------------------------
// Schedule: a replay of the reference's row walk, one record per row 2..last,
// handed to sink(row, R) in row order.
// ---------------------------------------------------------------------------
template <class Sink>
D2G_HD void vec_schedule(const VecGeom& geo, int end_gap, int gap, Sink&& sink) {
  const int len1 = geo.len1, len2 = geo.len2, band = geo.band, start_col = geo.start_col;
  const int lfl = geo.left_fill_limit, tfl = geo.top_fill_limit, top_limit = geo.top_limit;
  const int band_min_len1 = imin(band, len1);
  const bool do_recalc = end_gap > gap;
  int col_min = start_col, col_max = start_col, i_max = 0, j_min = 0;
  bool recalc_left = false, recalc_right = false, even = true;
  int row_free = end_gap;
  for (int row = 2; row <= len1 + len2; ++row) {
    RowRec R;
    const int n = col_max - col_min + 1;
    row_free = sat16(row_free + end_gap);
    int k_lo = i_max - len1 + 1;
    k_lo = k_lo < 0 ? 0 : (k_lo > n ? n : k_lo);
    int k_hi = imin(i_max + 1, len2 - j_min);
    k_hi = k_hi < 0 ? 0 : (k_hi > n ? n : k_hi);
    if (k_hi < k_lo) k_hi = k_lo;
    R.a1 = i_max + col_min;
    R.a2 = j_min - col_min;
    R.inr_lo = col_min + k_lo;
    R.inr_n = (uint32_t)(k_hi - k_lo);
    R.cmin = col_min;
    R.n = (uint32_t)n;
    R.edge_l = row < lfl ? imax(start_col - 1 - (row - 1) / 2, 0) : -1;
    R.edge_t = row < tfl ? start_col + row / 2 : -1;
    R.row_free = row_free;
    uint32_t flags = 0;
    // band transition (nwalign.rs:980-987)
    const bool trans_min = row == band_min_len1;
    if (trans_min) {
      col_min = imax(col_min - 1, 0);
      i_max += 1;
      j_min -= 1;
      flags |= 4u;
    }
    if (row == top_limit) col_max += 1;
    // ends-free recalculation flags (nwalign.rs:993-1026)
    if (do_recalc) {
      if (recalc_left) flags |= 1u;
      if (i_max == len1 - 1) recalc_left = true;
      if (recalc_right) flags |= 2u;
      if ((row + 1) / 2 + col_max - start_col == len2) recalc_right = true;
    }
    flags |= ((uint32_t)col_min & 0xFFu) << 8;
    flags |= ((uint32_t)col_max & 0xFFu) << 16;
    R.flags = flags;
    sink(row, R);
    // index walk (nwalign.rs:1032-1079)
    if (row < band && row < len1) {
      if (even) col_min = imax(col_min - 1, 0);
      i_max += 1;
    } else if (i_max < len1 - 1) {
      if (band % 2 == 0) {
        if (even) j_min += 1;
        else i_max += 1;
      } else if (even) {
        col_min = imax(col_min - 1, 0);
        i_max += 1;
      } else {
        col_min += 1;
        j_min += 1;
      }
    } else {
      if (!even) col_min += 1;
      j_min += 1;
    }
    if (row < top_limit) {
      if (!even) col_max += 1;
    } else if ((row + 1) / 2 + col_max - start_col < len2) {
      if ((band + len2 - len1) % 2 == 0) {
        if (even) col_max = imax(col_max - 1, 0);
        else col_max += 1;
      }
    } else if (even) {
      col_max = imax(col_max - 1, 0);
    }
    even = !even;
  }
}

__global__ void sched_kernel(BinSpace bs, const uint64_t* off, int end_gap, int gap,
                             RowRec* table, BinInfo* info) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bs.nbins) return;
  const uint64_t o = off[b];
  if (o == kSchedAbsent) return;  // a length pair no task can have
  int l1, l2;
  bin_lengths(bs, (uint32_t)b, &l1, &l2);
  const VecGeom g = vec_geom(l1, l2, bs.band);
  if (nw_groups(g, bs.tile) == 0) {  // the one-pair kernel's bins: only their cell count
    BinInfo bi;
    bi.steady_lo = 0;
    bi.steady_hi = 0;
    bi.cells = wide_fits(g) ? (uint32_t)vec_cells(l1, l2, bs.band) : 0u;
    bi.masked = 0;
    info[b] = bi;
    return;
  }
  RowRec* t = table + o;  // nrow records: rows 2 .. len1 + len2, two lookahead
  const int last = g.len1 + g.len2;
  const int G = nw_groups(g, bs.tile);
  // One pass over the walk: each record gets its lane masks and masked-row
  // flag and is stored; the steady-run search and the masked-row check keep
  // what they need of the two rows above in registers, so nothing is read
  // back from the table.
  uint64_t act_p1 = 0, act_p2 = 0;  // active lanes of rows row-1, row-2
  int a1_p1 = 0, a2_p1 = 0;         // s1 / s2 offsets of row row-1
  int best_lo = 0, best_hi = 0, run_lo = -1;
  int nm_first = -1, nm_last = -1;  // rows without the masked-row flag
  uint32_t cells = 0;
  RowRec last_rec;
  auto sink = [&](const int row, RowRec& R) {
    // active-lane masks over the G pairs of a wavefront (pair gi on lanes
    // gi * ncol .. gi * ncol + ncol - 1)
    uint64_t base = 0;
    if (R.n > 0 && R.cmin < 64)
      base = (R.n >= 64 ? ~0ull : ((1ull << R.n) - 1ull)) << R.cmin;
    uint64_t m = 0;
    for (int gi = 0; gi < G; ++gi) m |= base << (gi * g.ncol);
    R.act_lo = (uint32_t)m;
    R.act_hi = (uint32_t)(m >> 32);
    const bool window = R.inr_lo == R.cmin && R.inr_n == R.n;
    const bool inside = R.n == 0 || (R.cmin >= 1 && R.cmin + (int)R.n <= g.ncol - 1);
    const int cmin_new = (int)((R.flags >> 8) & 0xFFu), cmax_new = (int)((R.flags >> 16) & 0xFFu);
    const bool rec_l = !(R.flags & 1u) || cmin_new >= 1;
    const bool rec_r = !(R.flags & 2u) || cmax_new + 1 <= g.ncol - 1;
    if (window && inside && rec_l && rec_r) {
      R.flags |= kRowMasked;
    } else {
      if (nm_first < 0) nm_first = row;
      nm_last = row;
    }
    t[row - 2] = R;
    // longest run of steady rows whose s1 / s2 offsets advance by one after
    // each odd / even row.  A steady row's active cells keep off each pair's
    // outer columns (a lane-shift that leaves the wavefront lands on an
    // inactive lane), and its active lanes are those of the row two up (so
    // every active cell has an active diagonal predecessor, and a run needs
    // one mask per parity).
    cells += R.n;
    const uint64_t act2 = row >= 4 ? act_p2 : 0ull;
    const bool steady = R.edge_l < 0 && R.edge_t < 0 && (R.flags & 3u) == 0 && window &&
                        R.cmin >= 1 && R.cmin + (int)R.n <= g.ncol - 1 && m == act2;
    if (!steady) {
      run_lo = -1;
    } else {
      if (run_lo >= 0) {
        const bool odd_prev = ((row - 1) & 1) != 0;
        const bool linked = R.a1 - a1_p1 == (odd_prev ? 1 : 0) && R.a2 - a2_p1 == (odd_prev ? 0 : 1);
        if (!linked) run_lo = row;
      } else {
        run_lo = row;
      }
      if (row + 1 - run_lo > best_hi - best_lo) {
        best_lo = run_lo;
        best_hi = row + 1;
      }
    }
    act_p2 = act_p1;
    act_p1 = m;
    a1_p1 = R.a1;
    a2_p1 = R.a2;
    last_rec = R;
  };
  vec_schedule(g, end_gap, gap, sink);
  // two lookahead records past the last row (rows last+1, last+2)
  t[last - 1] = last_rec;
  t[last] = last_rec;
  // the run trimmed to even bounds and kept off the last row (whose history
  // word the generic and masked rows flush)
  int lo = (best_lo + 1) & ~1, hi = imin(best_hi, last) & ~1;
  if (hi <= lo) lo = hi = last + 1;
  // every row outside [lo, hi) carries the masked-row flag
  const bool all_masked = nm_first < 0 || (nm_first >= lo && nm_last < hi);
  BinInfo bi;
  bi.steady_lo = (uint32_t)lo;
  bi.steady_hi = (uint32_t)hi;
  bi.cells = cells;
  bi.masked = all_masked ? 1u : 0u;
  info[b] = bi;
}

hipError_t launch_sched(const BinSpace& bs, const uint64_t* off, int end_gap, int gap,
                        RowRec* table, BinInfo* info, hipStream_t s) {
  const int threads = 64;
  const int blocks = (bs.nbins + threads - 1) / threads;
  hipLaunchKernelGGL(sched_kernel, dim3(blocks), dim3(threads), 0, s, bs, off, end_gap, gap,
                     table, info);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// DP
// ---------------------------------------------------------------------------

__device__ __forceinline__ int clamp16(int v) { return v > 32767 ? 32767 : (v < -32768 ? -32768 : v); }

// Row constants of one wavefront (uniform).
struct RowConst {
  int match, mismatch, gap, end_gap, fill, last;
  int dmatch, dmis;  // match - 2 gap, mismatch - 2 gap: the steady rows' offset diagonal
};
// Per-lane column state.
struct LaneState {
  int prev2, prev;
  uint32_t hp;  // pointer history, 2 bits x 16 rows
};

// One anti-diagonal for every pair of the wavefront (nwalign.rs:875-1026).
// ca, cb: this lane's bases s1[i_max - k] and s2[j_min + k], k = c - col_min
// (nwalign.rs:922-953), loaded a row ahead by the caller.
template <bool EVEN, bool SWAP>
__device__ __forceinline__ void nw_row(const int row, const RowRec R, LaneState& ls,
                                       const RowConst K, const int ca, const int cb, const int c,
                                       const int lane, uint32_t* hist_p) {
  const bool inr = (unsigned)(c - R.inr_lo) < R.inr_n;
  const bool differ = ca != cb;
  const int score = inr ? (differ ? K.mismatch : K.match) : K.fill;
  const int dg = clamp16(ls.prev2 + score);
  const int lraw = EVEN ? from_lower(ls.prev, K.fill) : ls.prev;
  const int uraw = EVEN ? ls.prev : from_upper(ls.prev, K.fill);
  const int l = clamp16(lraw + K.gap), u = clamp16(uraw + K.gap);
  // dploop precedence (nwalign.rs:700-737); max(l, u) is the entry either way
  const bool upf = SWAP ? (l < u) : (u >= l);
  int e = l > u ? l : u;
  int p = upf ? 3 : 2;
  p = dg > e ? 1 : p;
  e = dg > e ? dg : e;
  // outside [col_min, col_max]: sentinel or ends-free edge (nwalign.rs:884-899)
  const bool act = (unsigned)(c - R.cmin) < R.n;
  const bool isT = c == R.edge_t, isL = c == R.edge_l;
  e = act ? e : ((isT || isL) ? R.row_free : K.fill);
  p = act ? p : (isT ? 2 : (isL ? 3 : 0));
  // ends-free recalculation (nwalign.rs:993-1026), only on the last rows
  if (R.flags & 3u) {
    const bool trans_min = (R.flags & 4u) != 0;
    const int cmin_new = (int)((R.flags >> 8) & 0xFFu);
    const int cmax_new = (int)((R.flags >> 16) & 0xFFu);
    if (R.flags & 1u) {
      const int d_free = clamp16((trans_min ? uraw : lraw) + K.end_gap);
      const bool here = c == cmin_new;
      const bool take = d_free > e;
      const bool tie = d_free == e && (SWAP ? p != 2 : p == 1);
      e = (here && take) ? d_free : e;
      p = (here && (take || tie)) ? 2 : p;
    }
    if (R.flags & 2u) {
      const int src = trans_min ? from_upper(uraw, K.fill) : uraw;
      const int d_free = clamp16(src + K.end_gap);
      const bool here = c == cmax_new;
      const bool take = d_free > e;
      const bool tie = d_free == e && (SWAP ? p == 1 : p != 3);
      e = (here && take) ? d_free : e;
      p = (here && (take || tie)) ? 3 : p;
    }
  }
  // pointer history.  A diagonal pointer over a mismatch is stored as 0;
  // the traceback tells it from an unwritten cell because only active cells
  // carry diagonal pointers.  (In the compare path every base is in 1..4,
  // so each such step is an al2subs substitution.)
  p = (p == 1 && inr && differ) ? 0 : p;
  ls.hp = ls.hp * 4u + (uint32_t)p;
  if ((row & 15) == 15 || row == K.last) {
    hist_p[(row >> 4) * 64 + lane] = ls.hp << (2 * (15 - (row & 15)));
    ls.hp = 0;
  }
  ls.prev2 = ls.prev;
  ls.prev = e;
}

// The schedule is read-only for the DP: reading it through the constant
// address space lets every record arrive by scalar loads.
typedef const __attribute__((address_space(4))) uint32_t CWord;
typedef CWord CRowRec;
constexpr int kRecWords = 12;
static_assert(sizeof(RowRec) == 4 * kRecWords, "RowRec is twelve words");

__device__ __forceinline__ RowRec load_rec(CWord* w) {
  RowRec R;
  R.a1 = (int32_t)w[0];
  R.a2 = (int32_t)w[1];
  R.cmin = (int32_t)w[2];
  R.n = w[3];
  R.inr_lo = (int32_t)w[4];
  R.inr_n = w[5];
  R.edge_l = (int32_t)w[6];
  R.edge_t = (int32_t)w[7];
  R.row_free = (int32_t)w[8];
  R.flags = w[9];
  R.act_lo = w[10];
  R.act_hi = w[11];
  return R;
}

// hp = 2 hp + (1 on the lanes of mask): one add with the mask as carry-in.
__device__ __forceinline__ uint32_t shift_in(uint32_t hp, uint64_t mask) {
  uint32_t r;
  uint64_t carry_out;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(hp), "s"(mask));
  return r;
}

// A kRowMasked row (see d2g_internal.h): nw_row's update with the window
// tests replaced by the record's lane masks.  Every active cell is inside
// both sequences (score = match / mismatch) and reads lanes of its own pair,
// so the values and pointers equal nw_row's: the same saturating i16
// candidates and dploop precedence for active lanes, the ends-free edge
// score (pointer 2 on the top edge, else 3 on the left) or fill elsewhere.
// Each condition becomes a lane mask where it is formed, and the pointer is
// shifted in as two mask bits (a diagonal over a mismatch is 0); on
// recalculation rows the pointer classes are lane masks too, updated by
// nw_row's take and tie rules on the recalculated columns.
template <bool EVEN, bool SWAP>
__device__ __forceinline__ void masked_row(const int row, const RowRec R, LaneState& ls,
                                           const RowConst K, const int ca, const int cb,
                                           const int c, const int lane, uint32_t* hist_p) {
  const bool eq = ca == cb;
  const uint64_t meq = __builtin_amdgcn_ballot_w64(eq);
  const int dg = clamp16(ls.prev2 + (eq ? K.match : K.mismatch));
  const int sh = EVEN ? shift_from_lower(ls.prev) : shift_from_upper(ls.prev);
  const int lraw = EVEN ? sh : ls.prev;
  const int uraw = EVEN ? ls.prev : sh;
  const int l = clamp16(lraw + K.gap), u = clamp16(uraw + K.gap);
  const bool upf = SWAP ? (l < u) : (u >= l);
  const uint64_t mup = __builtin_amdgcn_ballot_w64(upf);
  const int m = l > u ? l : u;
  const bool dw = dg > m;
  const uint64_t mdw = __builtin_amdgcn_ballot_w64(dw);
  const int val = dw ? dg : m;
  const uint64_t act = ((uint64_t)R.act_hi << 32) | R.act_lo;
  // ends-free edge lanes (no lane matches a column of -1)
  const uint64_t mt = __builtin_amdgcn_ballot_w64(c == R.edge_t);
  const uint64_t ml = __builtin_amdgcn_ballot_w64(c == R.edge_l);
  const int outside = __builtin_amdgcn_inverse_ballot_w64(mt | ml) ? R.row_free : K.fill;
  int e = __builtin_amdgcn_inverse_ballot_w64(act) ? val : outside;
  uint64_t hi, lo;
  if (R.flags & 3u) {
    // pointer classes as lane masks (p1: diagonal, before a mismatch makes
    // it 0; p2: left; p3: up; none: 0); the recalculated columns then follow
    // nw_row's take and tie rules
    uint64_t p1 = act & mdw;
    uint64_t p2 = (act & ~mdw & ~mup) | (~act & mt);
    uint64_t p3 = (act & ~mdw & mup) | (~act & ~mt & ml);
    const bool trans_min = (R.flags & 4u) != 0;
    if (R.flags & 1u) {
      const int cmin_new = (int)((R.flags >> 8) & 0xFFu);
      const int d_free = clamp16((trans_min ? uraw : lraw) + K.end_gap);
      const uint64_t here = __builtin_amdgcn_ballot_w64(c == cmin_new);
      const uint64_t take = here & __builtin_amdgcn_ballot_w64(d_free > e);
      const uint64_t tie = here & __builtin_amdgcn_ballot_w64(d_free == e) & (SWAP ? ~p2 : p1);
      e = __builtin_amdgcn_inverse_ballot_w64(take) ? d_free : e;
      const uint64_t upd = take | tie;
      p1 &= ~upd;
      p2 |= upd;
      p3 &= ~upd;
    }
    if (R.flags & 2u) {
      const int cmax_new = (int)((R.flags >> 16) & 0xFFu);
      const int src = trans_min ? shift_from_upper(uraw) : uraw;
      const int d_free = clamp16(src + K.end_gap);
      const uint64_t here = __builtin_amdgcn_ballot_w64(c == cmax_new);
      const uint64_t take = here & __builtin_amdgcn_ballot_w64(d_free > e);
      const uint64_t tie = here & __builtin_amdgcn_ballot_w64(d_free == e) & (SWAP ? p1 : ~p3);
      e = __builtin_amdgcn_inverse_ballot_w64(take) ? d_free : e;
      const uint64_t upd = take | tie;
      p1 &= ~upd;
      p2 &= ~upd;
      p3 |= upd;
    }
    hi = p2 | p3;
    lo = (p1 & meq) | p3;
  } else {
    hi = (act & ~mdw) | (~act & (mt | ml));
    lo = (act & ((mdw & meq) | (~mdw & mup))) | (~act & ml & ~mt);
  }
  ls.hp = shift_in(ls.hp, hi);
  ls.hp = shift_in(ls.hp, lo);
  if ((row & 15) == 15 || row == K.last) {
    hist_p[(row >> 4) * 64 + lane] = ls.hp << (2 * (15 - (row & 15)));
    ls.hp = 0;
  }
  ls.prev2 = ls.prev;
  ls.prev = e;
}

// Rows [r0, r1) through the full row update.  Software pipelined: the bases
// of row r+1 and the record of row r+2 are requested before row r is
// computed, so the one wait per row (LDS and scalar loads share a counter)
// finds both long arrived.  The schedule holds two lookahead records past
// the last row.
template <bool SWAP>
__device__ __forceinline__ void generic_rows(const int r0, const int r1, CRowRec* rec,
                                             LaneState& ls, const RowConst K, const uint8_t* s1l,
                                             const uint8_t* s2l, const int c, const int lane,
                                             uint32_t* hist_p) {
  if (r0 >= r1) return;
  RowRec cur = load_rec(rec + kRecWords * (r0 - 2));
  RowRec nxt = load_rec(rec + kRecWords * (r0 - 1));
  int ca = s1l[cur.a1], cb = s2l[cur.a2];
  for (int row = r0; row < r1; ++row) {
    const int ca1 = s1l[nxt.a1], cb1 = s2l[nxt.a2];
    const RowRec nx2 = load_rec(rec + kRecWords * row);  // row + 2
    if (row & 1)
      nw_row<false, SWAP>(row, cur, ls, K, ca, cb, c, lane, hist_p);
    else
      nw_row<true, SWAP>(row, cur, ls, K, ca, cb, c, lane, hist_p);
    cur = nxt;
    nxt = nx2;
    ca = ca1;
    cb = cb1;
  }
}

// The same rows of a bin whose rows are all kRowMasked, unrolled by row
// pairs: r0 is even (row 2, or the end of a steady run, which is even),
// so each pair is an even row and then an odd one, each with its own code
// and no branch between them (merged, the two parities' compares would
// leave their lane masks).  Loads are pipelined as in generic_rows.
template <bool SWAP>
__device__ __forceinline__ void masked_rows(const int r0, const int r1, CRowRec* rec,
                                            LaneState& ls, const RowConst K, const uint8_t* s1l,
                                            const uint8_t* s2l, const int c, const int lane,
                                            uint32_t* hist_p) {
  if (r0 >= r1) return;
  RowRec cur = load_rec(rec + kRecWords * (r0 - 2));
  RowRec nxt = load_rec(rec + kRecWords * (r0 - 1));
  int ca = s1l[cur.a1], cb = s2l[cur.a2];
  for (int row = r0; row < r1; row += 2) {
    {
      const int ca1 = s1l[nxt.a1], cb1 = s2l[nxt.a2];
      const RowRec nx2 = load_rec(rec + kRecWords * row);  // row + 2
      masked_row<true, SWAP>(row, cur, ls, K, ca, cb, c, lane, hist_p);
      cur = nxt;
      nxt = nx2;
      ca = ca1;
      cb = cb1;
    }
    if (row + 1 >= r1) break;
    {
      const int ca1 = s1l[nxt.a1], cb1 = s2l[nxt.a2];
      const RowRec nx2 = load_rec(rec + kRecWords * (row + 1));  // row + 3
      masked_row<false, SWAP>(row + 1, cur, ls, K, ca, cb, c, lane, hist_p);
      cur = nxt;
      nxt = nx2;
      ca = ca1;
      cb = cb1;
    }
  }
}


// A steady row: no ends-free edge or recalculation, every active cell
// inside both sequences and off its pair's outer columns, with an active
// diagonal predecessor.  nw_row reduced to what such a row can reach:
//  - act (the schedule's lane mask) sends inactive lanes to fill, and their
//    pointer bits are never read: no path of such a row enters them;
//  - without CLAMP every candidate is an active cell's score or fill plus
//    one increment, inside i16 by the launch's score bound, so the clamps
//    are the identity.  The run then keeps every score of row r offset by
//    -r gap (steady_rows converts on entry and exit): each candidate of a
//    row carries the same offset, so every comparison, pointer and maximum
//    is unchanged, a gap move adds nothing and a diagonal adds score - 2 gap.
//    Real cells stay at or above 0 there (no move costs more than -gap per
//    row) and fill far below, so a fill-derived candidate still never wins
//    or ties.  With the gap add gone, max(left, up) takes the shifted
//    neighbour as a DPP operand, and up-versus-left is read off it against
//    the unshifted one (or off the min on the other parity).
// The pointer (dploop precedence, nwalign.rs:700-737; a diagonal over a
// mismatch is 0) is shifted into hp as two bits.
template <bool EVEN, bool SWAP, bool CLAMP>
__device__ __forceinline__ void steady_cell(const int ca, const int cb, const uint64_t act,
                                            LaneState& ls, const RowConst K, uint32_t& hp) {
  const bool eq = ca == cb;
  int dg;
  const int sh = EVEN ? shift_from_lower(ls.prev) : shift_from_upper(ls.prev);
  bool upf;
  int m;
  if (CLAMP) {
    dg = clamp16(ls.prev2 + (eq ? K.match : K.mismatch));
    const int lraw = EVEN ? sh : ls.prev;
    const int uraw = EVEN ? ls.prev : sh;
    const int l = clamp16(lraw + K.gap), u = clamp16(uraw + K.gap);
    upf = SWAP ? (l < u) : (u >= l);
    m = l > u ? l : u;
  } else {
    dg = ls.prev2 + (eq ? K.dmatch : K.dmis);
    m = max(sh, ls.prev);
    // left/up: (sh, prev) on even rows, (prev, sh) on odd rows; upf is
    // u >= l, or l < u with SWAP
    if (EVEN)
      upf = SWAP ? min(sh, ls.prev) != ls.prev : m == ls.prev;
    else
      upf = SWAP ? m != ls.prev : min(sh, ls.prev) == ls.prev;
  }
  const bool dw = dg > m;
  const int val = dw ? dg : m;
  ls.prev2 = ls.prev;
  ls.prev = __builtin_amdgcn_inverse_ballot_w64(act) ? val : K.fill;
  // pointer bits as lane masks: hi = !dw, lo = dw ? eq : upf
  const uint64_t mdw = __builtin_amdgcn_ballot_w64(dw);
  const uint64_t meq = __builtin_amdgcn_ballot_w64(eq);
  const uint64_t mup = __builtin_amdgcn_ballot_w64(upf);
  hp = shift_in(hp, ~mdw);
  hp = shift_in(hp, (mdw & meq) | (~mdw & mup));
}

// Steady rows [r0, r1) inside one history block (r0 and r1 even), by row
// pairs: the pointer word continues the one in ls.hp and is flushed when r1
// ends the block.  p1 / p2 point at row r0's bases; the next pair's bases are
// read one pair ahead (the reads past the run stay in the padded rows).
template <bool SWAP, bool CLAMP>
__device__ __forceinline__ void steady_part(const int r0, const int r1, const uint8_t* p1,
                                            const uint8_t* p2, const uint64_t act_even,
                                            const uint64_t act_odd, LaneState& ls,
                                            const RowConst K, const int lane, uint32_t* hist_p) {
  if (r0 >= r1) return;
  uint32_t hp = ls.hp;
  int ca = p1[0], cb = p2[0], cb1 = p2[1];
  for (int row = r0; row < r1; row += 2) {
    const int na = p1[1], nb1 = p2[2];
    steady_cell<true, SWAP, CLAMP>(ca, cb, act_even, ls, K, hp);
    steady_cell<false, SWAP, CLAMP>(ca, cb1, act_odd, ls, K, hp);
    ca = na;
    cb = cb1;
    cb1 = nb1;
    ++p1;
    ++p2;
  }
  if ((r1 & 15) == 0) {
    hist_p[((r1 - 1) >> 4) * 64 + lane] = hp;
    hp = 0;
  }
  ls.hp = hp;
}

// Steady rows [b0, b1) (both even).  Within a run the s1 offset advances
// after odd rows and the s2 offset after even rows, so row b0 + k reads s1
// at a1(b0) + k/2 and s2 at a2(b0) + (k+1)/2, and the active lanes of every
// even / odd row are those of row b0 / b0 + 1: nothing is read from the
// schedule inside the run.  Whole 16-row history blocks run fully unrolled
// (immediate LDS offsets from per-block bases, one history word each); the
// rows before the first whole block and after the last go by row pairs.
template <bool SWAP, bool CLAMP>
__device__ __forceinline__ void steady_rows(const int b0, const int b1, CRowRec* rec,
                                            LaneState& ls, const RowConst K, const uint8_t* s1l,
                                            const uint8_t* s2l, const int lane, uint32_t* hist_p) {
  if (b0 >= b1) return;
  CWord* r0 = rec + kRecWords * (b0 - 2);
  const uint8_t* p1 = s1l + (int)r0[0];
  const uint8_t* p2 = s2l + (int)r0[1];
  const uint64_t act_even = ((uint64_t)r0[11] << 32) | r0[10];
  const uint64_t act_odd = ((uint64_t)r0[kRecWords + 11] << 32) | r0[kRecWords + 10];
  if (!CLAMP) {  // into the offset domain: row r's scores minus r gap
    ls.prev -= (b0 - 1) * K.gap;
    ls.prev2 -= (b0 - 2) * K.gap;
  }
  const int f0 = imin((b0 + 15) & ~15, b1);  // the first whole block
  const int f1 = imax(b1 & ~15, f0);         // the end of the whole blocks
  steady_part<SWAP, CLAMP>(b0, f0, p1, p2, act_even, act_odd, ls, K, lane, hist_p);
  p1 += (f0 - b0) >> 1;
  p2 += (f0 - b0) >> 1;
  for (int row = f0; row < f1; row += 16) {
    uint32_t hp = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int ca = p1[k >> 1], cb = p2[(k + 1) >> 1];
      if (k & 1)
        steady_cell<false, SWAP, CLAMP>(ca, cb, act_odd, ls, K, hp);
      else
        steady_cell<true, SWAP, CLAMP>(ca, cb, act_even, ls, K, hp);
    }
    hist_p[(row >> 4) * 64 + lane] = hp;
    p1 += 8;
    p2 += 8;
  }
  steady_part<SWAP, CLAMP>(f1, b1, p1, p2, act_even, act_odd, ls, K, lane, hist_p);
  if (!CLAMP) {
    // back from the offset domain: rows b1 - 1 (odd) and b1 - 2 (even);
    // their inactive lanes hold fill, as the plain update leaves them
    ls.prev = __builtin_amdgcn_inverse_ballot_w64(act_odd) ? ls.prev + (b1 - 1) * K.gap : K.fill;
    ls.prev2 = __builtin_amdgcn_inverse_ballot_w64(act_even) ? ls.prev2 + (b1 - 2) * K.gap : K.fill;
  }
}

template <bool SWAP, bool CLAMP>
__device__ __forceinline__ void forward(CRowRec* rec, const int b0, const int b1,
                                        const bool masked, LaneState& ls, const RowConst K,
                                        const uint8_t* s1l, const uint8_t* s2l, const int c,
                                        const int lane, uint32_t* hist_p) {
  if (masked) {
    masked_rows<SWAP>(2, b0, rec, ls, K, s1l, s2l, c, lane, hist_p);
    steady_rows<SWAP, CLAMP>(b0, b1, rec, ls, K, s1l, s2l, lane, hist_p);
    masked_rows<SWAP>(b1, K.last + 1, rec, ls, K, s1l, s2l, c, lane, hist_p);
  } else {
    generic_rows<SWAP>(2, b0, rec, ls, K, s1l, s2l, c, lane, hist_p);
    steady_rows<SWAP, CLAMP>(b0, b1, rec, ls, K, s1l, s2l, lane, hist_p);
    generic_rows<SWAP>(b1, K.last + 1, rec, ls, K, s1l, s2l, c, lane, hist_p);
  }
}

// ---------------------------------------------------------------------------
// Traceback
// ---------------------------------------------------------------------------
// Per-wave constants of the traceback.
struct BwConst {
  CRowRec* rec;
  const uint32_t* rec_v;  // the same records, for per-lane (vector) loads
  int gpos;  // lane of column 0 of this lane's pair
  int c, start_col;
  int st_lo;     // the bin's steady rows: [st_lo, st_lo + st_n)
  uint32_t st_n;
  bool swap, real, sub_ok;
  size_t my_sub;  // first substitution slot of this lane's pair
  const uint8_t* s1b;
  const uint8_t* s2b;
  uint64_t ao;  // alignment string offset (parity mode)
  int ncol;     // lanes of a pair (alignment-output mode writes a run's bytes over them)
};
// Path state of this lane's pair, replicated over the pair's lanes.
struct PathState {
  int rc;    // row of the path cell; -1 once the path has ended
  int addr;  // 4 x the lane owning the path cell's column (a ds_bpermute address)
  int ns, nt;
  bool bad;
};

// The traceback (nwalign.rs:1101-1139) by runs.  The path starts at the
// last cell and takes pointer p of its cell: 1 a diagonal (i-1, j-1), two
// rows up in the same compressed column; 2 left (i, j-1) and 3 up (i-1, j),
// one row up, i.e. one column down on even rows / up on odd rows; 0 a
// mismatching diagonal on an active cell (the al2subs substitution), else
// the reference's invalid pointer, which ends the path with the error flag.
// Each lane fetches its pair's path word from the lane owning the column
// (ds_bpermute).  Within one 16-row history word the pointers of rows r,
// r-2, r-4, ... of a column sit 4 bits apart: shifted to row r, a run of
// diagonal-over-match pointers (1) is the count of trailing zero fields of
// (t ^ 0x11111111) & 0x33333333.  Each iteration moves every pair's path
// over its run in the current block and then takes at most one other step,
// so a block costs one iteration per non-run step of the wavefront's
// busiest pair instead of one per row.
//
// Compare mode records each substitution (query position << 16 | center
// position) from the lane owning the path column.  Alignment-output mode
// (EMIT) writes the alignment strings in path order (reversed afterwards):
// a run's n diagonal columns hold s1[i-1-k], s2[j-1-k] (k < n), written by
// the pair's lanes in parallel; a gap or a mismatching diagonal writes one
// column from the pair's first lane.
//
// Block b of the history (rows 16b .. 16b+15), word hp of this lane.
template <bool EMIT>
__device__ __forceinline__ void runs_block(const int b, const uint32_t hp, PathState& ps,
                                           const BwConst& B, const NwDev& a) {
  const int base = b << 4;
  const int lo = b == 0 ? 1 : base;  // row 0 is the path's end
  for (;;) {
    // the path cell is in this block (rc <= base + 15 holds: rc starts at
    // last and leaves a block only downwards)
    const bool on = ps.rc >= lo;
    if (!__builtin_amdgcn_ballot_w64(on)) break;
    const uint32_t w = (uint32_t)__builtin_amdgcn_ds_bpermute(ps.addr, (int)hp);
    const int k = ps.rc - base;
    // field j of t: the pointer of row rc - 2j; fields past row 0 / 1 of
    // the block are zero, which ends a run
    const uint32_t t = w >> ((30 - 2 * k) & 31);
    const uint32_t m = (t ^ 0x11111111u) & 0x33333333u;
    // the lowest set bit sits at 4j or 4j + 1 of the first non-1 field j;
    // bit 31 stands for field 8 (all 8 fields 1, k = 15)
    const int n = (int)((__builtin_ctz(m | 0x80000000u) + 1u) >> 2);
    const int r = ps.rc - 2 * n;  // first cell after the run
    const bool step = on && r >= lo;  // a non-run pointer inside the block
    if (EMIT && on && n > 0) {
      // the run's columns: cell (i0 - k, j0 - k) for k < n, one byte each
      const int cc = (ps.addr >> 2) - B.gpos;
      const int dd = 2 * (cc - B.start_col) + (ps.rc & 1);  // j - i
      const int i0 = (ps.rc - dd) >> 1, j0 = (ps.rc + dd) >> 1;
      if (B.real) {
        for (int k = B.c; k < n; k += B.ncol) {
          const uint8_t x0 = B.s1b[i0 - 1 - k], x1 = B.s2b[j0 - 1 - k];
          a.al0[B.ao + ps.nt + k] = B.swap ? x1 : x0;
          a.al1[B.ao + ps.nt + k] = B.swap ? x0 : x1;
        }
      }
      ps.nt += n;
    }
    if (!__builtin_amdgcn_ballot_w64(step)) {
      // the common case: every path runs out of the block
      ps.rc = on ? r : ps.rc;
      break;
    }
    const int pp = (int)((t >> ((4 * n) & 31)) & 3u);  // 0, 2 or 3 there (a run stops at non-1)
    const bool odd = (r & 1) != 0;
    if (EMIT && step && pp != 0) {
      // a gap column at row r, in the path's column before the move: left
      // (2) pairs a gap with s2[j-1], up (3) s1[i-1] with a gap
      const int cc = (ps.addr >> 2) - B.gpos;
      const int dd = 2 * (cc - B.start_col) + (odd ? 1 : 0);
      const int i = (r - dd) >> 1, j = (r + dd) >> 1;
      if (B.real && B.c == 0) {
        const uint8_t x0 = pp == 2 ? (uint8_t)'-' : B.s1b[i - 1];
        const uint8_t x1 = pp == 3 ? (uint8_t)'-' : B.s2b[j - 1];
        a.al0[B.ao + ps.nt] = B.swap ? x1 : x0;
        a.al1[B.ao + ps.nt] = B.swap ? x0 : x1;
      }
      ps.nt += 1;
    }
    // gaps: up (3) one column up on odd rows, left (2) one column down on
    // even rows; both one row up.  Lanes that only ran move to r.
    const bool gap = step && pp != 0;
    const int delta = odd ? (pp == 3 ? 4 : 0) : (pp == 2 ? -4 : 0);
    ps.addr += gap ? delta : 0;
    ps.rc = on ? (gap ? r - 1 : r) : ps.rc;
    if (__builtin_amdgcn_ballot_w64(step && pp == 0)) {
      // pointer 0: a mismatching diagonal on an active cell, else invalid.
      // A path cell of a steady row is active (no path enters an inactive
      // lane there), so only rows outside the steady run read the record.
      const bool zero = step && pp == 0;
      const int cc = (ps.addr >> 2) - B.gpos;
      bool act = (unsigned)(r - B.st_lo) < B.st_n;
      if (zero && !act && r >= 2) {
        const int rcmin = (int)B.rec_v[kRecWords * (r - 2) + 2];
        const uint32_t rn = B.rec_v[kRecWords * (r - 2) + 3];
        act = (unsigned)(cc - rcmin) < rn;
      }
      const bool sub = zero && act;
      const bool invalid = zero && !act;
      const int dd = 2 * (cc - B.start_col) + (odd ? 1 : 0);  // j - i on row r
      const int i = (r - dd) >> 1, j = (r + dd) >> 1;
      if (EMIT) {
        // a mismatching diagonal column: both bases
        if (sub && B.c == 0 && B.real) {
          const uint8_t x0 = B.s1b[i - 1], x1 = B.s2b[j - 1];
          a.al0[B.ao + ps.nt] = B.swap ? x1 : x0;
          a.al1[B.ao + ps.nt] = B.swap ? x0 : x1;
        }
        ps.nt += sub ? 1 : 0;
      } else if (sub && B.c == cc && B.real && B.sub_ok) {
        const int qp = B.swap ? i - 1 : j - 1;
        const int cp = B.swap ? j - 1 : i - 1;
        a.sub_pool[B.my_sub + ps.ns] = ((uint32_t)qp << 16) | (uint32_t)cp;
      }
      ps.ns += sub ? 1 : 0;
      ps.rc = sub ? r - 2 : (invalid ? -1 : ps.rc);
      ps.bad = ps.bad || invalid;
    }
  }
}

// Blocks last/16 .. 0.  Four history words rotate, each reloaded (for the
// block four below) as soon as its block is done, so every load has three
// blocks of work to arrive in (a lone wavefront's blocks are short next to
// an L2 round trip) and no register copy waits for it.
template <bool EMIT>
__device__ __forceinline__ void backward_runs(const int last, const uint32_t* hist_p,
                                              const int lane, PathState& ps, const BwConst& B,
                                              const NwDev& a) {
  int b = last >> 4;
  uint32_t h0 = hist_p[b * 64 + lane];
  uint32_t h1 = hist_p[imax(b - 1, 0) * 64 + lane];
  uint32_t h2 = hist_p[imax(b - 2, 0) * 64 + lane];
  uint32_t h3 = hist_p[imax(b - 3, 0) * 64 + lane];
  for (;;) {
    runs_block<EMIT>(b, h0, ps, B, a);
    if (b == 0) break;
    h0 = hist_p[imax(b - 4, 0) * 64 + lane];
    --b;
    runs_block<EMIT>(b, h1, ps, B, a);
    if (b == 0) break;
    h1 = hist_p[imax(b - 4, 0) * 64 + lane];
    --b;
    runs_block<EMIT>(b, h2, ps, B, a);
    if (b == 0) break;
    h2 = hist_p[imax(b - 4, 0) * 64 + lane];
    --b;
    runs_block<EMIT>(b, h3, ps, B, a);
    if (b == 0) break;
    h3 = hist_p[imax(b - 4, 0) * 64 + lane];
    --b;
  }
}

constexpr int kStage = 4;  // sequence bytes per lane requested together when staging a pair

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void nw_vec_kernel(NwDev a) {
  extern __shared__ uint32_t lds_words[];
  const int lane = (int)__lane_id();
  const BinSpace bs = a.bins;
  // wavefronts are assigned statically (the grid covers the usual batch in
  // one pass; pairs of one launch cost about the same)
  const uint32_t total = (uint32_t)uni((int)a.ctr->total_waves);
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    // bin of wavefront w: last b with wave_start[b] <= w
    int lo = 0, hi = bs.nbins;  // invariant wave_start[lo] <= w < wave_start[hi]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (a.wave_start[mid] <= w) lo = mid; else hi = mid;
    }
    const int bin = uni(lo);
    int l1_in, l2_in;
    bin_lengths(bs, (uint32_t)bin, &l1_in, &l2_in);
    const VecGeom geo = vec_geom(l1_in, l2_in, bs.band);
    const int G = bin_groups(bs, l1_in, l2_in);
    const uint32_t wave_in_bin = w - a.wave_start[bin];
    const uint32_t first = a.bin_start[bin] + wave_in_bin * (uint32_t)G;
    const uint32_t bin_end = a.bin_start[bin + 1];
    const int count = uni((int)min((uint32_t)G, bin_end - first));

    const int len1 = geo.len1, len2 = geo.len2, ncol = geo.ncol;
    const int start_col = geo.start_col;
    const bool swap = geo.swap != 0;
    const int g_raw = lane / ncol;
    const int c = lane - g_raw * ncol;
    const bool real = g_raw < count;
    const int g = real ? g_raw : 0;  // idle lanes shadow group 0
    const DpTask task = a.tasks[first + (uint32_t)g];

    // ---- pointer history in this workgroup's global slot; LDS holds the
    // padded sequences of every pair ----
    uint32_t* hist_p = a.hist + (size_t)blockIdx.x * a.hist_stride;
    uint8_t* gbase0 = reinterpret_cast<uint8_t*>(lds_words);
    const int pad = nw_pad(ncol);
    const size_t gbytes = nw_group_bytes(geo);
    uint8_t* gbase = gbase0 + (size_t)g * gbytes;
    uint8_t* s1b = gbase + pad;
    uint8_t* s2b = gbase + pad + len1 + 2 * pad;
    if (real) {
      // kStage bytes of each sequence per lane are requested before any is
      // stored (indices clamped into the sequence, so every load is in
      // bounds): a 250-nt pair takes four round trips, not eight (more in
      // flight would spill at 64 registers)
      // (the pool is below 2^32 bytes: 32-bit offsets from its uniform base)
      const int lmax = imax(len1, len2);
      const uint32_t o1 = (uint32_t)(swap ? task.s2_off : task.s1_off);
      const uint32_t o2 = (uint32_t)(swap ? task.s1_off : task.s2_off);
      for (int t0 = c; t0 < lmax; t0 += kStage * ncol) {
        uint32_t b1[kStage], b2[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          const int t = t0 + k * ncol;
          b1[k] = a.pool[o1 + (uint32_t)imin(t, len1 - 1)];
          b2[k] = a.pool[o2 + (uint32_t)imin(t, len2 - 1)];
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          const int t = t0 + k * ncol;
          if (t < len1) s1b[t] = (uint8_t)b1[k];
          if (t < len2) s2b[t] = (uint8_t)b2[k];
        }
      }
    }
    __syncthreads();

    // ---- scoring constants (nwalign.rs:817-818) ----
    RowConst K;
    K.match = a.match;
    K.mismatch = a.mismatch;
    K.gap = a.gap;
    K.end_gap = a.end_gap;
    int min_score = K.mismatch < K.gap ? K.mismatch : K.gap;
    min_score = K.match < min_score ? K.match : min_score;
    min_score = min_score < 0 ? min_score : 0;
    K.fill = (int)(int16_t)(uint16_t)(uint32_t)(-32768 - min_score);
    K.last = len1 + len2;
    K.dmatch = K.match - 2 * K.gap;
    K.dmis = K.mismatch - 2 * K.gap;
    const int last = K.last;

    // ---- rows 0 and 1 (nwalign.rs:842-861) ----
    LaneState ls;
    ls.prev2 = (c == start_col) ? 0 : K.fill;
    {
      const bool isL = 1 < geo.left_fill_limit && c == imax(start_col - 1, 0);
      const bool isT = 1 < geo.top_fill_limit && c == start_col;
      ls.prev = (isL || isT) ? K.end_gap : K.fill;
      const int p1 = isT ? 2 : (isL ? 3 : 0);
      ls.hp = (uint32_t)p1;  // rows 0 (only 0 pointers) and 1 shifted in
    }
    if (last < 2) hist_p[lane] = ls.hp << 28;

    // ---- forward: anti-diagonals 2..len1+len2 ----
    CRowRec* rec = (CRowRec*)(a.sched + a.sched_off[bin]);
    const BinInfo bi = a.info[bin];
    const int b0 = uni((int)bi.steady_lo), b1 = uni((int)bi.steady_hi);
    const bool masked = uni((int)bi.masked) != 0;
    const uint8_t* s1l = s1b - c;
    const uint8_t* s2l = s2b + c;
    if (a.steady_clamp) {
      if (swap)
        forward<true, true>(rec, b0, b1, masked, ls, K, s1l, s2l, c, lane, hist_p);
      else
        forward<false, true>(rec, b0, b1, masked, ls, K, s1l, s2l, c, lane, hist_p);
    } else {
      if (swap)
        forward<true, false>(rec, b0, b1, masked, ls, K, s1l, s2l, c, lane, hist_p);
      else
        forward<false, false>(rec, b0, b1, masked, ls, K, s1l, s2l, c, lane, hist_p);
    }
    __syncthreads();

    // ---- backward sweep: every pair's traceback (nwalign.rs:1101-1139) ----
    // Substitution records go straight to the pool: sorted task ti owns
    // sub_stride slots (more than its diagonal steps) at ti * sub_stride.
    const uint32_t ti = first + (uint32_t)g;
    const size_t sub0 = (size_t)ti * a.sub_stride;
    const bool sub_ok = a.emit || sub0 + a.sub_stride <= a.sub_capacity;
    if (!a.emit && __builtin_amdgcn_ballot_w64(!sub_ok && real) && lane == 0)
      atomicOr(&a.ctr->err_flags, kErrSubOverflow);
    BwConst B;
    B.rec = rec;
    B.rec_v = reinterpret_cast<const uint32_t*>(a.sched + a.sched_off[bin]);
    B.gpos = g_raw * ncol;
    B.ncol = ncol;
    B.c = c;
    B.start_col = start_col;
    B.st_lo = b0;
    B.st_n = (uint32_t)(b1 - b0);
    B.swap = swap;
    B.real = real;
    B.sub_ok = sub_ok;
    B.my_sub = sub_ok ? sub0 : 0;
    B.s1b = s1b;
    B.s2b = s2b;
    B.ao = (a.emit && real) ? a.al_off[task.out_idx] : 0;
    PathState ps;
    ps.rc = real ? last : -1;
    ps.addr = (B.gpos + ((2 * start_col + len2 - len1) >> 1)) * 4;
    ps.ns = 0;
    ps.nt = 0;
    ps.bad = false;
    if (a.emit)
      backward_runs<true>(last, hist_p, lane, ps, B, a);
    else
      backward_runs<false>(last, hist_p, lane, ps, B, a);
    const int ns = ps.ns, nt = ps.nt;
    const bool bad = ps.bad;
    if (__builtin_amdgcn_ballot_w64(bad && real) && lane == 0)
      atomicOr(&a.ctr->err_flags, kErrTraceback);
    __syncthreads();

    // ---- outputs (column 0 lane of each pair) ----
    if (a.emit) {
      if (real) {
        uint8_t* o0 = a.al0 + B.ao;
        uint8_t* o1 = a.al1 + B.ao;
        for (int x = c; x < nt / 2; x += ncol) {
          const int y = nt - 1 - x;
          const uint8_t t0 = o0[x], t1 = o1[x];
          o0[x] = o0[y];
          o1[x] = o1[y];
          o0[y] = t0;
          o1[y] = t1;
        }
        if (c == 0) a.al_len[task.out_idx] = bad ? 0u : (uint32_t)nt;
      }
    } else if (real && c == 0) {
      a.sub_cnt[ti] = sub_ok ? ((uint32_t)ns | (bad ? 0x80000000u : 0u)) : 0u;
    }
    __syncthreads();
  }
}

// Whether a steady row's i16 clamps can change a result for this bin space:
// a cell (i, j) of a banded alignment scores within [-(i + j) pen, min(i, j)
// match] (pen the largest penalty), so when that range, widened by one
// increment, fits i16 every clamp of a real score is the identity (fill plus
// one increment never saturates by the choice of fill).
bool steady_needs_clamp(const BinSpace& bs, int match, int mismatch, int gap, int end_gap) {
  const long long lmax = (long long)bs.lmin + bs.range - 1;
  long long pen = 0;
  for (int x : {mismatch, gap, end_gap, match}) pen = std::max(pen, (long long)-x);
  const long long up = std::max(0, match);
  const long long lo_bound = (2 * lmax + 1) * pen;
  const long long hi_bound = (lmax + 1) * up;
  return lo_bound > 32768 - 64 || hi_bound > 32767 - 64;
}

hipError_t launch_nw_vec(const NwArgs& s, hipStream_t st) {
  NwDev d;
  d.bins = s.bins;
  d.bin_start = s.bin_start;
  d.wave_start = s.wave_start;
  d.tasks = s.tasks;
  d.pool = s.pool;
  d.sched = s.sched;
  d.info = s.info;
  d.sched_off = s.sched_off;
  d.match = s.match;
  d.mismatch = s.mismatch;
  d.gap = s.gap_p;
  d.end_gap = s.end_gap_p;
  d.steady_clamp = steady_needs_clamp(s.bins, s.match, s.mismatch, s.gap_p, s.end_gap_p) ? 1 : 0;
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
  if (s.grid == 0 || s.hist == nullptr || s.sched == nullptr || s.sched_off == nullptr)
    return hipErrorInvalidValue;
  if (!s.emit_alignment && s.sub_stride == 0) return hipErrorInvalidValue;
  const size_t lds = s.lds_bytes;
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)nw_vec_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(nw_vec_kernel, dim3(s.grid), dim3(64), lds, st, d);
  return hipGetLastError();
}

// Dynamic LDS of a launch: the largest wavefront footprint over every bin of
// the space (bins are not known on the host, only their range).
size_t nw_vec_lds_for(const BinSpace& bs, const uint8_t* present) {
  size_t lds = 0;
  for (int b = 0; b < bs.nbins; ++b) {
    int l1, l2;
    bin_lengths(bs, (uint32_t)b, &l1, &l2);
    if (!bin_present(bs, present, l1, l2)) continue;
    VecGeom g = vec_geom(l1, l2, bs.band);
    int G = nw_groups(g, bs.tile);
    if (G > 0) {
      size_t need = nw_wave_bytes(g, G);
      if (need > lds) lds = need;
    }
  }
  lds = (lds + 255) & ~(size_t)255;
  return lds < 256 ? 256 : lds;
}

}  // namespace d2g
