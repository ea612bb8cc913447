// nw_vec.hip — the compressed anti-diagonal banded Needleman-Wunsch of
// align_vectorized_with_buf (nwalign.rs:761-1140), bit-exact, on gfx950.
//
// Mapping.  The reference stores cell (i, j) of anti-diagonal row = i + j in
// compressed column (2*start_col + j - i) / 2 of an ncol-wide row
// (nwalign.rs:783-787).  Here a wavefront holds G = 64 / ncol pairs; lane c
// of group g owns column c of pair g for every row, so a column's values
// stay in registers: prev2 (row-2, the diagonal source) and prev (row-1).
// The left/up operands (nwalign.rs:957-964) are column c-1/c on even rows
// and c/c+1 on odd rows: a one-lane DPP wavefront shift.  Columns 0 and
// ncol-1 never become active (their lanes hold the sentinel), so a shift
// across a group boundary is never consumed.
//
// The row schedule (col_min, col_max, i_max, j_min, band wedges, ends-free
// recalculation flags; nwalign.rs:865-1088) depends only on (len1, len2,
// band); every pair of a wavefront shares them, so it runs in scalar
// registers.  Saturating i16 arithmetic is emulated on i32 with clamps.
//
// Traceback pointers (2 bits per cell) are ballots: two 64-bit masks per row
// collected one row per lane and flushed to LDS every 64 rows as one 16-byte
// record per row.  One lane per pair then walks the pointers
// (nwalign.rs:1101-1139) and either emits the alignment strings (parity
// mode) or the substitution list al2subs would build (nwalign.rs:1169-1224),
// as (query position << 8 | 4*(center nt-1) + (query nt-1)) records for the
// lambda kernel.
#include "d2g_launch.h"

namespace d2g {

struct NwDev {
  BinSpace bins;
  const uint32_t* bin_start;
  const uint32_t* wave_start;
  const DpTask* tasks;
  const uint8_t* pool;
  int match, mismatch, gap, end_gap;
  Counters* ctr;
  int emit;
  uint32_t* sub_pool;
  uint32_t sub_capacity;
  uint32_t* sub_top;
  uint32_t* sub_off;
  uint32_t* sub_cnt;
  uint8_t* al0;
  uint8_t* al1;
  const uint64_t* al_off;
  uint32_t* al_len;
};

// lane i <- lane i-1 (DPP wave_shr:1); lane 0 keeps `old`
__device__ __forceinline__ int from_lower(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}
// lane i <- lane i+1 (DPP wave_shl:1); lane 63 keeps `old`
__device__ __forceinline__ int from_upper(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Row constants of one wavefront (uniform).
struct RowConst {
  int len1, len2, start_col, lfl, tfl, top_limit, band_min_len1, band_mod2, full_band, last, band;
  int match, mismatch, gap, end_gap, fill;
  bool swap, do_recalc;
};
// Row walk state (uniform).
struct RowState {
  int col_min, col_max, i_max, j_min, row_free;
  bool recalc_left, recalc_right;
  uint64_t cells;
};
// Per-lane column state.
struct LaneState {
  int prev2, prev, acc0, acc1, acc2, acc3;
};

// One anti-diagonal (nwalign.rs:875-1088) for every pair of the wavefront.
template <bool EVEN>
__device__ __forceinline__ void nw_row(const int row, RowState& st, LaneState& ls, const RowConst K,
                                       const uint8_t* s1b, const uint8_t* s2b, const int c,
                                       const int lane, uint4* lds_ptr) {
  const int n = st.col_max - st.col_min + 1;
  st.row_free = sat16(st.row_free + K.end_gap);
  const int edgeL = row < K.lfl ? imax(K.start_col - 1 - (row - 1) / 2, 0) : -1;
  const int edgeT = row < K.tfl ? K.start_col + row / 2 : -1;
  int k_lo = st.i_max - K.len1 + 1;
  k_lo = k_lo < 0 ? 0 : (k_lo > n ? n : k_lo);
  int k_hi = imin(st.i_max + 1, K.len2 - st.j_min);
  k_hi = k_hi < 0 ? 0 : (k_hi > n ? n : k_hi);
  if (k_hi < k_lo) k_hi = k_lo;
  st.cells += (uint64_t)n;

  // cell (i, j) of this lane: si = i_max - k, sj = j_min + k (nwalign.rs:922-953)
  const int k = c - st.col_min;
  const int ca = s1b[st.i_max - k];
  const int cb = s2b[st.j_min + k];
  const bool inr = (unsigned)(k - k_lo) < (unsigned)(k_hi - k_lo);
  const int score = inr ? (ca == cb ? K.match : K.mismatch) : K.fill;
  const int dg = sat16(ls.prev2 + score);
  const int lraw = EVEN ? from_lower(ls.prev, K.fill) : ls.prev;
  const int uraw = EVEN ? ls.prev : from_upper(ls.prev, K.fill);
  const int l = sat16(lraw + K.gap), u = sat16(uraw + K.gap);
  // dploop precedence (nwalign.rs:700-737)
  const bool upf = K.swap ? (l < u) : (u >= l);
  int e = upf ? u : l;
  int p = upf ? 3 : 2;
  if (dg > e) {
    e = dg;
    p = 1;
  }
  if ((unsigned)k >= (unsigned)n) {
    // outside [col_min, col_max]: sentinel or ends-free edge (nwalign.rs:884-899)
    const bool isT = c == edgeT, isL = c == edgeL;
    e = (isT || isL) ? st.row_free : K.fill;
    p = isT ? 2 : (isL ? 3 : 0);
  }
  // band transition (nwalign.rs:980-987)
  const bool trans_min = row == K.band_min_len1;
  if (trans_min) {
    st.col_min = imax(st.col_min - 1, 0);
    st.i_max += 1;
    st.j_min -= 1;
  }
  if (row == K.top_limit) st.col_max += 1;
  // ends-free recalculation (nwalign.rs:993-1026)
  if (K.do_recalc) {
    const bool rl = st.recalc_left;
    if (st.i_max == K.len1 - 1) st.recalc_left = true;
    const bool rr = st.recalc_right;
    if ((row + 1) / 2 + st.col_max - K.start_col == K.len2) st.recalc_right = true;
    if (rl) {
      const int d_free = sat16((trans_min ? uraw : lraw) + K.end_gap);
      if (c == st.col_min) {
        if (d_free > e) {
          e = d_free;
          p = 2;
        } else if (d_free == e && (K.swap ? p != 2 : p == 1)) {
          p = 2;
        }
      }
    }
    if (rr) {
      const int src = trans_min ? from_upper(uraw, K.fill) : uraw;
      const int d_free = sat16(src + K.end_gap);
      if (c == st.col_max) {
        if (d_free > e) {
          e = d_free;
          p = 3;
        } else if (d_free == e && (K.swap ? p == 1 : p != 3)) {
          p = 3;
        }
      }
    }
  }
  // pointer record of this row
  const uint64_t b0 = __builtin_amdgcn_ballot_w64((p & 1) != 0);
  const uint64_t b1 = __builtin_amdgcn_ballot_w64((p & 2) != 0);
  const int slot = row & 63;
  const bool mine = lane == slot;
  ls.acc0 = mine ? (int)(uint32_t)b0 : ls.acc0;
  ls.acc1 = mine ? (int)(uint32_t)b1 : ls.acc1;
  ls.acc2 = mine ? (int)(uint32_t)(b0 >> 32) : ls.acc2;
  ls.acc3 = mine ? (int)(uint32_t)(b1 >> 32) : ls.acc3;
  if (slot == 63 || row == K.last) {
    const int rbase = row & ~63;
    if (rbase + lane <= row) lds_ptr[rbase + lane] = make_uint4(ls.acc0, ls.acc1, ls.acc2, ls.acc3);
  }
  ls.prev2 = ls.prev;
  ls.prev = e;
  // index walk (nwalign.rs:1032-1079)
  if (row < K.band && row < K.len1) {
    if (EVEN) st.col_min = imax(st.col_min - 1, 0);
    st.i_max += 1;
  } else if (st.i_max < K.len1 - 1) {
    if (K.band_mod2 == 0) {
      if (EVEN) st.j_min += 1;
      else st.i_max += 1;
    } else if (EVEN) {
      st.col_min = imax(st.col_min - 1, 0);
      st.i_max += 1;
    } else {
      st.col_min += 1;
      st.j_min += 1;
    }
  } else {
    if (!EVEN) st.col_min += 1;
    st.j_min += 1;
  }
  if (row < K.top_limit) {
    if (!EVEN) st.col_max += 1;
  } else if ((row + 1) / 2 + st.col_max - K.start_col < K.len2) {
    if (K.full_band % 2 == 0) {
      if (EVEN) st.col_max = imax(st.col_max - 1, 0);
      else st.col_max += 1;
    }
  } else if (EVEN) {
    st.col_max = imax(st.col_max - 1, 0);
  }
}

__global__ __launch_bounds__(64) void nw_vec_kernel(NwDev a) {
  extern __shared__ uint4 lds_ptr[];  // [nrow] pointer records, then group buffers
  const int lane = (int)__lane_id();
  const BinSpace bs = a.bins;
  for (;;) {
    uint32_t w = 0;
    if (lane == 0) w = atomicAdd(&a.ctr->wave_counter, 1u);
    w = (uint32_t)uni((int)w);
    const uint32_t total = (uint32_t)uni((int)a.ctr->total_waves);
    if (w >= total) break;
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
    const int G = nw_groups(geo);
    const uint32_t wave_in_bin = w - a.wave_start[bin];
    const uint32_t first = a.bin_start[bin] + wave_in_bin * (uint32_t)G;
    const uint32_t bin_end = a.bin_start[bin + 1];
    const int count = (int)min((uint32_t)G, bin_end - first);

    const int len1 = geo.len1, len2 = geo.len2, ncol = geo.ncol, nrow = geo.nrow;
    const int start_col = geo.start_col, band = geo.band;
    const bool swap = geo.swap != 0;
    const int g_raw = lane / ncol;
    const int c = lane - g_raw * ncol;
    const int g = g_raw < count ? g_raw : 0;  // idle lanes shadow group 0
    const DpTask task = a.tasks[first + (uint32_t)g];

    // ---- stage the pair's sequences (internal s1 = shorter) in LDS ----
    uint8_t* lbytes = reinterpret_cast<uint8_t*>(lds_ptr);
    const int pad = nw_pad(ncol);
    const size_t gbytes = nw_group_bytes(geo);
    uint8_t* gbase = lbytes + (size_t)nrow * 16 + (size_t)g * gbytes;
    uint8_t* s1b = gbase + pad;
    uint8_t* s2b = gbase + pad + len1 + 2 * pad;
    uint32_t* subl = reinterpret_cast<uint32_t*>(
        gbase + (((size_t)len1 + len2 + 4 * pad + 15) & ~(size_t)15));
    const uint8_t* src1 = a.pool + (swap ? task.s2_off : task.s1_off);
    const uint8_t* src2 = a.pool + (swap ? task.s1_off : task.s2_off);
    if (g_raw < count) {
      for (int t = c; t < len1; t += ncol) s1b[t] = src1[t];
      for (int t = c; t < len2; t += ncol) s2b[t] = src2[t];
    }
    __syncthreads();

    // ---- scoring constants (nwalign.rs:817-818) ----
    const int match = a.match, mismatch = a.mismatch, gap = a.gap, end_gap = a.end_gap;
    int min_score = mismatch < gap ? mismatch : gap;
    min_score = match < min_score ? match : min_score;
    min_score = min_score < 0 ? min_score : 0;
    const int fill = (int)(int16_t)(uint16_t)(uint32_t)(-32768 - min_score);

    const int lfl = geo.left_fill_limit, tfl = geo.top_fill_limit, top_limit = geo.top_limit;
    const int band_min_len1 = imin(band, len1);
    const int band_mod2 = band % 2;
    const int full_band = band + len2 - len1;
    const bool do_recalc = end_gap > gap;
    const int last = len1 + len2;

    // ---- rows 0 and 1 (nwalign.rs:842-861) ----
    int prev2 = (c == start_col) ? 0 : fill;
    int prev, p1;
    {
      const bool isL = 1 < lfl && c == imax(start_col - 1, 0);
      const bool isT = 1 < tfl && c == start_col;
      prev = (isL || isT) ? end_gap : fill;
      p1 = isT ? 2 : (isL ? 3 : 0);
    }
    int acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    {
      const uint64_t b0 = __builtin_amdgcn_ballot_w64((p1 & 1) != 0);
      const uint64_t b1 = __builtin_amdgcn_ballot_w64((p1 & 2) != 0);
      if (lane == 1) {
        acc0 = (int)(uint32_t)b0;
        acc1 = (int)(uint32_t)b1;
        acc2 = (int)(uint32_t)(b0 >> 32);
        acc3 = (int)(uint32_t)(b1 >> 32);
      }
    }
    if (last < 2) {  // no DP rows: flush rows 0..last
      if (lane <= last) lds_ptr[lane] = make_uint4(acc0, acc1, acc2, acc3);
    }

    // ---- main anti-diagonal loop (nwalign.rs:875-1088) ----
    int col_min = start_col, col_max = start_col, i_max = 0, j_min = 0;
    bool recalc_left = false, recalc_right = false;
    int row_free = end_gap;
    uint64_t cells = 0;

    RowConst K;
    K.len1 = len1;
    K.len2 = len2;
    K.start_col = start_col;
    K.lfl = lfl;
    K.tfl = tfl;
    K.top_limit = top_limit;
    K.band_min_len1 = band_min_len1;
    K.band_mod2 = band_mod2;
    K.full_band = full_band;
    K.last = last;
    K.band = band;
    K.match = match;
    K.mismatch = mismatch;
    K.gap = gap;
    K.end_gap = end_gap;
    K.fill = fill;
    K.swap = swap;
    K.do_recalc = do_recalc;
    RowState st;
    st.col_min = col_min;
    st.col_max = col_max;
    st.i_max = i_max;
    st.j_min = j_min;
    st.row_free = row_free;
    st.recalc_left = recalc_left;
    st.recalc_right = recalc_right;
    st.cells = cells;
    LaneState ls;
    ls.prev2 = prev2;
    ls.prev = prev;
    ls.acc0 = acc0;
    ls.acc1 = acc1;
    ls.acc2 = acc2;
    ls.acc3 = acc3;
    for (int row = 2; row <= last; row += 2) {
      nw_row<true>(row, st, ls, K, s1b, s2b, c, lane, lds_ptr);
      if (row + 1 <= last) nw_row<false>(row + 1, st, ls, K, s1b, s2b, c, lane, lds_ptr);
    }
    cells = st.cells;
    __syncthreads();

    if (lane == 0) atomicAdd(&a.ctr->dp_cells, (unsigned long long)(cells * (uint64_t)count));

    // ---- traceback, one lane per pair (nwalign.rs:1101-1139) ----
    if (c == 0 && g_raw < count) {
      int i = len1, j = len2, nstep = 0, nsub = 0;
      bool bad = false;
      uint8_t* out0 = nullptr;
      uint8_t* out1 = nullptr;
      if (a.emit) {
        const uint64_t o = a.al_off[task.out_idx];
        out0 = a.al0 + o;
        out1 = a.al1 + o;
      }
      while (i > 0 || j > 0) {
        const int cs = 2 * start_col + j - i;
        const int col = cs >> 1;
        if (cs < 0 || col >= ncol) {
          bad = true;
          break;
        }
        const int pos = g * ncol + col;
        const uint4 m = lds_ptr[i + j];
        const uint32_t w0 = pos < 32 ? m.x : m.z;
        const uint32_t w1 = pos < 32 ? m.y : m.w;
        const int sh = pos & 31;
        const int pp = (int)((w0 >> sh) & 1u) | (int)(((w1 >> sh) & 1u) << 1);
        int x0, x1;  // internal s1 side, s2 side
        if (pp == 1) {
          x0 = s1b[i - 1];
          x1 = s2b[j - 1];
          if (!a.emit && x0 != x1 && x0 >= 1 && x0 <= 4 && x1 >= 1 && x1 <= 4) {
            const int qp = swap ? i - 1 : j - 1;
            const int cn = swap ? x1 : x0, qn = swap ? x0 : x1;
            subl[nsub++] = ((uint32_t)qp << 8) | (uint32_t)(4 * (cn - 1) + (qn - 1));
          }
          --i;
          --j;
        } else if (pp == 2) {
          x0 = '-';
          x1 = s2b[j - 1];
          --j;
        } else if (pp == 3) {
          x0 = s1b[i - 1];
          x1 = '-';
          --i;
        } else {
          bad = true;
          break;
        }
        if (a.emit) {
          out0[nstep] = (uint8_t)(swap ? x1 : x0);
          out1[nstep] = (uint8_t)(swap ? x0 : x1);
        }
        ++nstep;
      }
      if (bad) atomicOr(&a.ctr->err_flags, kErrTraceback);
      const uint32_t ti = first + (uint32_t)g;
      if (a.emit) {
        for (int x = 0, y = nstep - 1; x < y; ++x, --y) {
          uint8_t t0 = out0[x], t1 = out1[x];
          out0[x] = out0[y];
          out1[x] = out1[y];
          out0[y] = t0;
          out1[y] = t1;
        }
        a.al_len[task.out_idx] = bad ? 0u : (uint32_t)nstep;
      } else {
        const uint32_t off = atomicAdd(a.sub_top, (uint32_t)nsub);
        if (off + (uint32_t)nsub > a.sub_capacity) {
          atomicOr(&a.ctr->err_flags, kErrSubOverflow);
          a.sub_off[ti] = 0;
          a.sub_cnt[ti] = 0;
        } else {
          for (int s = 0; s < nsub; ++s) a.sub_pool[off + s] = subl[s];
          a.sub_off[ti] = off;
          a.sub_cnt[ti] = (uint32_t)nsub | (bad ? 0x80000000u : 0u);
        }
      }
    }
    __syncthreads();
  }
}

hipError_t launch_nw_vec(const NwArgs& s, hipStream_t st) {
  NwDev d;
  d.bins = s.bins;
  d.bin_start = s.bin_start;
  d.wave_start = s.wave_start;
  d.tasks = s.tasks;
  d.pool = s.pool;
  d.match = s.match;
  d.mismatch = s.mismatch;
  d.gap = s.gap_p;
  d.end_gap = s.end_gap_p;
  d.ctr = s.ctr;
  d.emit = s.emit_alignment;
  d.sub_pool = s.sub_pool;
  d.sub_capacity = s.sub_capacity;
  d.sub_top = s.sub_top;
  d.sub_off = s.sub_off;
  d.sub_cnt = s.sub_cnt;
  d.al0 = s.al0;
  d.al1 = s.al1;
  d.al_off = s.al_off;
  d.al_len = s.al_len;
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
size_t nw_vec_lds_for(const BinSpace& bs) {
  size_t lds = 0;
  for (int b = 0; b < bs.nbins; ++b) {
    int l1, l2;
    bin_lengths(bs, (uint32_t)b, &l1, &l2);
    VecGeom g = vec_geom(l1, l2, bs.band);
    int G = nw_groups(g);
    if (G > 0) {
      size_t need = nw_wave_bytes(g, G);
      if (need > lds) lds = need;
    }
  }
  lds = (lds + 255) & ~(size_t)255;
  return lds < 256 ? 256 : lds;
}

}  // namespace d2g
