// nw_wide.hip — align_vectorized_with_buf (nwalign.rs:761-1140) for the
// pairs whose compressed band is too wide for nw_vec's wavefront (more than
// 64 columns: |len1 - len2| above ~90 at band 16, wide or unbanded bands).
//
// One wavefront per pair.  The row walk (sentinels, ends-free edge fills,
// band wedges, the end-gap recalculation and the col_min / col_max / i_max /
// j_min index walk) is wavefront-uniform scalar state, restated from the
// reference step by step; the cells of an anti-diagonal are strided over the
// lanes: each lane forms its cells' diagonal candidate from the row two up
// and the bases, then dploop's update with the swap-dependent tie precedence
// (nwalign.rs:700-737), all in saturating i16 arithmetic on i32.  The three
// rolling score rows live in LDS; the pointer matrix (ncol x nrow bytes,
// zeroed per pair as the reference resets it) in a global slot per
// wavefront.  Lane 0 walks the traceback (nwalign.rs:1101-1132) and writes
// the same outputs as nw_vec: al2subs records (query position << 16 |
// center position, in path order) and their count, or the alignment strings
// un-swapped in align mode.  These pairs are rare; the kernel favours a
// literal restatement over speed.
#include "d2g_device.h"
#include "d2g_launch.h"

namespace d2g {

namespace {

struct WideDev {
  BinSpace bins;
  const DpTask* tasks;
  const uint8_t* pool;
  int match, mismatch, gap, end_gap;
  Counters* ctr;
  int emit;
  uint32_t* sub_pool;
  size_t sub_capacity;
  uint32_t sub_stride;
  uint32_t* sub_cnt;
  uint8_t* al0;
  uint8_t* al1;
  const uint64_t* al_off;
  uint32_t* al_len;
  uint8_t* pbuf;     // one slot of pslot bytes per wavefront
  size_t pslot;
  int max_ncol;      // LDS rows are sized for it
};

__device__ __forceinline__ int sat16(int x) { return x < -32768 ? -32768 : (x > 32767 ? 32767 : x); }

__device__ __forceinline__ void wsync() { __syncthreads(); }

__global__ __launch_bounds__(64) void nw_wide_kernel(WideDev a) {
  extern __shared__ int rows_lds[];  // 3 rolling rows of max_ncol scores
  const int lane = (int)__lane_id();
  const uint32_t n_dp = a.ctr->n_dp;
  uint8_t* p = a.pbuf + (size_t)blockIdx.x * a.pslot;
  for (uint32_t ti = blockIdx.x; ti < n_dp; ti += gridDim.x) {
    const DpTask task = a.tasks[ti];
    int l1_in, l2_in;
    bin_lengths(a.bins, task.bin, &l1_in, &l2_in);
    const VecGeom g = vec_geom(l1_in, l2_in, a.bins.band);
    if (nw_groups(g) != 0 || !wide_fits(g)) continue;  // nw_vec's, or beyond this kernel
    const bool swap = g.swap != 0;
    const uint8_t* s1 = a.pool + (swap ? task.s2_off : task.s1_off);
    const uint8_t* s2 = a.pool + (swap ? task.s1_off : task.s2_off);
    const int len1 = g.len1, len2 = g.len2, band = g.band, start_col = g.start_col;
    const int ncol = g.ncol, nrow = g.nrow;
    int* slot[3] = {rows_lds, rows_lds + a.max_ncol, rows_lds + 2 * a.max_ncol};
    // reset_buf of d16 and p8 (nwalign.rs:808-809)
    for (int x = lane; x < 3 * ncol; x += 64) slot[x / ncol][x % ncol] = 0;
    const size_t pbytes = (size_t)ncol * (size_t)nrow;
    for (size_t x = (size_t)lane * 16; x < pbytes; x += 64 * 16) {
      if (x + 16 <= pbytes) {
        *reinterpret_cast<uint4*>(p + x) = make_uint4(0u, 0u, 0u, 0u);
      } else {
        for (size_t y = x; y < pbytes; ++y) p[y] = 0;
      }
    }
    wsync();
    int min_score = a.mismatch;
    if (a.gap < min_score) min_score = a.gap;
    if (a.match < min_score) min_score = a.match;
    if (0 < min_score) min_score = 0;
    const int fill = (int)(int16_t)(uint16_t)(uint32_t)(-32768 - min_score);
    const int lfl = g.left_fill_limit, tfl = g.top_fill_limit;
    auto left_col = [&](int row) { const int v = (start_col - 1) - (row - 1) / 2; return v > 0 ? v : 0; };
    auto top_col = [&](int row) { return start_col + row / 2; };
    // rows 0 and 1 (nwalign.rs:842-861)
    if (lane == 0) {
      for (int row = 0; row < 2; ++row) {
        int* s = slot[row];
        s[0] = fill;
        s[1] = fill;
        s[ncol - 2] = fill;
        s[ncol - 1] = fill;
        if (row == 0) {
          s[start_col] = 0;
        } else {
          if (row < lfl) {
            s[left_col(row)] = a.end_gap;
            p[(size_t)row * ncol + left_col(row)] = 3;
          }
          if (row < tfl) {
            s[top_col(row)] = a.end_gap;
            p[(size_t)row * ncol + top_col(row)] = 2;
          }
        }
      }
      p[start_col] = 0;
    }
    wsync();
    int row_free = a.end_gap;
    int col_min = start_col, col_max = start_col, i_max = 0, j_min = 0;
    bool even = true, recalc_left = false, recalc_right = false, geom_bad = false;
    for (int row = 2; row <= len1 + len2; ++row) {
      const int n = col_max - col_min + 1;
      if (n < 0 || col_min < 0 || col_max >= ncol) {
        geom_bad = true;
        break;
      }
      row_free = sat16(row_free + a.end_gap);
      int* prev2 = slot[0];
      int* prev = slot[1];
      int* cur = slot[2];
      uint8_t* prow = p + (size_t)row * ncol;
      if (lane == 0) {  // sentinels, then the edge fills (nwalign.rs:884-899)
        cur[0] = fill;
        cur[1] = fill;
        cur[ncol - 2] = fill;
        cur[ncol - 1] = fill;
        if (row < lfl) {
          cur[left_col(row)] = row_free;
          prow[left_col(row)] = 3;
        }
        if (row < tfl) {
          cur[top_col(row)] = row_free;
          prow[top_col(row)] = 2;
        }
      }
      wsync();
      // diag candidates over the in-range window [k_lo, k_hi) (nwalign.rs:910-954),
      // then dploop (nwalign.rs:665-739)
      int k_lo = i_max - len1 + 1;
      k_lo = k_lo < 0 ? 0 : (k_lo > n ? n : k_lo);
      int k_hi = imin(i_max + 1, len2 - j_min);
      k_hi = k_hi < 0 ? 0 : (k_hi > n ? n : k_hi);
      if (k_hi < k_lo) k_hi = k_lo;
      if (k_lo < k_hi && j_min + k_lo < 0) {  // the reference's slice would panic
        geom_bad = true;
        break;
      }
      const int left_off = col_min - (even ? 1 : 0);
      const int up_off = col_min + 1 - (even ? 1 : 0);
      if (left_off < 0 || up_off + n > ncol) {
        geom_bad = true;
        break;
      }
      for (int k = lane; k < n; k += 64) {
        int dg;
        if (k >= k_lo && k < k_hi)
          dg = sat16(prev2[col_min + k] + (s1[i_max - k] == s2[j_min + k] ? a.match : a.mismatch));
        else
          dg = sat16(prev2[col_min + k] + fill);
        const int l = sat16(prev[left_off + k] + a.gap);
        const int u = sat16(prev[up_off + k] + a.gap);
        int e, pe;
        if (swap) {
          e = l >= u ? l : u;
          pe = l >= u ? 2 : 3;
        } else {
          e = u >= l ? u : l;
          pe = u >= l ? 3 : 2;
        }
        if (dg > e) {
          e = dg;
          pe = 1;
        }
        cur[col_min + k] = e;
        prow[col_min + k] = (uint8_t)pe;
      }
      wsync();
      // band transition (nwalign.rs:980-987)
      if (row == imin(band, len1)) {
        col_min = col_min > 0 ? col_min - 1 : 0;
        i_max += 1;
        j_min -= 1;
      }
      if (row == imin(band + len2 - len1, len2)) col_max += 1;
      // end-gap recalculation (nwalign.rs:993-1026)
      if (a.end_gap > a.gap) {
        if (lane == 0) {
          if (recalc_left) {
            const int d_free = sat16(prev[left_off] + a.end_gap);
            const int cv = cur[col_min];
            const int pc = prow[col_min];
            if (d_free > cv) {
              cur[col_min] = d_free;
              prow[col_min] = 2;
            } else if (!(d_free != cv || (!swap && pc != 1) || (swap && pc == 2))) {
              prow[col_min] = 2;
            }
          }
        }
        if (i_max == len1 - 1) recalc_left = true;
        if (recalc_right) {
          const int idx = up_off + col_max - col_min;
          if (idx < 0 || idx >= ncol) {
            geom_bad = true;
            break;
          }
          if (lane == 0) {
            const int d_free = sat16(prev[idx] + a.end_gap);
            const int cv = cur[col_max];
            const int pc = prow[col_max];
            if (d_free > cv) {
              cur[col_max] = d_free;
              prow[col_max] = 3;
            } else if (!(d_free != cv || (!swap && pc == 3) || (swap && pc != 1))) {
              prow[col_max] = 3;
            }
          }
        }
        if ((row + 1) / 2 + col_max - start_col == len2) recalc_right = true;
      }
      wsync();
      // the index walk (nwalign.rs:1032-1079)
      if (row < band && row < len1) {
        if (even) col_min = col_min > 0 ? col_min - 1 : 0;
        i_max += 1;
      } else if (i_max < len1 - 1) {
        if (band % 2 == 0) {
          if (even) j_min += 1;
          else i_max += 1;
        } else if (even) {
          col_min = col_min > 0 ? col_min - 1 : 0;
          i_max += 1;
        } else {
          col_min += 1;
          j_min += 1;
        }
      } else {
        if (!even) col_min += 1;
        j_min += 1;
      }
      const int top_limit = g.top_limit;
      if (row < top_limit) {
        if (!even) col_max += 1;
      } else if ((row + 1) / 2 + col_max - start_col < len2) {
        if ((band + len2 - len1) % 2 == 0) {
          if (even) col_max = col_max > 0 ? col_max - 1 : 0;
          else col_max += 1;
        }
      } else if (even) {
        col_max = col_max > 0 ? col_max - 1 : 0;
      }
      int* t = slot[0];  // rotate_left(1)
      slot[0] = slot[1];
      slot[1] = slot[2];
      slot[2] = t;
      even = !even;
    }
    // traceback (nwalign.rs:1101-1132), lane 0
    const size_t sub0 = (size_t)ti * a.sub_stride;
    const bool sub_ok = a.emit || sub0 + a.sub_stride <= a.sub_capacity;
    const uint64_t ao = a.emit ? a.al_off[task.out_idx] : 0;
    int nt = 0, ns = 0;
    bool bad = geom_bad;
    if (lane == 0 && !bad) {
      int i = len1, j = len2;
      while (i > 0 || j > 0) {
        const int cs = 2 * start_col + j - i;
        if (cs < 0) {
          bad = true;
          break;
        }
        const size_t at = (size_t)(i + j) * ncol + (size_t)(cs / 2);
        if (at >= pbytes) {  // past the matrix: the reference's index panics
          bad = true;
          break;
        }
        const int v = p[at];
        uint8_t x0, x1;
        if (v == 1) {
          x0 = s1[i - 1];
          x1 = s2[j - 1];
          if (!a.emit && x0 != x1 && sub_ok) {
            const int qp = swap ? i - 1 : j - 1;
            const int cp = swap ? j - 1 : i - 1;
            a.sub_pool[sub0 + ns] = ((uint32_t)qp << 16) | (uint32_t)cp;
          }
          ns += x0 != x1 ? 1 : 0;
          --i;
          --j;
        } else if (v == 2) {
          x0 = (uint8_t)'-';
          x1 = s2[j - 1];
          --j;
        } else if (v == 3) {
          x0 = s1[i - 1];
          x1 = (uint8_t)'-';
          --i;
        } else {
          bad = true;
          break;
        }
        if (a.emit) {
          a.al0[ao + nt] = swap ? x1 : x0;
          a.al1[ao + nt] = swap ? x0 : x1;
        }
        ++nt;
      }
    }
    nt = __shfl(nt, 0);
    ns = __shfl(ns, 0);
    bad = __shfl((int)bad, 0) != 0;
    if (lane == 0) {
      if (bad) atomicOr(&a.ctr->err_flags, geom_bad ? kErrGeometry : kErrTraceback);
      if (!a.emit && !sub_ok) atomicOr(&a.ctr->err_flags, kErrSubOverflow);
    }
    wsync();
    if (a.emit) {
      uint8_t* o0 = a.al0 + ao;
      uint8_t* o1 = a.al1 + ao;
      for (int x = lane; x < nt / 2; x += 64) {
        const int y = nt - 1 - x;
        const uint8_t t0 = o0[x], t1 = o1[x];
        o0[x] = o0[y];
        o1[x] = o1[y];
        o0[y] = t0;
        o1[y] = t1;
      }
      if (lane == 0) a.al_len[task.out_idx] = bad ? 0u : (uint32_t)nt;
    } else if (lane == 0) {
      a.sub_cnt[ti] = sub_ok ? ((uint32_t)ns | (bad ? 0x80000000u : 0u)) : 0u;
    }
    wsync();
  }
}

}  // namespace

WideNeed nw_wide_need(const BinSpace& bs, const uint8_t* present) {
  WideNeed w{0, 0};
  if (bs.kind != kDpVectorized) return w;
  for (int b = 0; b < bs.nbins; ++b) {
    int l1, l2;
    bin_lengths(bs, (uint32_t)b, &l1, &l2);
    if (!bin_present(bs, present, l1, l2)) continue;
    const VecGeom g = vec_geom(l1, l2, bs.band);
    if (nw_groups(g) != 0 || !wide_fits(g)) continue;
    w.max_ncol = imax(w.max_ncol, g.ncol);
    w.pslot = std::max(w.pslot, ((size_t)g.ncol * (size_t)g.nrow + 255) & ~(size_t)255);
  }
  return w;
}

hipError_t launch_nw_wide(const NwArgs& s, const WideNeed& need, uint8_t* pbuf, uint32_t grid,
                          hipStream_t st) {
  if (need.max_ncol == 0) return hipSuccess;
  if (pbuf == nullptr || grid == 0) return hipErrorInvalidValue;
  WideDev d;
  d.bins = s.bins;
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
  d.sub_stride = s.sub_stride;
  d.sub_cnt = s.sub_cnt;
  d.al0 = s.al0;
  d.al1 = s.al1;
  d.al_off = s.al_off;
  d.al_len = s.al_len;
  d.pbuf = pbuf;
  d.pslot = need.pslot;
  d.max_ncol = need.max_ncol;
  const size_t lds = (size_t)3 * need.max_ncol * sizeof(int);
  if (lds > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)nw_wide_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(nw_wide_kernel, dim3(grid), dim3(64), lds, st, d);
  return hipGetLastError();
}

}  // namespace d2g
