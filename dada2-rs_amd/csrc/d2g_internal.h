// d2g_internal.h — records and geometry shared by the host driver and the
// gfx950 kernels of the dada2-rs compare path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define D2G_HD __host__ __device__ __forceinline__

namespace d2g {

constexpr int kWave = 64;
constexpr uint32_t kHammingNone = 0xFFFFFFFFu;
// nwalign.rs:1383: above this the reference leaves the i16 vectorized path.
constexpr uint32_t kVectorizedMaxLen = 3500;
// error flags raised by kernels (OR-ed into a device word)
constexpr uint32_t kErrTraceback = 1u;
constexpr uint32_t kErrGeometry = 2u;
constexpr uint32_t kErrQual = 4u;
constexpr uint32_t kErrSubOverflow = 8u;

D2G_HD int imin(int a, int b) { return a < b ? a : b; }
D2G_HD int imax(int a, int b) { return a > b ? a : b; }
D2G_HD int sat16(int v) { return v > 32767 ? 32767 : (v < -32768 ? -32768 : v); }

// Compressed anti-diagonal geometry of align_vectorized_with_buf
// (nwalign.rs:775-787, 833-838).  s1 is the shorter input after the swap.
struct VecGeom {
  int len1, len2, band, start_col, ncol, nrow, swap;
  int left_fill_limit, top_fill_limit, top_limit;
};

D2G_HD VecGeom vec_geom(int l1_in, int l2_in, int band_in) {
  VecGeom g;
  g.swap = l1_in > l2_in;
  g.len1 = g.swap ? l2_in : l1_in;
  g.len2 = g.swap ? l1_in : l2_in;
  g.band = band_in < 0 ? g.len2 : band_in;
  int m = imin(g.band, g.len1);
  g.start_col = 1 + (m + 1) / 2;
  g.top_limit = imin(g.band + g.len2 - g.len1, g.len2);
  g.ncol = 2 + g.start_col + g.top_limit / 2;
  g.nrow = g.len1 + g.len2 + 1;
  g.left_fill_limit = 1 + m;
  g.top_fill_limit = 1 + g.top_limit;
  return g;
}

// Pairs of one geometry share a wavefront: G = 64 / ncol groups of ncol
// lanes, lane c of a group holding compressed column c.
D2G_HD int vec_groups(int ncol) { return ncol <= kWave ? kWave / ncol : 0; }

// LDS budget of one DP wavefront: per pair the padded sequences.  The
// pointer history lives in a global slot per workgroup.
constexpr int kNwMaxGroups = 16;
constexpr size_t kNwLdsSoft = 64 * 1024;
constexpr size_t kNwLdsHard = 160 * 1024 - 1024;

D2G_HD int nw_pad(int ncol) { return ((ncol + 4) + 3) & ~3; }
// pointer history: one word per lane per 16 rows
D2G_HD size_t nw_hist_p_words(int nrow) { return (size_t)((nrow + 15) / 16) * 64; }
// per pair: both sequences with padding either side
D2G_HD size_t nw_group_bytes(const VecGeom& g) {
  size_t seq = (size_t)g.len1 + g.len2 + 4 * nw_pad(g.ncol);
  return (seq + 15) & ~(size_t)15;
}
D2G_HD size_t nw_wave_bytes(const VecGeom& g, int G) { return (size_t)G * nw_group_bytes(g); }
// One anti-diagonal of the vectorized aligner's row walk, precomputed per
// geometry (the schedule of nwalign.rs:875-1088 depends only on (len1, len2,
// band, end_gap_p > gap_p)).  Fields are in compressed columns.
struct RowRec {
  int32_t a1;       // i_max + col_min: s1 byte offset of column 0's cell
  int32_t a2;       // j_min - col_min: s2 byte offset of column 0's cell
  int32_t cmin;     // col_min
  uint32_t n;       // active width col_max - col_min + 1
  int32_t inr_lo;   // col_min + k_lo: first column whose cell is inside both sequences
  uint32_t inr_n;   // k_hi - k_lo
  int32_t edge_l;   // left ends-free edge column, -1 when none (nwalign.rs:891-894)
  int32_t edge_t;   // top ends-free edge column, -1 when none (nwalign.rs:895-898)
  int32_t row_free; // accumulated ends-free edge score (nwalign.rs:884)
  uint32_t flags;   // bit0 left recalc, bit1 right recalc, bit2 band wedge,
                    // bit3 kRowMasked (the DP may take the masked row update),
                    // bits 8-15 col_min after the wedge, bits 16-23 col_max after it
  uint32_t act_lo;  // lanes of the bin's wavefront whose column is active
  uint32_t act_hi;  // (columns col_min..col_max of each of its pairs), bits 0-31 / 32-63
};

// A row whose active columns are exactly its in-sequence columns, lie off
// each pair's outer columns, and whose recalculated columns (if any) read
// only lanes of their own pair: the DP updates it with lane masks from the
// record instead of the per-lane window tests.
constexpr uint32_t kRowMasked = 8u;

// Per bin: the rows [steady_lo, steady_hi) (even bounds, before the last row) where no
// ends-free edge or recalculation applies, every active cell lies inside
// both sequences, and the s1 / s2 offsets advance by one after each odd /
// even row, so the DP runs its short row update there; the DP cells of
// one pair (sum of the active widths); and masked = 1 when every other row
// of the bin is kRowMasked, so the DP runs its masked row update on all of
// them (vectorized bins; 0 elsewhere).
struct BinInfo {
  uint32_t steady_lo, steady_hi, cells, masked;
};

// Records of the schedule table per bin of a BinSpace whose longest length
// is lmax: rows 2..len1+len2 and two lookahead records.
D2G_HD int nw_sched_rows(int lmax) { return 2 * lmax + 1; }
// Schedule offset of a bin whose length pair no task of the call can have.
constexpr uint64_t kSchedAbsent = ~0ull;

// Geometries nw_vec cannot hold (more than 64 compressed columns) go to the
// one-pair-per-wavefront kernel (nw_wide.hip) up to these limits: its three
// score rows in LDS and its pointer matrix in one global slot.
constexpr int kWideMaxCol = 4096;
constexpr size_t kWideMaxCells = (size_t)32 << 20;
D2G_HD bool wide_fits(const VecGeom& g) {
  return g.ncol <= kWideMaxCol && (size_t)g.ncol * (size_t)g.nrow <= kWideMaxCells;
}

// DP cells of the vectorized loop for one pair: the sum of the active widths
// over the anti-diagonals, a replay of the index walk (nwalign.rs:865-1088).
D2G_HD unsigned long long vec_cells(int l1_in, int l2_in, int band_in) {
  const VecGeom g = vec_geom(l1_in, l2_in, band_in);
  const int len1 = g.len1, len2 = g.len2, b = g.band, sc = g.start_col;
  int row = 2, col_min = sc, col_max = sc, i_max = 0;
  bool even = true;
  unsigned long long total = 0;
  while (row <= len1 + len2) {
    const int n = col_max - col_min + 1;
    if (n > 0) total += (unsigned long long)n;
    if (row == imin(b, len1)) {
      col_min = imax(col_min - 1, 0);
      i_max += 1;
    }
    if (row == g.top_limit) col_max += 1;
    if (row < b && row < len1) {
      if (even) col_min = imax(col_min - 1, 0);
      i_max += 1;
    } else if (i_max < len1 - 1) {
      if (b % 2 == 0) {
        if (!even) i_max += 1;
      } else if (even) {
        col_min = imax(col_min - 1, 0);
        i_max += 1;
      } else {
        col_min += 1;
      }
    } else if (!even) {
      col_min += 1;
    }
    if (row < g.top_limit) {
      if (!even) col_max += 1;
    } else if ((row + 1) / 2 + col_max - sc < len2) {
      if ((b + len2 - len1) % 2 == 0) {
        if (even) col_max = imax(col_max - 1, 0);
        else col_max += 1;
      }
    } else if (even) {
      col_max = imax(col_max - 1, 0);
    }
    ++row;
    even = !even;
  }
  return total;
}

// Pairs per wavefront for one geometry (0 = unsupported).
D2G_HD int nw_groups(const VecGeom& g, int cap = 0) {
  int G = vec_groups(g.ncol);
  if (cap > 0 && G > cap) G = cap;
  if (G > kNwMaxGroups) G = kNwMaxGroups;
  while (G > 1 && nw_wave_bytes(g, G) > kNwLdsSoft) --G;
  if (G >= 1 && nw_wave_bytes(g, G) > kNwLdsHard) G = 0;
  return G;
}

// Row-band geometry of the scalar aligners (align_endsfree_with_buf,
// nwalign.rs:336-417; align_endsfree_homo_with_buf, nwalign.rs:435-528):
// row i computes j in [max(1, i-lband), min(len2, i+rband)] (band_adjust,
// nwalign.rs:276-284), i.e. diagonals d = j - i in [-lband, rband], with
// -9999 sentinels on d = -lband-1 and d = rband+1 (fill_band_sentinels,
// nwalign.rs:287-308).  Diagonal index D = d + lband + 1 runs over
// [0, dmax].  Lane c holds the 2k diagonals D in [2kc, 2kc + 2k): on each
// anti-diagonal the k of them whose parity matches it (k = 1, 2 or 4,
// the smallest that fits one wavefront).
constexpr int kBandMaxK = 4;

struct BandGeom {
  int len1, len2, lband, rband, dmax, lanes, nrow, k;
  bool sentinels;
};

D2G_HD BandGeom band_geom(int l1, int l2, int band, int kmin = 1) {
  BandGeom g;
  g.len1 = l1;
  g.len2 = l2;
  if (band < 0) {  // unbanded: every row spans j in [1, len2]
    g.lband = l1;
    g.rband = l2;
    g.sentinels = false;
  } else {
    g.lband = band + (l1 > l2 ? l1 - l2 : 0);
    g.rband = band + (l2 > l1 ? l2 - l1 : 0);
    g.sentinels = band < l1 || band < l2;
  }
  g.dmax = g.lband + g.rband + 2;
  g.k = 1;
  while (g.k < kBandMaxK && g.k < kmin) g.k *= 2;
  while (g.k < kBandMaxK && (g.dmax + 1 + 2 * g.k - 1) / (2 * g.k) > kWave) g.k *= 2;
  g.lanes = (g.dmax + 1 + 2 * g.k - 1) / (2 * g.k);
  g.nrow = l1 + l2 + 1;
  return g;
}

// Per pair in LDS: one u16 per position of each sequence (the base in the
// low byte; with homopolymer gaps, bit 8 flags a position inside a run of
// >= 3 identical bytes, homopolymer_mask_into, nwalign.rs:211-227), the two
// rows with 16 bytes of padding before, between (32) and after them.
D2G_HD size_t band_group_bytes(const BandGeom& g, bool) {
  return ((size_t)64 + 2 * (size_t)(g.len1 + g.len2) + 15) & ~(size_t)15;
}
// Byte offsets of the two rows within a pair's LDS.
D2G_HD size_t band_q1_offset(const BandGeom&) { return 16; }
D2G_HD size_t band_q2_offset(const BandGeom& g) { return 48 + 2 * (size_t)g.len1; }
D2G_HD size_t band_wave_bytes(const BandGeom& g, int G, bool homo) {
  return (size_t)G * band_group_bytes(g, homo);
}
D2G_HD int band_groups(const BandGeom& g, bool homo) {
  int G = g.lanes <= kWave ? kWave / g.lanes : 0;
  if (G > kNwMaxGroups) G = kNwMaxGroups;
  while (G > 1 && band_wave_bytes(g, G, homo) > kNwLdsSoft) --G;
  if (G >= 1 && band_wave_bytes(g, G, homo) > kNwLdsHard) G = 0;
  return G;
}
// Cells the reference's row-band loop computes (sum over rows i = 1..len1
// of the band width, nwalign.rs:381-384 / 481-484).
D2G_HD uint32_t band_cells(const BandGeom& g, int band) {
  uint32_t cells = 0;
  for (int i = 1; i <= g.len1; ++i) {
    const int l = band >= 0 ? imax(i - g.lband, 1) : 1;
    const int r = band >= 0 ? imin(i + g.rband, g.len2) : g.len2;
    if (r >= l) cells += (uint32_t)(r - l + 1);
  }
  return cells;
}

// Which DP kernel a bin space feeds.
constexpr int kDpVectorized = 0;
constexpr int kDpRowBand = 1;

// One pair queued for the vectorized DP.  Sequences live in a byte pool; the
// bin groups pairs of identical (len1_in, len2_in) so a wavefront's row
// walk is uniform.
struct DpTask {
  uint64_t s1_off;   // s1_in (the center in the compare path)
  uint64_t s2_off;   // s2_in (the query)
  uint32_t out_idx;  // compare: c_slot * n + original unique index; align: pair index
  uint32_t bin;
  uint32_t u_idx;    // compare: query (sorted index)
  uint32_t c_idx;    // compare: center (sorted index)
};

// One pair routed to the gapless aligner (nwalign.rs:1324-1327).
struct GaplessTask {
  uint32_t c_idx;
  uint32_t u_idx;
  uint32_t out_idx;
  uint32_t _pad;
};

// Binning of DP tasks: bin = (len1_in - lmin) * range + (len2_in - lmin).
// kind: the DP kernel of every bin (kDpVectorized or kDpRowBand); homo: the
// row-band kernel applies homopolymer gap penalties.
// tile: the DP tile shape the autotuner picks (results never depend on
// it): for the vectorized kernel a cap on pairs per wavefront (0: as many
// as fit), for the row-band kernel the least cells per lane (0 or 1: the
// fewest that fit the band into one wavefront).
struct BinSpace {
  int lmin;
  int range;
  int nbins;
  int band;
  int kind;
  int homo;
  int tile;
};

D2G_HD void bin_lengths(const BinSpace& bs, uint32_t bin, int* l1, int* l2) {
  *l1 = bs.lmin + (int)(bin / (uint32_t)bs.range);
  *l2 = bs.lmin + (int)(bin % (uint32_t)bs.range);
}

// Whether a call's tasks can fall in the bin of lengths (l1, l2): present
// (nullable: every length) flags the lengths lmin .. lmin + range - 1 the
// call's sequences have.
inline bool bin_present(const BinSpace& bs, const uint8_t* present, int l1, int l2) {
  return !present || (present[l1 - bs.lmin] && present[l2 - bs.lmin]);
}

// Pairs per wavefront of a bin for its kernel (0 = unsupported geometry).
D2G_HD int bin_groups(const BinSpace& bs, int l1, int l2) {
  if (bs.kind == kDpRowBand) return band_groups(band_geom(l1, l2, bs.band, bs.tile), bs.homo != 0);
  return nw_groups(vec_geom(l1, l2, bs.band), bs.tile);
}

// Device-side counters of one compare call.
struct Counters {
  uint32_t n_gapless;
  uint32_t n_dp;
  uint32_t n_skipped;
  uint32_t n_shrouded;
  uint32_t n_exact_fallback;  // kmer_dist8 overflow pairs (kmers.rs:259-266)
  uint32_t err_flags;
  uint32_t wave_counter;      // unused (kept for layout)
  uint32_t total_waves;
  unsigned long long dp_cells;
};

}  // namespace d2g
