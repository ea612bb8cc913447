/*
 * nwalign.c — restatement of dada2-rs src/nwalign.rs (TEST INFRASTRUCTURE ONLY).
 *
 * Pointer encoding 1 = diag, 2 = left (gap in s1), 3 = up (gap in s2)
 * (nwalign.rs:11-12).  Gap glyph b'-' (45).  i16 arithmetic in the vectorized
 * aligner is saturating everywhere (nwalign.rs:690-694).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GAPC ((uint8_t)'-')

struct orc_buf {
  int32_t* d32;
  size_t d32_cap;
  uint8_t* p32;
  size_t p32_cap;
  int16_t* d16;
  size_t d16_cap;
  uint8_t* p8;
  size_t p8_cap;
  int16_t* diag;
  size_t diag_cap;
  uint8_t* homo1;
  uint8_t* homo2;
  size_t homo_cap;
  uint8_t* al0;
  uint8_t* al1;
  size_t al_cap;
  uint32_t alen;
  uint16_t* map;
  uint16_t* pos;
  uint8_t* nt0;
  uint8_t* nt1;
  uint8_t* q0;
  uint8_t* q1;
  size_t sub_cap;
  uint16_t* kv1;
  uint16_t* kv2;
  size_t kv_cap;
  uint8_t* written;
  size_t written_cap;
  int last_method; /* 0 gapless, 1 vectorized, 2 homo, 3 ends-free */
};

orc_buf* orc_buf_new(void) { return (orc_buf*)calloc(1, sizeof(orc_buf)); }

void orc_buf_free(orc_buf* b) {
  if (!b) return;
  free(b->d32);
  free(b->p32);
  free(b->d16);
  free(b->p8);
  free(b->diag);
  free(b->homo1);
  free(b->homo2);
  free(b->al0);
  free(b->al1);
  free(b->map);
  free(b->pos);
  free(b->nt0);
  free(b->nt1);
  free(b->q0);
  free(b->q1);
  free(b->kv1);
  free(b->kv2);
  free(b->written);
  free(b);
}

const uint8_t* orc_buf_al0(const orc_buf* b) { return b->al0; }
const uint8_t* orc_buf_al1(const orc_buf* b) { return b->al1; }
uint32_t orc_buf_alen(const orc_buf* b) { return b->alen; }
int orc_buf_last_method(const orc_buf* b) { return b->last_method; }

/* grow helper: ensure capacity for n elements of size sz */
static int grow(void** p, size_t* cap, size_t n, size_t sz) {
  if (*cap >= n && *p) return 0;
  size_t nc = n < 64 ? 64 : n;
  void* q = realloc(*p, nc * sz);
  if (!q) return -1;
  *p = q;
  *cap = nc;
  return 0;
}

static int ensure_al(orc_buf* b, size_t n) {
  size_t c = b->al_cap;
  if (grow((void**)&b->al0, &c, n + 1, 1)) return -1;
  c = b->al_cap;
  if (grow((void**)&b->al1, &c, n + 1, 1)) return -1;
  b->al_cap = c;
  return 0;
}

static void reverse_bytes(uint8_t* a, uint32_t n) {
  for (uint32_t i = 0, j = n ? n - 1 : 0; i < j; ++i, --j) {
    uint8_t t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
}

static inline int16_t sat16(int32_t v) {
  return v > 32767 ? (int16_t)32767 : (v < -32768 ? (int16_t)-32768 : (int16_t)v);
}
static inline int16_t sadd16(int16_t a, int16_t b) { return sat16((int32_t)a + (int32_t)b); }
static inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
static inline int64_t sat_sub1(int64_t v) { return v > 0 ? v - 1 : 0; } /* usize saturating_sub(1) */

/* homopolymer_mask_into, nwalign.rs:211-227: runs of >= 3 identical bytes */
static void homopolymer_mask(const uint8_t* seq, uint32_t n, uint8_t* mask) {
  memset(mask, 0, n);
  uint32_t rs = 0;
  while (rs < n) {
    uint32_t re = rs + 1;
    while (re < n && seq[re] == seq[rs]) ++re;
    if (re - rs >= 3)
      for (uint32_t k = rs; k < re; ++k) mask[k] = 1;
    rs = re;
  }
}

/* traceback_into, nwalign.rs:234-273 */
static int traceback_full(const uint8_t* p, size_t ncol, const uint8_t* s1, const uint8_t* s2,
                          uint32_t len1, uint32_t len2, orc_buf* b) {
  if (ensure_al(b, (size_t)len1 + len2)) return ORC_ERR_ALLOC;
  uint32_t n = 0;
  size_t i = len1, j = len2;
  while (i > 0 || j > 0) {
    switch (p[i * ncol + j]) {
      case 1:
        b->al0[n] = s1[i - 1];
        b->al1[n] = s2[j - 1];
        --i;
        --j;
        break;
      case 2:
        b->al0[n] = GAPC;
        b->al1[n] = s2[j - 1];
        --j;
        break;
      case 3:
        b->al0[n] = s1[i - 1];
        b->al1[n] = GAPC;
        --i;
        break;
      default:
        return ORC_ERR_TRACEBACK;
    }
    ++n;
  }
  reverse_bytes(b->al0, n);
  reverse_bytes(b->al1, n);
  b->alen = n;
  return ORC_OK;
}

/* band_adjust, nwalign.rs:276-284 */
static void band_adjust(uint32_t len1, uint32_t len2, int32_t band, int32_t* lb, int32_t* rb) {
  if (len2 > len1) {
    *lb = band;
    *rb = band + (int32_t)(len2 - len1);
  } else if (len1 > len2) {
    *lb = band + (int32_t)(len1 - len2);
    *rb = band;
  } else {
    *lb = band;
    *rb = band;
  }
}

/* fill_band_sentinels, nwalign.rs:287-308 */
static void fill_band_sentinels(int32_t* d, size_t ncol, uint32_t len1, uint32_t len2,
                                int32_t lband, int32_t rband, int32_t band) {
  if (band >= 0 && (band < (int32_t)len1 || band < (int32_t)len2)) {
    for (uint32_t i = 0; i <= len1; ++i) {
      int32_t li = (int32_t)i - lband - 1;
      if (li >= 0) d[i * ncol + (size_t)li] = ORC_BAND_SENTINEL;
      int32_t ri = (int32_t)i + rband + 1;
      if (ri <= (int32_t)len2) d[i * ncol + (size_t)ri] = ORC_BAND_SENTINEL;
    }
  }
}

/* shared scalar row-band DP: mode 0 ends-free (nwalign.rs:336-417),
 * mode 1 homopolymer (435-528), mode 2 standard (552-628). */
static int scalar_dp(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                     int32_t match, int32_t mismatch, int32_t gap_p, int32_t homo_gap_p,
                     int32_t band, int mode, orc_buf* b) {
  size_t ncol = (size_t)len2 + 1, nrow = (size_t)len1 + 1;
  size_t cells = ncol * nrow;
  if (grow((void**)&b->d32, &b->d32_cap, cells, sizeof(int32_t))) return ORC_ERR_ALLOC;
  if (grow((void**)&b->p32, &b->p32_cap, cells, 1)) return ORC_ERR_ALLOC;
  int32_t* d = b->d32;
  uint8_t* p = b->p32;
  memset(d, 0, cells * sizeof(int32_t));
  memset(p, 0, cells);
  uint8_t *h1 = NULL, *h2 = NULL;
  if (mode == 1) {
    size_t hc = (len1 > len2 ? len1 : len2) + 1;
    size_t c1 = b->homo_cap;
    if (grow((void**)&b->homo1, &c1, hc, 1)) return ORC_ERR_ALLOC;
    size_t c2 = b->homo_cap;
    if (grow((void**)&b->homo2, &c2, hc, 1)) return ORC_ERR_ALLOC;
    b->homo_cap = c1 < c2 ? c1 : c2;
    h1 = b->homo1;
    h2 = b->homo2;
    homopolymer_mask(s1, len1, h1);
    homopolymer_mask(s2, len2, h2);
  }
  if (mode == 2) {
    for (uint32_t i = 1; i <= len1; ++i) {
      d[i * ncol] = d[(i - 1) * ncol] + gap_p;
      p[i * ncol] = 3;
    }
    for (uint32_t j = 1; j <= len2; ++j) {
      d[j] = d[j - 1] + gap_p;
      p[j] = 2;
    }
  } else {
    for (uint32_t i = 0; i <= len1; ++i) p[i * ncol] = 3;
    for (uint32_t j = 0; j <= len2; ++j) p[j] = 2;
  }
  int32_t lband, rband;
  band_adjust(len1, len2, band, &lband, &rband);
  fill_band_sentinels(d, ncol, len1, len2, lband, rband, band);
  for (uint32_t i = 1; i <= len1; ++i) {
    int64_t l = band >= 0 ? imax64((int64_t)i - lband, 1) : 1;
    int64_t r = band >= 0 ? imin64((int64_t)i + rband, len2) : len2;
    for (int64_t j = l; j <= r; ++j) {
      int32_t left, up;
      int32_t dl = d[i * ncol + (size_t)j - 1];
      int32_t du = d[(i - 1) * ncol + (size_t)j];
      if (mode == 2) {
        left = dl + gap_p;
        up = du + gap_p;
      } else if (mode == 1) {
        left = (i == len1) ? dl : (h2[j - 1] ? dl + homo_gap_p : dl + gap_p);
        up = ((uint32_t)j == len2) ? du : (h1[i - 1] ? du + homo_gap_p : du + gap_p);
      } else {
        left = (i == len1) ? dl : dl + gap_p;
        up = ((uint32_t)j == len2) ? du : du + gap_p;
      }
      int32_t diag =
          d[(i - 1) * ncol + (size_t)j - 1] + (s1[i - 1] == s2[j - 1] ? match : mismatch);
      size_t o = i * ncol + (size_t)j;
      if (up >= diag && up >= left) {
        d[o] = up;
        p[o] = 3;
      } else if (left >= diag) {
        d[o] = left;
        p[o] = 2;
      } else {
        d[o] = diag;
        p[o] = 1;
      }
    }
  }
  return traceback_full(p, ncol, s1, s2, len1, len2, b);
}

int orc_align_endsfree(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                       int32_t match, int32_t mismatch, int32_t gap_p, int32_t band, orc_buf* b) {
  return scalar_dp(s1, len1, s2, len2, match, mismatch, gap_p, gap_p, band, 0, b);
}

int orc_align_endsfree_homo(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                            const orc_align_params* pr, orc_buf* b) {
  return scalar_dp(s1, len1, s2, len2, pr->match_score, pr->mismatch, pr->gap_p, pr->homo_gap_p,
                   pr->band, 1, b);
}

int orc_align_standard(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                       int32_t match, int32_t mismatch, int32_t gap_p, int32_t band, orc_buf* b) {
  return scalar_dp(s1, len1, s2, len2, match, mismatch, gap_p, gap_p, band, 2, b);
}

/* align_gapless_with_buf, nwalign.rs:645-655 */
int orc_align_gapless(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                      orc_buf* b) {
  uint32_t len = len1 > len2 ? len1 : len2;
  if (ensure_al(b, len)) return ORC_ERR_ALLOC;
  for (uint32_t i = 0; i < len; ++i) {
    b->al0[i] = i < len1 ? s1[i] : GAPC;
    b->al1[i] = i < len2 ? s2[i] : GAPC;
  }
  b->alen = len;
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Vectorized (compressed anti-diagonal) NW                                */
/* ---------------------------------------------------------------------- */

/* Geometry shared by both vectorized transcriptions, nwalign.rs:775-787. */
typedef struct {
  const uint8_t* s1;
  const uint8_t* s2;
  int64_t len1, len2, band, start_col, ncol, nrow;
  int swap;
} vgeom;

static void vgeom_init(vgeom* g, const uint8_t* s1_in, uint32_t l1, const uint8_t* s2_in,
                       uint32_t l2, int32_t band) {
  g->swap = l1 > l2;
  g->s1 = g->swap ? s2_in : s1_in;
  g->s2 = g->swap ? s1_in : s2_in;
  g->len1 = g->swap ? l2 : l1;
  g->len2 = g->swap ? l1 : l2;
  g->band = band < 0 ? g->len2 : band;
  int64_t m = imin64(g->band, g->len1);
  g->start_col = 1 + (m + 1) / 2; /* div_ceil */
  g->ncol = 2 + g->start_col + imin64(g->band + g->len2 - g->len1, g->len2) / 2;
  g->nrow = g->len1 + g->len2 + 1;
}

/* dploop, nwalign.rs:665-739, for one anti-diagonal. */
static void dploop(int16_t* dcur, uint8_t* pcur, const int16_t* dprev, const int16_t* diag,
                   int64_t left_off, int64_t up_off, int64_t col_min, int64_t n, int16_t gap_p,
                   int swap) {
  for (int64_t k = 0; k < n; ++k) {
    int16_t l = sadd16(dprev[left_off + k], gap_p);
    int16_t u = sadd16(dprev[up_off + k], gap_p);
    int16_t dg = diag[col_min + k];
    int16_t entry;
    uint8_t pe;
    if (swap) {
      if (l >= u) {
        entry = l;
        pe = 2;
      } else {
        entry = u;
        pe = 3;
      }
    } else {
      if (u >= l) {
        entry = u;
        pe = 3;
      } else {
        entry = l;
        pe = 2;
      }
    }
    if (dg > entry) {
      entry = dg;
      pe = 1;
    }
    dcur[col_min + k] = entry;
    pcur[col_min + k] = pe;
  }
}

/* Traceback through the compressed p matrix, nwalign.rs:1090-1139. */
static int vtraceback(const vgeom* g, const uint8_t* p, orc_buf* b) {
  if (ensure_al(b, (size_t)(g->len1 + g->len2))) return ORC_ERR_ALLOC;
  uint32_t n = 0;
  int64_t i = g->len1, j = g->len2;
  while (i > 0 || j > 0) {
    int64_t cs = 2 * g->start_col + j - i;
    if (cs < 0) return ORC_ERR_TRACEBACK;
    int64_t col = cs / 2;
    if (col >= g->ncol) return ORC_ERR_TRACEBACK;
    switch (p[(i + j) * g->ncol + col]) {
      case 1:
        b->al0[n] = g->s1[i - 1];
        b->al1[n] = g->s2[j - 1];
        --i;
        --j;
        break;
      case 2:
        b->al0[n] = GAPC;
        b->al1[n] = g->s2[j - 1];
        --j;
        break;
      case 3:
        b->al0[n] = g->s1[i - 1];
        b->al1[n] = GAPC;
        --i;
        break;
      default:
        return ORC_ERR_TRACEBACK;
    }
    ++n;
  }
  reverse_bytes(b->al0, n);
  reverse_bytes(b->al1, n);
  if (g->swap) {
    uint8_t* t = b->al0;
    b->al0 = b->al1;
    b->al1 = t;
  }
  b->alen = n;
  return ORC_OK;
}

/*
 * Production rolling-buffer aligner, nwalign.rs:761-1140.
 * The d matrix is three rolling rows [row-2, row-1, cur]; p is full size.
 */
int orc_align_vectorized(const uint8_t* s1_in, uint32_t l1, const uint8_t* s2_in, uint32_t l2,
                         const orc_vec_scores* sc, orc_buf* b) {
  vgeom g;
  vgeom_init(&g, s1_in, l1, s2_in, l2, sc->band);
  const int64_t len1 = g.len1, len2 = g.len2, band = g.band, start_col = g.start_col,
                ncol = g.ncol, nrow = g.nrow;
  const int swap = g.swap;
  const int16_t match = sc->match_score, mismatch = sc->mismatch, gap_p = sc->gap_p,
                end_gap_p = sc->end_gap_p;
  /* nwalign.rs:808-813 */
  if (grow((void**)&b->d16, &b->d16_cap, (size_t)(ncol * 3), sizeof(int16_t))) return ORC_ERR_ALLOC;
  if (grow((void**)&b->p8, &b->p8_cap, (size_t)(ncol * nrow), 1)) return ORC_ERR_ALLOC;
  if (grow((void**)&b->diag, &b->diag_cap, (size_t)ncol, sizeof(int16_t))) return ORC_ERR_ALLOC;
  memset(b->d16, 0, (size_t)(ncol * 3) * sizeof(int16_t));
  memset(b->p8, 0, (size_t)(ncol * nrow));
  memset(b->diag, 0, (size_t)ncol * sizeof(int16_t));
  uint8_t* p = b->p8;
  int16_t* diag_buf = b->diag;
  /* nwalign.rs:817-818 */
  int32_t min_score = mismatch;
  if (gap_p < min_score) min_score = gap_p;
  if (match < min_score) min_score = match;
  if (0 < min_score) min_score = 0;
  const int16_t fill_val = (int16_t)(uint16_t)(uint32_t)(-32768 - min_score);
  int16_t* slots[3] = {b->d16, b->d16 + ncol, b->d16 + 2 * ncol};
  /* nwalign.rs:833-838 */
  const int64_t left_fill_limit = 1 + imin64(band, len1);
  const int64_t top_fill_limit = 1 + imin64(band + len2 - len1, len2);
#define LEFT_COL(row) (((start_col - 1) - ((row)-1) / 2) > 0 ? ((start_col - 1) - ((row)-1) / 2) : 0)
#define TOP_COL(row) (start_col + (row) / 2)
  /* rows 0 and 1, nwalign.rs:842-861 */
  for (int64_t row = 0; row < 2; ++row) {
    int16_t* s = slots[row];
    s[0] = fill_val;
    s[1] = fill_val;
    s[ncol - 2] = fill_val;
    s[ncol - 1] = fill_val;
    if (row == 0) {
      s[start_col] = 0;
    } else {
      if (row < left_fill_limit) {
        s[LEFT_COL(row)] = end_gap_p;
        p[row * ncol + LEFT_COL(row)] = 3;
      }
      if (row < top_fill_limit) {
        s[TOP_COL(row)] = end_gap_p;
        p[row * ncol + TOP_COL(row)] = 2;
      }
    }
  }
  p[start_col] = 0;
  int16_t row_free = end_gap_p;
  int64_t row = 2, col_min = start_col, col_max = start_col, i_max = 0, j_min = 0;
  int even = 1, recalc_left = 0, recalc_right = 0;
  while (row <= len1 + len2) {
    int64_t n = col_max - col_min + 1;
    if (n < 0 || col_min < 0 || col_max >= ncol) return ORC_ERR_ARG;
    /* nwalign.rs:884-899 */
    row_free = sadd16(row_free, end_gap_p);
    {
      int16_t* cur = slots[2];
      cur[0] = fill_val;
      cur[1] = fill_val;
      cur[ncol - 2] = fill_val;
      cur[ncol - 1] = fill_val;
      if (row < left_fill_limit) {
        cur[LEFT_COL(row)] = row_free;
        p[row * ncol + LEFT_COL(row)] = 3;
      }
      if (row < top_fill_limit) {
        cur[TOP_COL(row)] = row_free;
        p[row * ncol + TOP_COL(row)] = 2;
      }
    }
    /* diag fill, nwalign.rs:910-954 */
    {
      const int16_t* prev2 = slots[0] + col_min;
      int16_t* dout = diag_buf + col_min;
      int64_t k_lo = clamp64(i_max - len1 + 1, 0, n);
      int64_t k_hi = clamp64(imin64(i_max + 1, len2 - j_min), 0, n);
      if (k_hi < k_lo) k_hi = k_lo;
      for (int64_t k = 0; k < k_lo; ++k) dout[k] = sadd16(prev2[k], fill_val);
      for (int64_t k = k_hi; k < n; ++k) dout[k] = sadd16(prev2[k], fill_val);
      if (k_lo < k_hi) {
        /* the reference slices s2[j_min + k_lo ..]; a negative start would
         * panic (usize overflow) in its debug build */
        if (j_min + k_lo < 0) return ORC_ERR_ARG;
        for (int64_t k = k_lo; k < k_hi; ++k) {
          uint8_t a = g.s1[i_max - k], bb = g.s2[j_min + k];
          dout[k] = sadd16(prev2[k], a == bb ? match : mismatch);
        }
      }
    }
    /* nwalign.rs:962-973 */
    int64_t even_off = even ? 1 : 0;
    int64_t left_off = col_min - even_off;
    int64_t up_off = col_min + 1 - even_off;
    if (left_off < 0 || up_off + n > ncol) return ORC_ERR_ARG;
    dploop(slots[2], p + row * ncol, slots[1], diag_buf, left_off, up_off, col_min, n, gap_p,
           swap);
    /* band transition, nwalign.rs:980-987 */
    if (row == imin64(band, len1)) {
      col_min = sat_sub1(col_min);
      i_max += 1;
      j_min -= 1;
    }
    if (row == imin64(band + len2 - len1, len2)) col_max += 1;
    /* ends-free recalculation, nwalign.rs:993-1026 */
    if (end_gap_p > gap_p) {
      if (recalc_left) {
        int16_t d_free = sadd16(slots[1][left_off], end_gap_p);
        int16_t cur = slots[2][col_min];
        uint8_t pc = p[row * ncol + col_min];
        if (d_free > cur) {
          slots[2][col_min] = d_free;
          p[row * ncol + col_min] = 2;
        } else if (!(d_free != cur || (!swap && pc != 1) || (swap && pc == 2))) {
          p[row * ncol + col_min] = 2;
        }
      }
      if (i_max == len1 - 1) recalc_left = 1;
      if (recalc_right) {
        int64_t idx = up_off + col_max - col_min;
        if (idx < 0 || idx >= ncol) return ORC_ERR_ARG;
        int16_t d_free = sadd16(slots[1][idx], end_gap_p);
        int16_t cur = slots[2][col_max];
        uint8_t pc = p[row * ncol + col_max];
        if (d_free > cur) {
          slots[2][col_max] = d_free;
          p[row * ncol + col_max] = 3;
        } else if (!(d_free != cur || (!swap && pc == 3) || (swap && pc != 1))) {
          p[row * ncol + col_max] = 3;
        }
      }
      int64_t j_max_1idx = (row + 1) / 2 + col_max - start_col;
      if (j_max_1idx == len2) recalc_right = 1;
    }
    /* index walk, nwalign.rs:1032-1079 */
    int64_t band_mod2 = band % 2;
    if (row < band && row < len1) {
      if (even) col_min = sat_sub1(col_min);
      i_max += 1;
    } else if (i_max < len1 - 1) {
      if (band_mod2 == 0) {
        if (even)
          j_min += 1;
        else
          i_max += 1;
      } else if (even) {
        col_min = sat_sub1(col_min);
        i_max += 1;
      } else {
        col_min += 1;
        j_min += 1;
      }
    } else {
      if (!even) col_min += 1;
      j_min += 1;
    }
    int64_t top_limit = imin64(band + len2 - len1, len2);
    if (row < top_limit) {
      if (!even) col_max += 1;
    } else if ((row + 1) / 2 + col_max - start_col < len2) {
      int64_t full_band = band + len2 - len1;
      if (full_band % 2 == 0) {
        if (even)
          col_max = sat_sub1(col_max);
        else
          col_max += 1;
      }
    } else if (even) {
      col_max = sat_sub1(col_max);
    }
    /* rotate_left(1), nwalign.rs:1084 */
    int16_t* t = slots[0];
    slots[0] = slots[1];
    slots[1] = slots[2];
    slots[2] = t;
    row += 1;
    even = !even;
  }
#undef LEFT_COL
#undef TOP_COL
  return vtraceback(&g, p, b);
}

/*
 * Full-matrix twin, restated from the reference's retained copy
 * (nwalign.rs:2375-2719): d is ncol x nrow, edges filled up front.
 */
int orc_align_vectorized_full(const uint8_t* s1_in, uint32_t l1, const uint8_t* s2_in,
                              uint32_t l2, const orc_vec_scores* sc, orc_buf* b,
                              uint64_t* stale_reads) {
  vgeom g;
  vgeom_init(&g, s1_in, l1, s2_in, l2, sc->band);
  const int64_t len1 = g.len1, len2 = g.len2, band = g.band, start_col = g.start_col,
                ncol = g.ncol, nrow = g.nrow;
  const int swap = g.swap;
  const int16_t match = sc->match_score, mismatch = sc->mismatch, gap_p = sc->gap_p,
                end_gap_p = sc->end_gap_p;
  size_t cells = (size_t)(ncol * nrow);
  if (grow((void**)&b->d16, &b->d16_cap, cells, sizeof(int16_t))) return ORC_ERR_ALLOC;
  if (grow((void**)&b->p8, &b->p8_cap, cells, 1)) return ORC_ERR_ALLOC;
  if (grow((void**)&b->diag, &b->diag_cap, (size_t)ncol, sizeof(int16_t))) return ORC_ERR_ALLOC;
  if (grow((void**)&b->written, &b->written_cap, cells, 1)) return ORC_ERR_ALLOC;
  int16_t* d = b->d16;
  uint8_t* p = b->p8;
  uint8_t* w = b->written;
  int16_t* diag_buf = b->diag;
  memset(d, 0, cells * sizeof(int16_t));
  memset(p, 0, cells);
  memset(w, 0, cells);
  memset(diag_buf, 0, (size_t)ncol * sizeof(int16_t));
  uint64_t stale = 0;
#define WR(idx, v) \
  do {             \
    d[idx] = (v);  \
    w[idx] = 1;    \
  } while (0)
#define RD(idx) (stale += !w[idx], d[idx])
  int32_t min_score = mismatch;
  if (gap_p < min_score) min_score = gap_p;
  if (match < min_score) min_score = match;
  if (0 < min_score) min_score = 0;
  const int16_t fill_val = (int16_t)(uint16_t)(uint32_t)(-32768 - min_score);
  for (int64_t row = 0; row < nrow; ++row) {
    WR(row * ncol, fill_val);
    WR(row * ncol + 1, fill_val);
    WR(row * ncol + ncol - 2, fill_val);
    WR(row * ncol + ncol - 1, fill_val);
  }
  WR(start_col, 0);
  p[start_col] = 0;
  { /* left column, nwalign.rs:2426-2440 */
    int64_t row = 1, col = start_col - 1;
    int16_t d_free = end_gap_p;
    int64_t limit = 1 + imin64(band, len1);
    while (row < limit) {
      WR(row * ncol + col, d_free);
      p[row * ncol + col] = 3;
      if (row % 2 == 0) col = sat_sub1(col);
      row += 1;
      d_free = sadd16(d_free, end_gap_p);
    }
  }
  { /* top row, nwalign.rs:2442-2457 */
    int64_t row = 1, col = start_col;
    int16_t d_free = end_gap_p;
    int64_t limit = 1 + imin64(band + len2 - len1, len2);
    while (row < limit) {
      WR(row * ncol + col, d_free);
      p[row * ncol + col] = 2;
      if (row % 2 == 1) col += 1;
      row += 1;
      d_free = sadd16(d_free, end_gap_p);
    }
  }
  int64_t row = 2, col_min = start_col, col_max = start_col, i_max = 0, j_min = 0;
  int even = 1, recalc_left = 0, recalc_right = 0;
  while (row <= len1 + len2) {
    int64_t n = col_max - col_min + 1;
    if (n < 0 || col_min < 0 || col_max >= ncol) return ORC_ERR_ARG;
    {
      int64_t base = (row - 2) * ncol + col_min;
      int16_t* dout = diag_buf + col_min;
      int64_t k_lo = clamp64(i_max - len1 + 1, 0, n);
      int64_t k_hi = clamp64(imin64(i_max + 1, len2 - j_min), 0, n);
      if (k_hi < k_lo) k_hi = k_lo;
      for (int64_t k = 0; k < k_lo; ++k) dout[k] = sadd16(RD(base + k), fill_val);
      for (int64_t k = k_hi; k < n; ++k) dout[k] = sadd16(RD(base + k), fill_val);
      if (k_lo < k_hi) {
        if (j_min + k_lo < 0) return ORC_ERR_ARG;
        for (int64_t k = k_lo; k < k_hi; ++k) {
          uint8_t a = g.s1[i_max - k], bb = g.s2[j_min + k];
          dout[k] = sadd16(RD(base + k), a == bb ? match : mismatch);
        }
      }
    }
    int64_t even_off = even ? 1 : 0;
    int64_t prev_base = (row - 1) * ncol;
    int64_t left_off = prev_base + col_min - even_off;
    int64_t up_off = prev_base + col_min + 1 - even_off;
    if (col_min - even_off < 0 || col_min + 1 - even_off + n > ncol) return ORC_ERR_ARG;
    for (int64_t k = 0; k < n; ++k) {
      int16_t l = sadd16(RD(left_off + k), gap_p);
      int16_t u = sadd16(RD(up_off + k), gap_p);
      int16_t dg = diag_buf[col_min + k];
      int16_t entry;
      uint8_t pe;
      if (swap) {
        if (l >= u) {
          entry = l;
          pe = 2;
        } else {
          entry = u;
          pe = 3;
        }
      } else {
        if (u >= l) {
          entry = u;
          pe = 3;
        } else {
          entry = l;
          pe = 2;
        }
      }
      if (dg > entry) {
        entry = dg;
        pe = 1;
      }
      WR(row * ncol + col_min + k, entry);
      p[row * ncol + col_min + k] = pe;
    }
    if (row == imin64(band, len1)) {
      col_min = sat_sub1(col_min);
      i_max += 1;
      j_min -= 1;
    }
    if (row == imin64(band + len2 - len1, len2)) col_max += 1;
    if (end_gap_p > gap_p) {
      if (recalc_left) {
        int16_t d_free = sadd16(RD(left_off), end_gap_p);
        int16_t cur = RD(row * ncol + col_min);
        uint8_t pc = p[row * ncol + col_min];
        if (d_free > cur) {
          WR(row * ncol + col_min, d_free);
          p[row * ncol + col_min] = 2;
        } else if (!(d_free != cur || (!swap && pc != 1) || (swap && pc == 2))) {
          p[row * ncol + col_min] = 2;
        }
      }
      if (i_max == len1 - 1) recalc_left = 1;
      if (recalc_right) {
        int64_t idx = up_off + col_max - col_min;
        if (idx < 0 || idx >= (int64_t)cells) return ORC_ERR_ARG;
        int16_t d_free = sadd16(RD(idx), end_gap_p);
        int16_t cur = RD(row * ncol + col_max);
        uint8_t pc = p[row * ncol + col_max];
        if (d_free > cur) {
          WR(row * ncol + col_max, d_free);
          p[row * ncol + col_max] = 3;
        } else if (!(d_free != cur || (!swap && pc == 3) || (swap && pc != 1))) {
          p[row * ncol + col_max] = 3;
        }
      }
      int64_t j_max_1idx = (row + 1) / 2 + col_max - start_col;
      if (j_max_1idx == len2) recalc_right = 1;
    }
    int64_t band_mod2 = band % 2;
    if (row < band && row < len1) {
      if (even) col_min = sat_sub1(col_min);
      i_max += 1;
    } else if (i_max < len1 - 1) {
      if (band_mod2 == 0) {
        if (even)
          j_min += 1;
        else
          i_max += 1;
      } else if (even) {
        col_min = sat_sub1(col_min);
        i_max += 1;
      } else {
        col_min += 1;
        j_min += 1;
      }
    } else {
      if (!even) col_min += 1;
      j_min += 1;
    }
    int64_t top_limit = imin64(band + len2 - len1, len2);
    if (row < top_limit) {
      if (!even) col_max += 1;
    } else if ((row + 1) / 2 + col_max - start_col < len2) {
      int64_t full_band = band + len2 - len1;
      if (full_band % 2 == 0) {
        if (even)
          col_max = sat_sub1(col_max);
        else
          col_max += 1;
      }
    } else if (even) {
      col_max = sat_sub1(col_max);
    }
    row += 1;
    even = !even;
  }
#undef WR
#undef RD
  if (stale_reads) *stale_reads = stale;
  return vtraceback(&g, p, b);
}

/* Replays only the index walk of the vectorized aligner (nwalign.rs:865-1088)
 * to count DP cells and report the active-column envelope. */
static void vwalk(uint32_t l1, uint32_t l2, int32_t band_in, uint64_t* cells, int64_t* ncol_out,
                  int64_t* mn, int64_t* mx) {
  vgeom g;
  vgeom_init(&g, NULL, l1, NULL, l2, band_in);
  const int64_t len1 = g.len1, len2 = g.len2, band = g.band, start_col = g.start_col;
  int64_t row = 2, col_min = start_col, col_max = start_col, i_max = 0;
  int even = 1;
  uint64_t total = 0;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  while (row <= len1 + len2) {
    int64_t n = col_max - col_min + 1;
    if (n > 0) {
      total += (uint64_t)n;
      if (col_min < lo) lo = col_min;
      if (col_max > hi) hi = col_max;
    }
    if (row == imin64(band, len1)) {
      col_min = sat_sub1(col_min);
      i_max += 1;
    }
    if (row == imin64(band + len2 - len1, len2)) col_max += 1;
    int64_t band_mod2 = band % 2;
    if (row < band && row < len1) {
      if (even) col_min = sat_sub1(col_min);
      i_max += 1;
    } else if (i_max < len1 - 1) {
      if (band_mod2 == 0) {
        if (!even) i_max += 1;
      } else if (even) {
        col_min = sat_sub1(col_min);
        i_max += 1;
      } else {
        col_min += 1;
      }
    } else {
      if (!even) col_min += 1;
    }
    int64_t top_limit = imin64(band + len2 - len1, len2);
    if (row < top_limit) {
      if (!even) col_max += 1;
    } else if ((row + 1) / 2 + col_max - start_col < len2) {
      int64_t full_band = band + len2 - len1;
      if (full_band % 2 == 0) {
        if (even)
          col_max = sat_sub1(col_max);
        else
          col_max += 1;
      }
    } else if (even) {
      col_max = sat_sub1(col_max);
    }
    row += 1;
    even = !even;
  }
  if (cells) *cells = total;
  if (ncol_out) *ncol_out = g.ncol;
  if (mn) *mn = lo;
  if (mx) *mx = hi;
}

uint64_t orc_vectorized_cells(uint32_t len1_in, uint32_t len2_in, int32_t band) {
  uint64_t c = 0;
  vwalk(len1_in, len2_in, band, &c, NULL, NULL, NULL);
  return c;
}

void orc_vectorized_geometry(uint32_t len1_in, uint32_t len2_in, int32_t band, int64_t* ncol,
                             int64_t* min_col, int64_t* max_col) {
  vwalk(len1_in, len2_in, band, NULL, ncol, min_col, max_col);
}

/* score_alignment test helper, nwalign.rs:1814-1855 */
int32_t orc_score_alignment(const uint8_t* al0, const uint8_t* al1, uint32_t n, int32_t match,
                            int32_t mismatch, int32_t gap_p) {
  int32_t score = 0;
  for (uint32_t k = 0; k < n; ++k) {
    int g0 = al0[k] == GAPC, g1 = al1[k] == GAPC;
    if (!g0 && !g1) {
      score += al0[k] == al1[k] ? match : mismatch;
    } else if (g0) {
      int at_start = 1, at_end = 1;
      for (uint32_t t = 0; t < k; ++t) at_start &= al0[t] == GAPC;
      for (uint32_t t = k + 1; t < n; ++t) at_end &= al0[t] == GAPC;
      if (!at_start && !at_end) score += gap_p;
    } else {
      int at_start = 1, at_end = 1;
      for (uint32_t t = 0; t < k; ++t) at_start &= al1[t] == GAPC;
      for (uint32_t t = k + 1; t < n; ++t) at_end &= al1[t] == GAPC;
      if (!at_start && !at_end) score += gap_p;
    }
  }
  return score;
}

/* ---------------------------------------------------------------------- */
/* al2subs / raw_align / sub_new                                           */
/* ---------------------------------------------------------------------- */

static inline int is_nt(uint8_t c) { return c >= 1 && c <= 5; } /* nwalign.rs:203-206 */

/* al2subs, nwalign.rs:1169-1224 */
int orc_al2subs(const uint8_t* al0, const uint8_t* al1, uint32_t alen, orc_buf* b,
                orc_sub* out) {
  uint32_t len0 = 0, nsubs = 0;
  for (uint32_t i = 0; i < alen; ++i) {
    int n0 = is_nt(al0[i]), n1 = is_nt(al1[i]);
    if (n0) ++len0;
    if (n0 && n1 && al0[i] != al1[i] && al0[i] != 5 && al1[i] != 5) ++nsubs;
  }
  size_t need = (size_t)(len0 > nsubs ? len0 : nsubs) + 1;
  size_t c;
  c = b->sub_cap;
  if (grow((void**)&b->map, &c, need, sizeof(uint16_t))) return ORC_ERR_ALLOC;
  c = b->sub_cap;
  if (grow((void**)&b->pos, &c, need, sizeof(uint16_t))) return ORC_ERR_ALLOC;
  c = b->sub_cap;
  if (grow((void**)&b->nt0, &c, need, 1)) return ORC_ERR_ALLOC;
  c = b->sub_cap;
  if (grow((void**)&b->nt1, &c, need, 1)) return ORC_ERR_ALLOC;
  c = b->sub_cap;
  if (grow((void**)&b->q0, &c, need, 1)) return ORC_ERR_ALLOC;
  c = b->sub_cap;
  if (grow((void**)&b->q1, &c, need, 1)) return ORC_ERR_ALLOC;
  b->sub_cap = c;
  for (uint32_t i = 0; i < len0; ++i) b->map[i] = (uint16_t)ORC_GAP_GLYPH;
  int64_t i0 = -1, i1 = -1;
  uint32_t s = 0;
  for (uint32_t i = 0; i < alen; ++i) {
    int n0 = is_nt(al0[i]), n1 = is_nt(al1[i]);
    if (n0) ++i0;
    if (n1) ++i1;
    if (n0) b->map[i0] = n1 ? (uint16_t)i1 : (uint16_t)ORC_GAP_GLYPH;
    if (n0 && n1 && al0[i] != al1[i] && al0[i] != 5 && al1[i] != 5) {
      b->pos[s] = (uint16_t)i0;
      b->nt0[s] = al0[i];
      b->nt1[s] = al1[i];
      ++s;
    }
  }
  out->len0 = len0;
  out->nsubs = nsubs;
  out->map = b->map;
  out->pos = b->pos;
  out->nt0 = b->nt0;
  out->nt1 = b->nt1;
  out->q0 = NULL;
  out->q1 = NULL;
  return ORC_OK;
}

/* raw_align_with_buf + raw_align_dp, nwalign.rs:1244-1414 (NW backend). */
int orc_raw_align(const orc_raw* r1, const orc_raw* r2, const orc_align_params* p, orc_buf* b,
                  double* kdist_out, double* kodist_out) {
  double kdist = 0.0, kodist = -1.0;
  if (p->use_kmers) {
    int k = p->kmer_size;
    if (r1->kmer8 && r2->kmer8) {
      double d8 = orc_kmer_dist8(r1->kmer8, r1->len, r2->kmer8, r2->len, k);
      if (d8 < 0.0) {
        size_t nk = (size_t)1 << (2 * k);
        if (grow((void**)&b->kv1, &b->kv_cap, nk, sizeof(uint16_t))) return ORC_ERR_ALLOC;
        size_t c2 = b->kv_cap;
        if (grow((void**)&b->kv2, &c2, nk, sizeof(uint16_t))) return ORC_ERR_ALLOC;
        orc_assign_kmer(r1->seq, r1->len, k, b->kv1);
        orc_assign_kmer(r2->seq, r2->len, k, b->kv2);
        kdist = orc_kmer_dist(b->kv1, r1->len, b->kv2, r2->len, k);
      } else {
        kdist = d8;
      }
    }
    if (p->gapless && r1->kord && r2->kord)
      kodist = orc_kord_dist(r1->kord, r1->len, r2->kord, r2->len, k);
  }
  if (kdist_out) *kdist_out = kdist;
  if (kodist_out) *kodist_out = kodist;
  if (p->use_kmers && kdist > p->kdist_cutoff) return 0;
  /* method selection, nwalign.rs:1324-1413 */
  double diff = kodist - kdist;
  if (diff < 0) diff = -diff;
  if (p->band == 0 || (p->gapless && diff < 2.220446049250313080847e-16)) {
    b->last_method = 0;
    int rc = orc_align_gapless(r1->seq, r1->len, r2->seq, r2->len, b);
    return rc ? rc : 1;
  }
  int use_homo = p->homo_gap_p != p->gap_p && p->homo_gap_p <= 0;
  int too_long = r1->len > ORC_VECTORIZED_MAX_LEN || r2->len > ORC_VECTORIZED_MAX_LEN;
  int rc;
  if (p->vectorized && !too_long && !use_homo) {
    orc_vec_scores sc = {(int16_t)p->match_score, (int16_t)p->mismatch, (int16_t)p->gap_p, 0,
                         p->band};
    b->last_method = 1;
    rc = orc_align_vectorized(r1->seq, r1->len, r2->seq, r2->len, &sc, b);
  } else if (use_homo) {
    b->last_method = 2;
    rc = orc_align_endsfree_homo(r1->seq, r1->len, r2->seq, r2->len, p, b);
  } else {
    b->last_method = 3;
    rc = orc_align_endsfree(r1->seq, r1->len, r2->seq, r2->len, p->match_score, p->mismatch,
                            p->gap_p, p->band, b);
  }
  return rc ? rc : 1;
}

/* sub_new_with_buf, nwalign.rs:1429-1455 */
int orc_sub_new(const orc_raw* r0, const orc_raw* r1, const orc_align_params* p, orc_buf* b,
                orc_sub* out) {
  int rc = orc_raw_align(r0, r1, p, b, NULL, NULL);
  if (rc <= 0) return rc;
  rc = orc_al2subs(b->al0, b->al1, b->alen, b, out);
  if (rc) return rc;
  if (r0->qual && r1->qual) {
    for (uint32_t s = 0; s < out->nsubs; ++s) {
      b->q0[s] = r0->qual[b->pos[s]];
      b->q1[s] = r1->qual[b->map[b->pos[s]]];
    }
    out->q0 = b->q0;
    out->q1 = b->q1;
  }
  return 1;
}

/* cells visited by the scalar row-band loops (nwalign.rs:365-376) */
uint64_t orc_scalar_cells(uint32_t len1, uint32_t len2, int32_t band) {
  int32_t lband, rband;
  band_adjust(len1, len2, band, &lband, &rband);
  uint64_t total = 0;
  for (uint32_t i = 1; i <= len1; ++i) {
    int64_t l = band >= 0 ? imax64((int64_t)i - lband, 1) : 1;
    int64_t r = band >= 0 ? imin64((int64_t)i + rband, len2) : len2;
    if (r >= l) total += (uint64_t)(r - l + 1);
  }
  return total;
}
