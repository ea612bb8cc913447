/*
 * chimera.c — restatement of the bimera parent search in dada2-rs
 * src/chimera.rs:54-433 (TEST INFRASTRUCTURE ONLY).  Sequences are ASCII
 * bytes; alignment is the vectorized NW with end_gap_p = 0 and
 * band = max_shift (chimera.rs:234-240, 336-342).
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GAPC ((uint8_t)'-')

/* get_ham_endsfree, chimera.rs:54-86 */
uint32_t orc_get_ham_endsfree(const uint8_t* s0, const uint8_t* s1, uint32_t len) {
  uint32_t i = 0;
  int gap0 = 1, gap1 = 1;
  while ((gap0 || gap1) && i < len) {
    gap0 = gap0 && s0[i] == GAPC;
    gap1 = gap1 && s1[i] == GAPC;
    if (gap0 || gap1) ++i;
  }
  int64_t j = (int64_t)len - 1;
  gap0 = 1;
  gap1 = 1;
  while ((gap0 || gap1) && j >= (int64_t)i) {
    gap0 = gap0 && s0[j] == GAPC;
    gap1 = gap1 && s1[j] == GAPC;
    if (gap0 || gap1) --j;
  }
  if (j < (int64_t)i) return 0;
  uint32_t c = 0;
  for (int64_t p = i; p <= j; ++p) c += s0[p] != s1[p];
  return c;
}

/* get_lr, chimera.rs:99-175; out = {left, right, left_oo, right_oo} */
void orc_get_lr(const uint8_t* s0, const uint8_t* s1, uint32_t len, int allow_one_off,
                uint32_t max_shift, uint32_t out[4]) {
  uint32_t pos = 0;
  while (pos < len && s0[pos] == GAPC) ++pos;
  uint32_t left = 0;
  while (pos < len && s1[pos] == GAPC && left < max_shift) {
    ++pos;
    ++left;
  }
  while (pos < len && s0[pos] == s1[pos]) {
    ++pos;
    ++left;
  }
  uint32_t left_oo = 0;
  if (allow_one_off) {
    uint32_t loo = left;
    pos += 1;
    if (pos < len && s0[pos] != GAPC) ++loo;
    while (pos < len && s0[pos] == s1[pos]) {
      ++pos;
      ++loo;
    }
    left_oo = loo;
  }
  int64_t rp = (int64_t)len - 1;
  while (rp >= 0 && s0[rp] == GAPC) --rp;
  uint32_t right = 0;
  while (rp >= 0 && s1[rp] == GAPC && right < max_shift) {
    --rp;
    ++right;
  }
  while (rp >= 0 && s0[rp] == s1[rp]) {
    --rp;
    ++right;
  }
  uint32_t right_oo = 0;
  if (allow_one_off) {
    uint32_t roo = right;
    rp -= 1;
    if (rp >= 0 && s0[rp] != GAPC) ++roo;
    while (rp >= 0 && s0[rp] == s1[rp]) {
      --rp;
      ++roo;
    }
    right_oo = roo;
  }
  out[0] = left;
  out[1] = right;
  out[2] = left_oo;
  out[3] = right_oo;
}

/* is_bimera_with_buf, chimera.rs:218-292 */
int orc_is_bimera(const uint8_t* sq, uint32_t sqlen, const uint8_t* const* parents,
                  const uint32_t* parent_lens, uint32_t npar, const orc_bimera_params* p) {
  orc_vec_scores sc = {p->match_score, p->mismatch, p->gap_p, 0, p->max_shift};
  orc_buf* b = orc_buf_new();
  uint32_t max_left = 0, max_right = 0, oml = 0, omr = 0, omloo = 0, omroo = 0;
  int result = 0;
  for (uint32_t k = 0; k < npar; ++k) {
    int rc = orc_align_vectorized(sq, sqlen, parents[k], parent_lens[k], &sc, b);
    if (rc) {
      result = rc;
      break;
    }
    const uint8_t* al0 = orc_buf_al0(b);
    const uint8_t* al1 = orc_buf_al1(b);
    uint32_t alen = orc_buf_alen(b);
    uint32_t lr[4];
    orc_get_lr(al0, al1, alen, p->allow_one_off, (uint32_t)p->max_shift, lr);
    if (lr[0] + lr[1] >= sqlen) continue;
    if (lr[0] > max_left) max_left = lr[0];
    if (lr[1] > max_right) max_right = lr[1];
    if (p->allow_one_off &&
        orc_get_ham_endsfree(al0, al1, alen) >= (uint32_t)p->min_one_off_par_dist) {
      if (lr[0] > oml) oml = lr[0];
      if (lr[1] > omr) omr = lr[1];
      if (lr[2] > omloo) omloo = lr[2];
      if (lr[3] > omroo) omroo = lr[3];
    }
    if (max_left + max_right >= sqlen) {
      result = 1;
      break;
    }
    if (p->allow_one_off && (oml + omroo >= sqlen || omloo + omr >= sqlen)) {
      result = 1;
      break;
    }
  }
  orc_buf_free(b);
  return result;
}

typedef struct {
  const uint32_t* mat;
  uint32_t nrow, ncol;
  const uint8_t* seqs;
  const uint64_t* off;
  double min_fold;
  uint32_t min_abund;
  const orc_bimera_params* p;
  uint32_t* nflag;
  uint32_t* nsam;
  atomic_uint next;
  atomic_int err;
} tb_job;

/* the per-column body of table_bimera2, chimera.rs:351-431 */
static int tb_column(tb_job* J, orc_buf* b, uint32_t j, uint32_t* cache /* 5*ncol */,
                     uint8_t* have) {
  const orc_bimera_params* p = J->p;
  orc_vec_scores sc = {p->match_score, p->mismatch, p->gap_p, 0, p->max_shift};
  uint32_t nrow = J->nrow, ncol = J->ncol;
  const uint8_t* sj = J->seqs + J->off[j];
  uint32_t sqlen = (uint32_t)(J->off[j + 1] - J->off[j]);
  uint32_t nsam = 0, nflag = 0;
  memset(have, 0, ncol);
  for (uint32_t i = 0; i < nrow; ++i) {
    uint32_t j_abund = J->mat[i + (size_t)j * nrow];
    if (j_abund == 0) continue;
    ++nsam;
    uint32_t max_left = 0, max_right = 0, oml = 0, omr = 0, omloo = 0, omroo = 0;
    for (uint32_t k = 0; k < ncol; ++k) {
      uint32_t k_abund = J->mat[i + (size_t)k * nrow];
      if ((double)k_abund <= J->min_fold * (double)j_abund || k_abund < J->min_abund) continue;
      uint32_t* e = cache + 5 * (size_t)k;
      if (!have[k]) {
        const uint8_t* sk = J->seqs + J->off[k];
        uint32_t lk = (uint32_t)(J->off[k + 1] - J->off[k]);
        int rc = orc_align_vectorized(sj, sqlen, sk, lk, &sc, b);
        if (rc) return rc;
        const uint8_t* al0 = orc_buf_al0(b);
        const uint8_t* al1 = orc_buf_al1(b);
        uint32_t alen = orc_buf_alen(b);
        uint32_t lr[4];
        orc_get_lr(al0, al1, alen, p->allow_one_off, (uint32_t)p->max_shift, lr);
        uint32_t allowed = p->allow_one_off &&
                           orc_get_ham_endsfree(al0, al1, alen) >= (uint32_t)p->min_one_off_par_dist;
        if (lr[0] + lr[1] < sqlen) {
          e[0] = lr[0];
          e[1] = lr[1];
          e[2] = lr[2];
          e[3] = lr[3];
        } else {
          e[0] = e[1] = e[2] = e[3] = 0;
        }
        e[4] = allowed;
        have[k] = 1;
      }
      if (e[0] > max_left) max_left = e[0];
      if (e[1] > max_right) max_right = e[1];
      if (p->allow_one_off && e[4]) {
        if (e[0] > oml) oml = e[0];
        if (e[1] > omr) omr = e[1];
        if (e[2] > omloo) omloo = e[2];
        if (e[3] > omroo) omroo = e[3];
      }
    }
    if (max_left + max_right >= sqlen ||
        (p->allow_one_off && (oml + omroo >= sqlen || omloo + omr >= sqlen)))
      ++nflag;
  }
  J->nflag[j] = nflag;
  J->nsam[j] = nsam;
  return ORC_OK;
}

static void* tb_worker(void* arg) {
  tb_job* J = (tb_job*)arg;
  orc_buf* b = orc_buf_new();
  uint32_t* cache = (uint32_t*)malloc(sizeof(uint32_t) * 5 * (size_t)J->ncol + 4);
  uint8_t* have = (uint8_t*)malloc((size_t)J->ncol + 1);
  for (;;) {
    uint32_t j = atomic_fetch_add(&J->next, 1u);
    if (j >= J->ncol) break;
    int rc = tb_column(J, b, j, cache, have);
    if (rc) atomic_store(&J->err, rc);
  }
  free(cache);
  free(have);
  orc_buf_free(b);
  return NULL;
}

/* table_bimera2, chimera.rs:317-433 */
int orc_table_bimera2(const uint32_t* mat, uint32_t nrow, uint32_t ncol, const uint8_t* seqs,
                      const uint64_t* offsets, double min_fold, uint32_t min_abund,
                      const orc_bimera_params* p, uint32_t* out_nflag, uint32_t* out_nsam,
                      int nthreads) {
  tb_job J;
  J.mat = mat;
  J.nrow = nrow;
  J.ncol = ncol;
  J.seqs = seqs;
  J.off = offsets;
  J.min_fold = min_fold;
  J.min_abund = min_abund;
  J.p = p;
  J.nflag = out_nflag;
  J.nsam = out_nsam;
  atomic_init(&J.next, 0u);
  atomic_init(&J.err, 0);
  if (nthreads <= 1) {
    tb_worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, tb_worker, &J);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
  }
  return atomic_load(&J.err);
}
