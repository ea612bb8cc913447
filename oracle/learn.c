/*
 * learn.c — restatement of build_trans_mat (dada2-rs
 * src/learn_errors.rs:618-690).  TEST INFRASTRUCTURE ONLY: a second checker
 * of the device's d2g_build_trans_mat and the timed CPU baseline of the
 * transition workload.
 *
 * As in the reference's rayon fold, raws are handed to worker threads in
 * chunks; each thread owns its alignment buffer and a private matrix, and
 * the private matrices are summed at the end.  The reference's u32 cells
 * saturate; a saturating sum of non-negative terms is min(total, 2^32 - 1)
 * in any order, so the private sums are kept exact (u64) and saturated once.
 * Each pair: raw_align_with_buf(center, raw) over Raws without k-mer vectors,
 * then the alignment walked column by column: both bases 1..4 count at
 * quality min(q, nq - 1) of the query position, which advances on every
 * query base (learn_errors.rs:655-676).
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  const orc_raw* raws;
  uint32_t nraw;
  const int64_t* center_of;
  const uint32_t* weight;
  const orc_align_params* p;
  uint32_t nq;
  atomic_uint next;
  atomic_int err_code;
  pthread_mutex_t mu;
  uint64_t* total;  /* 16 * nq */
} tm_job;

static void* tm_worker(void* arg) {
  tm_job* J = (tm_job*)arg;
  orc_buf* b = orc_buf_new();
  uint64_t* m = (uint64_t*)calloc((size_t)16 * J->nq, sizeof(uint64_t));
  for (;;) {
    const uint32_t start = atomic_fetch_add(&J->next, 32u);
    if (start >= J->nraw) break;
    const uint32_t end = start + 32u < J->nraw ? start + 32u : J->nraw;
    for (uint32_t i = start; i < end; ++i) {
      const int64_t c = J->center_of[i];
      const uint32_t w = J->weight[i];
      if (c < 0 || w == 0 || !J->raws[i].qual) continue;
      const int rc = orc_raw_align(&J->raws[c], &J->raws[i], J->p, b, NULL, NULL);
      if (rc < 0) {
        atomic_store(&J->err_code, rc);
        continue;
      }
      if (rc == 0) continue;  /* no alignment: nothing counted */
      const uint8_t* a0 = orc_buf_al0(b);
      const uint8_t* a1 = orc_buf_al1(b);
      const uint32_t L = orc_buf_alen(b);
      const uint8_t* q = J->raws[i].qual;
      uint32_t qpos = 0;
      for (uint32_t k = 0; k < L; ++k) {
        if (a0[k] >= 1 && a0[k] <= 4 && a1[k] >= 1 && a1[k] <= 4) {
          const uint32_t qq = q[qpos] < J->nq - 1 ? q[qpos] : J->nq - 1;
          m[((uint32_t)(a0[k] - 1) * 4 + (a1[k] - 1)) * J->nq + qq] += w;
        }
        if (a1[k] != '-') ++qpos;
      }
    }
  }
  pthread_mutex_lock(&J->mu);
  for (size_t k = 0; k < (size_t)16 * J->nq; ++k) J->total[k] += m[k];
  pthread_mutex_unlock(&J->mu);
  free(m);
  orc_buf_free(b);
  return NULL;
}

int orc_build_trans_mat(const orc_raw* raws, uint32_t nraw, const int64_t* center_of,
                        const uint32_t* weight, const orc_align_params* p, uint32_t nq,
                        int nthreads, uint32_t* out) {
  if (nq == 0) return ORC_ERR_ARG;
  tm_job J;
  J.raws = raws;
  J.nraw = nraw;
  J.center_of = center_of;
  J.weight = weight;
  J.p = p;
  J.nq = nq;
  atomic_init(&J.next, 0u);
  atomic_init(&J.err_code, 0);
  pthread_mutex_init(&J.mu, NULL);
  J.total = (uint64_t*)calloc((size_t)16 * nq, sizeof(uint64_t));
  if (nthreads <= 1) {
    tm_worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, tm_worker, &J);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
  }
  for (size_t k = 0; k < (size_t)16 * nq; ++k)
    out[k] = J.total[k] > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)J.total[k];
  free(J.total);
  pthread_mutex_destroy(&J.mu);
  return atomic_load(&J.err_code);
}
