/*
 * cluster.c — restatement of the b_compare / b_compare_parallel map
 * (dada2-rs src/cluster.rs:106-164, 239-330).  TEST INFRASTRUCTURE ONLY.
 *
 * The multithreaded form mirrors the rayon map: raws are split into chunks of
 * `grain` (cluster.rs:19-28, default 32) handed to worker threads, each owning
 * its own scratch buffer (cluster.rs:279-290).  Results are written by raw
 * index, so the output is independent of scheduling.  This is the timed CPU
 * baseline's compute body as well as the checker.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>

#include "oracle.h"

typedef struct {
  const orc_raw* raws;
  uint32_t nraw, center;
  const double* err;
  uint32_t ncol;
  const orc_align_params* p;
  int use_quals, greedy, grain;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
  atomic_uint next;
  atomic_int err_code;
  atomic_ullong cells;
  atomic_ullong aligned;
} cmp_job;

/* one comparison, cluster.rs:294-304 */
static int compare_one(cmp_job* J, orc_buf* b, uint32_t index, uint64_t* cells,
                       uint64_t* aligned) {
  const orc_raw* raw = &J->raws[index];
  const orc_raw* c = &J->raws[J->center];
  if (J->greedy && (raw->reads > c->reads || raw->lock)) {
    J->out_lambda[index] = 0.0; /* compute_lambda(raw, None) */
    J->out_hamming[index] = UINT32_MAX;
    J->out_skipped[index] = 1;
    return ORC_OK;
  }
  orc_sub sub;
  int rc = orc_sub_new(c, raw, J->p, b, &sub);
  if (rc < 0) return rc;
  double lam;
  int rc2 = orc_compute_lambda(raw, rc == 1 ? &sub : NULL, J->err, J->ncol, J->use_quals, &lam);
  if (rc2) return rc2;
  J->out_lambda[index] = lam;
  J->out_hamming[index] = rc == 1 ? sub.nsubs : UINT32_MAX;
  J->out_skipped[index] = 0;
  if (rc == 1) {
    int m = orc_buf_last_method(b);
    if (m != 0) *aligned += 1;
    if (m == 1) *cells += orc_vectorized_cells(c->len, raw->len, J->p->band);
    if (m >= 2) *cells += orc_scalar_cells(c->len, raw->len, J->p->band);
  }
  return ORC_OK;
}

static void* worker(void* arg) {
  cmp_job* J = (cmp_job*)arg;
  orc_buf* b = orc_buf_new();
  uint64_t cells = 0, aligned = 0;
  for (;;) {
    uint32_t start = atomic_fetch_add(&J->next, (unsigned)J->grain);
    if (start >= J->nraw) break;
    uint32_t end = start + (uint32_t)J->grain;
    if (end > J->nraw) end = J->nraw;
    for (uint32_t i = start; i < end; ++i) {
      int rc = compare_one(J, b, i, &cells, &aligned);
      if (rc) atomic_store(&J->err_code, rc);
    }
  }
  atomic_fetch_add(&J->cells, cells);
  atomic_fetch_add(&J->aligned, aligned);
  orc_buf_free(b);
  return NULL;
}

int orc_compare_center(const orc_raw* raws, uint32_t nraw, uint32_t center, const double* err,
                       uint32_t ncol, const orc_align_params* p, int use_quals, int greedy,
                       double* out_lambda, uint32_t* out_hamming, uint8_t* out_skipped,
                       int nthreads, int grain, uint64_t* out_dp_cells, uint64_t* out_aligned) {
  if (center >= nraw) return ORC_ERR_ARG;
  cmp_job J;
  J.raws = raws;
  J.nraw = nraw;
  J.center = center;
  J.err = err;
  J.ncol = ncol;
  J.p = p;
  J.use_quals = use_quals;
  J.greedy = greedy;
  J.grain = grain > 0 ? grain : 32;
  J.out_lambda = out_lambda;
  J.out_hamming = out_hamming;
  J.out_skipped = out_skipped;
  atomic_init(&J.next, 0u);
  atomic_init(&J.err_code, 0);
  atomic_init(&J.cells, 0ull);
  atomic_init(&J.aligned, 0ull);
  if (nthreads <= 1) {
    worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &J);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
  }
  if (out_dp_cells) *out_dp_cells = atomic_load(&J.cells);
  if (out_aligned) *out_aligned = atomic_load(&J.aligned);
  return atomic_load(&J.err_code);
}
