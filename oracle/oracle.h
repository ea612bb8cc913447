/*
 * oracle.h — CPU restatement of the dada2-rs compare hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in dada2-rs_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and never as the thing measured or
 * shipped.  The product path never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the dada2-rs v0.2.0 source tree).  Integer/byte results must be bit-exact
 * with the reference; f64 results follow the same operation order.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   k-mers        self-distance 0, overflow -1, brute force, sparse==dense
 *   aligners      known-answer scores 60/51/47/255/10/275/272/347,
 *                 exhaustive enumeration on tiny inputs, rolling==full-matrix
 *                 on the reference fuzz grid, vectorized==endsfree score sweep,
 *                 homo-vs-vectorized divergence on the reference's fixed pair
 *   al2subs       brute-force mismatch count
 *   lambda        identical sequences == product of diagonal entries
 *   calc_pA       Poisson identities + the reference's own small-lambda series
 *                 (statrs bit parity is NOT pinned: statrs is not vendored)
 *   bimera        the reference's chimera constructs (max_left/max_right 20)
 */
#ifndef DADA2_ORACLE_H
#define DADA2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* nwalign.rs:27 / nwalign.rs:30 */
#define ORC_GAP_GLYPH 9999u
#define ORC_BAND_SENTINEL (-9999)
/* nwalign.rs:1383 */
#define ORC_VECTORIZED_MAX_LEN 3500u

/* status codes */
#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_TRACEBACK (-2)  /* reference panics: invalid pointer */
#define ORC_ERR_NONACGT (-3)    /* reference panics in compute_lambda */
#define ORC_ERR_ALLOC (-4)

/* AlignParams, nwalign.rs:63-90 (backend is always Nw here). */
typedef struct {
  int32_t match_score;
  int32_t mismatch;
  int32_t gap_p;
  int32_t homo_gap_p;
  int32_t use_kmers;
  int32_t kmer_size;
  double kdist_cutoff;
  int32_t band;
  int32_t vectorized;
  int32_t gapless;
  int32_t _pad;
} orc_align_params;

/* VectorizedAlignScores, nwalign.rs:96-103 */
typedef struct {
  int16_t match_score;
  int16_t mismatch;
  int16_t gap_p;
  int16_t end_gap_p;
  int32_t band;
} orc_vec_scores;

/* Raw, containers.rs:78-108 (only the fields the compare reads). */
typedef struct {
  const uint8_t* seq; /* 1..4 (dada) or ASCII (bimera) */
  const uint8_t* qual; /* nullable, rounded Phred */
  const uint8_t* kmer8; /* nullable, 4^k bytes */
  const uint16_t* kord; /* nullable, len-k+1 entries */
  uint32_t len;
  uint32_t reads;
  uint8_t lock;
  uint8_t prior;
  uint8_t _pad[6];
} orc_raw;

/* Reusable scratch, the counterpart of AlignBuffers (nwalign.rs:118-164). */
typedef struct orc_buf orc_buf;
orc_buf* orc_buf_new(void);
void orc_buf_free(orc_buf* b);
/* most recent alignment (valid until the next call on this buffer) */
const uint8_t* orc_buf_al0(const orc_buf* b);
const uint8_t* orc_buf_al1(const orc_buf* b);
uint32_t orc_buf_alen(const orc_buf* b);
/* method of the last orc_raw_align: 0 gapless, 1 vectorized, 2 homo, 3 ends-free */
int orc_buf_last_method(const orc_buf* b);
/* DP cells of the scalar row-band aligners (nwalign.rs:365-376 loop bounds) */
uint64_t orc_scalar_cells(uint32_t len1, uint32_t len2, int32_t band);

/* Sub, containers.rs:30-51. Arrays point into orc_buf-owned storage. */
typedef struct {
  uint32_t len0;
  uint32_t nsubs;
  const uint16_t* map;
  const uint16_t* pos;
  const uint8_t* nt0;
  const uint8_t* nt1;
  const uint8_t* q0;
  const uint8_t* q1;
} orc_sub;

/* ---------------- kmers.rs ---------------- */
void orc_assign_kmer(const uint8_t* seq, uint32_t len, int k, uint16_t* out);
void orc_assign_kmer8(const uint8_t* seq, uint32_t len, int k, uint8_t* out);
void orc_assign_kmer_order(const uint8_t* seq, uint32_t len, int k, uint16_t* out);
double orc_kmer_dist(const uint16_t* kv1, uint32_t len1, const uint16_t* kv2, uint32_t len2, int k);
double orc_kmer_dist8(const uint8_t* kv1, uint32_t len1, const uint8_t* kv2, uint32_t len2, int k);
double orc_kmer_dist8_sparse(const uint16_t* i1, const uint8_t* c1, uint32_t n1, uint32_t len1,
                             const uint16_t* i2, const uint8_t* c2, uint32_t n2, uint32_t len2,
                             int k);
double orc_kord_dist(const uint16_t* kord1, uint32_t len1, const uint16_t* kord2, uint32_t len2,
                     int k);

/* ---------------- nwalign.rs aligners ---------------- */
int orc_align_endsfree(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                       int32_t match, int32_t mismatch, int32_t gap_p, int32_t band, orc_buf* b);
int orc_align_endsfree_homo(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                            const orc_align_params* p, orc_buf* b);
int orc_align_standard(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                       int32_t match, int32_t mismatch, int32_t gap_p, int32_t band, orc_buf* b);
int orc_align_gapless(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                      orc_buf* b);
/* production rolling-buffer version, nwalign.rs:761-1140 */
int orc_align_vectorized(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                         const orc_vec_scores* sc, orc_buf* b);
/* retained full-matrix twin, nwalign.rs:2375-2719.  *stale_reads (nullable)
 * receives the number of DP reads of cells never written in the same call. */
int orc_align_vectorized_full(const uint8_t* s1, uint32_t len1, const uint8_t* s2, uint32_t len2,
                              const orc_vec_scores* sc, orc_buf* b, uint64_t* stale_reads);
/* DP cells the vectorized loop computes: sum over anti-diagonals of n. */
uint64_t orc_vectorized_cells(uint32_t len1_in, uint32_t len2_in, int32_t band);
/* geometry probe: ncol, and min/max active column over all rows */
void orc_vectorized_geometry(uint32_t len1_in, uint32_t len2_in, int32_t band, int64_t* ncol,
                             int64_t* min_col, int64_t* max_col);
/* Optimal score of an alignment under ends-free rules (test helper in
 * nwalign.rs:1814-1855, restated). */
int32_t orc_score_alignment(const uint8_t* al0, const uint8_t* al1, uint32_t alen, int32_t match,
                            int32_t mismatch, int32_t gap_p);

/* ---------------- substitutions / dispatch ---------------- */
int orc_al2subs(const uint8_t* al0, const uint8_t* al1, uint32_t alen, orc_buf* b, orc_sub* out);
/* returns 1 aligned, 0 shrouded, <0 error; kdist_out and kodist_out nullable */
int orc_raw_align(const orc_raw* r1, const orc_raw* r2, const orc_align_params* p, orc_buf* b,
                  double* kdist_out, double* kodist_out);
/* returns 1 Some(sub), 0 None (shrouded), <0 error */
int orc_sub_new(const orc_raw* r0, const orc_raw* r1, const orc_align_params* p, orc_buf* b,
                orc_sub* out);

/* ---------------- pval.rs ---------------- */
/* returns ORC_OK, lambda in *out. sub == NULL means None (lambda 0). */
int orc_compute_lambda(const orc_raw* raw, const orc_sub* sub, const double* err, uint32_t ncol,
                       int use_quals, double* out);
double orc_calc_pA(uint32_t reads, double e_reads, int prior);
double orc_poisson_upper_tail_direct(uint32_t reads, double lambda);
double orc_get_pA(uint32_t reads, int prior, double lambda, uint32_t hamming, uint32_t bi_reads,
                  int detect_singletons);
/* regularised upper incomplete gamma Q(a, x) used for the Poisson tail */
double orc_gamma_q(double a, double x);
/* statrs 0.18 ln_gamma and gamma_lr (the regularised lower incomplete gamma
 * P(a, x) behind Poisson::sf), restated */
double orc_ln_gamma(double x);
double orc_gamma_lr(double a, double x);

/* ---------------- cluster.rs compare ---------------- */
/* The per-raw map of b_compare_parallel (cluster.rs:279-330).  Outputs are
 * caller-allocated arrays of length nraw.  nthreads <= 1 runs serially
 * (b_compare order); otherwise pthreads over chunks of `grain` raws. */
int orc_compare_center(const orc_raw* raws, uint32_t nraw, uint32_t center, const double* err,
                       uint32_t ncol, const orc_align_params* p, int use_quals, int greedy,
                       double* out_lambda, uint32_t* out_hamming, uint8_t* out_skipped,
                       int nthreads, int grain, uint64_t* out_dp_cells, uint64_t* out_aligned);

/* ---------------- chimera.rs ---------------- */
typedef struct {
  int32_t allow_one_off;
  int32_t min_one_off_par_dist;
  int16_t match_score;
  int16_t mismatch;
  int16_t gap_p;
  int16_t _pad;
  int32_t max_shift;
} orc_bimera_params;

void orc_get_lr(const uint8_t* s0, const uint8_t* s1, uint32_t len, int allow_one_off,
                uint32_t max_shift, uint32_t out[4]);
uint32_t orc_get_ham_endsfree(const uint8_t* s0, const uint8_t* s1, uint32_t len);
int orc_is_bimera(const uint8_t* sq, uint32_t sqlen, const uint8_t* const* parents,
                  const uint32_t* parent_lens, uint32_t npar, const orc_bimera_params* p);
/* mat column-major nrow x ncol; seqs concatenated with offsets[ncol+1];
 * out_nflag/out_nsam length ncol. */
int orc_table_bimera2(const uint32_t* mat, uint32_t nrow, uint32_t ncol, const uint8_t* seqs,
                      const uint64_t* offsets, double min_fold, uint32_t min_abund,
                      const orc_bimera_params* p, uint32_t* out_nflag, uint32_t* out_nsam,
                      int nthreads);

/* ---------------- learn_errors.rs ---------------- */
/* build_trans_mat (learn_errors.rs:618-690): center_of[i] < 0 for none,
 * weight[i] the abundance (0: skipped); raws without k-mer vectors and with
 * qualities; out 16 * nq, saturating. */
int orc_build_trans_mat(const orc_raw* raws, uint32_t nraw, const int64_t* center_of,
                        const uint32_t* weight, const orc_align_params* p, uint32_t nq,
                        int nthreads, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
