/*
 * dada2_amd.h — C ABI of the MI355X (gfx950) compare hot path for dada2-rs.
 *
 * This is the drop-in boundary: plain pointers and sizes, no torch or Rust
 * types.  Each entry point names the reference function it replaces
 * (paths relative to the dada2-rs v0.2.0 source tree).
 *
 * Conventions
 *  - Every entry point returns an int status: D2G_OK (0) or a negative code.
 *    Invariant violations that panic in the reference (an invalid traceback
 *    pointer, nwalign.rs:268/1130; a non-ACGT base reaching compute_lambda,
 *    pval.rs:281-284) become error codes, never aborts.
 *  - Host arrays are caller-owned and copied.  Device buffers are owned by the
 *    context.  Outputs are caller-allocated.
 *  - A context is not thread-safe; use one per host thread (the reference's
 *    AlignBuffers are likewise per worker, nwalign.rs:112-113).
 *  - Sequences for the compare path use the reference encoding A=1 C=2 G=3
 *    T=4 (misc.rs:346-356); bimera sequences are ASCII (remove_bimera.rs:70-75).
 */
#ifndef DADA2_AMD_H
#define DADA2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D2G_OK 0
#define D2G_ERR_ARG (-1)         /* invalid argument / size */
#define D2G_ERR_HIP (-2)         /* HIP runtime error (no device, OOM, launch) */
#define D2G_ERR_NONACGT (-3)     /* non-ACGT base in the compare path */
#define D2G_ERR_TRACEBACK (-4)   /* invalid traceback pointer (reference panics) */
#define D2G_ERR_UNSUPPORTED (-5) /* geometry beyond the kernels' limits */
#define D2G_ERR_STATE (-6)       /* call order (e.g. compare before load) */

#define D2G_HAMMING_NONE 0xFFFFFFFFu /* shrouded or skipped (cluster.rs:262, 302) */

typedef struct d2g_ctx d2g_ctx;

/* AlignParams, nwalign.rs:63-90.  The WFA backend is not part of this
 * library (it is off by default and not NW-identical, nwalign.rs:1340-1348). */
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
  int32_t _reserved;
} d2g_align_params;

/* VectorizedAlignScores, nwalign.rs:96-103. */
typedef struct {
  int16_t match_score;
  int16_t mismatch;
  int16_t gap_p;
  int16_t end_gap_p;
  int32_t band;
} d2g_vec_scores;

/* Per-call counters (the GPU counterpart of CompareTiming, cluster.rs:190-231).
 * Kernel times are measured with HIP events on the context's stream. */
typedef struct {
  uint64_t comparisons; /* (center, unique) pairs processed */
  uint64_t skipped;     /* greedy skips (cluster.rs:295) */
  uint64_t shrouded;    /* k-mer screen rejections (nwalign.rs:1299) */
  uint64_t gapless;     /* gapless dispatch (nwalign.rs:1324) */
  uint64_t aligned_dp;  /* pairs that ran a DP aligner */
  uint64_t dp_cells;    /* exact DP cells computed (reference loop count) */
  double ms_screen;
  double ms_plan;
  double ms_dp;
  double ms_lambda;
  double ms_total;      /* first to last event on the stream */
} d2g_stats;

/* ---- context ---------------------------------------------------------- */
int d2g_ctx_create(int device, d2g_ctx** out);
int d2g_ctx_destroy(d2g_ctx* ctx);
/* Enable per-kernel HIP-event timing in d2g_stats (default off). */
int d2g_ctx_set_timing(d2g_ctx* ctx, int enable);
/* Replay the compare's launch sequence as a HIP graph, captured once per
 * shape (centers count, bin space, parameters, buffers) (default on; timed
 * calls always launch eagerly). */
int d2g_ctx_set_graphs(d2g_ctx* ctx, int enable);
/* Launch shape of the k-mer screen: a workgroup compares qblock = 64, 128
 * or 256 uniques with 16 centers, split over split = 1, 2 or 4 wavefront
 * groups (16/split centers each), each lane taking per_lane = 1 or 2
 * uniques (2 needs qblock >= 128).  Results do not depend on it (default
 * 64, 4, 1).  No counterpart in the reference. */
int d2g_ctx_set_screen_shape(d2g_ctx* ctx, int qblock, int split, int per_lane);
/* Autotuning of the screen's launch shape (default off): the first compare
 * of each workload class (uniques and centers by power of two, k, band and
 * aligner) runs once per candidate shape, timed, and the class keeps the
 * fastest.  Results do not depend on it.  No counterpart in the
 * reference. */
int d2g_ctx_set_autotune(d2g_ctx* ctx, int enable);
/* The DP tile shape the autotuner would pick, fixed: for the vectorized
 * kernel a cap on pairs per wavefront (0: as many as fit), for the scalar
 * row-band kernel the least cells per lane (1, 2 or 4); -1 (the default)
 * leaves it to the autotuner (with autotuning off: as many pairs / as few
 * cells as fit).  Results do not depend on it.  No counterpart in the
 * reference. */
int d2g_ctx_set_dp_tile(d2g_ctx* ctx, int tile);
/* The hipStream_t every kernel of this context is launched on. */
void* d2g_ctx_stream(d2g_ctx* ctx);
/* Device properties used for roofline figures. */
int d2g_device_info(d2g_ctx* ctx, int* cu_count, int* clock_khz, int* mem_clock_khz,
                    int* mem_bus_width);
const char* d2g_status_string(int status);

/* ---- uniques ------------------------------------------------------------
 * Replaces the per-raw preparation of dada_uniques_cached (dada.rs:243-402):
 * nt_encode'd sequences (1..4), rounded qualities (containers.rs:128-137),
 * read counts, and raw_assign_kmers (kmers.rs:213-221) which the device
 * builds itself (dense 4^k kmer8 screen + positional kord).
 * seq_concat holds n sequences back to back; offsets has n+1 entries.
 * qual_concat is nullable (no qualities: use_quals has no effect).
 * Validation mirrors dada.rs:248-317: k in 3..=8, every length > k,
 * every base in 1..4, lengths < 9999. */
int d2g_load_uniques(d2g_ctx* ctx, uint32_t n, const uint8_t* seq_concat,
                     const uint64_t* offsets, const uint8_t* qual_concat,
                     const uint32_t* reads, uint32_t kmer_size);

/* ---- the compare ---------------------------------------------------------
 * Plain path: one center against every loaded unique.  Replaces the parallel
 * map of b_compare_parallel (cluster.rs:279-330; the serial twin b_compare,
 * cluster.rs:106-164).  For each unique index:
 *   out_lambda[i]  compute_lambda(raw_i, sub) (pval.rs:253), 0.0 when
 *                  shrouded or greedy-skipped
 *   out_hamming[i] nsubs, or D2G_HAMMING_NONE when shrouded/skipped
 *   out_skipped[i] 1 when greedy && (reads_i > reads_center || lock_i)
 * err is row-major 16 x ncol f64 (pval.rs:246-250); every quality must be
 * < ncol.  lock is nullable (all unlocked).  The serial store
 * (cluster.rs:351-397) stays with the caller. */
int d2g_compare_center(d2g_ctx* ctx, uint32_t center, const double* err, uint32_t ncol,
                       const d2g_align_params* params, int use_quals, int greedy,
                       const uint8_t* lock, double* out_lambda, uint32_t* out_hamming,
                       uint8_t* out_skipped, d2g_stats* stats);

/* Batched path: n_centers centers x all uniques in one device pipeline.
 * Outputs are n_centers x n_uniques, row c for centers[c]. */
int d2g_compare_batch(d2g_ctx* ctx, const uint32_t* centers, uint32_t n_centers,
                      const double* err, uint32_t ncol, const d2g_align_params* params,
                      int use_quals, int greedy, const uint8_t* lock, double* out_lambda,
                      uint32_t* out_hamming, uint8_t* out_skipped, d2g_stats* stats);

/* Device-resident variant of d2g_compare_batch: results stay in context-owned
 * device buffers (fetch with d2g_fetch_results).  This is what a caller
 * looping compare rounds on the device uses; it performs no D2H copy. */
int d2g_compare_batch_device(d2g_ctx* ctx, const uint32_t* centers, uint32_t n_centers,
                             const double* err, uint32_t ncol, const d2g_align_params* params,
                             int use_quals, int greedy, const uint8_t* lock, d2g_stats* stats);
int d2g_fetch_results(d2g_ctx* ctx, double* out_lambda, uint32_t* out_hamming,
                      uint8_t* out_skipped);

/* ---- the store on the device ---------------------------------------------
 * Replaces the serial store of b_compare_parallel (cluster.rs:351-397): a
 * compare keeps comparison (i, index) when lambda * B.reads > e_minmax[index]
 * (strict), raises e_minmax[index] to lambda * center reads when that is
 * larger, and pushes Comparison { i, index, lambda, hamming (0 when
 * shrouded) } onto cluster i's comp vector in raw-index order.  e_minmax and
 * the kept comparisons stay resident; only the kept ones come back.
 * A record of the store: Comparison (containers.rs:30-51). */
typedef struct d2g_comparison {
  uint32_t index;    /* raw (unique, load order) */
  uint32_t cluster;  /* Comparison.i */
  uint32_t hamming;  /* 0 when shrouded (cluster.rs:389) */
  uint32_t _pad;
  double lambda;
} d2g_comparison;

/* B::new's store (containers.rs:342): e_minmax = -999.0 for every loaded
 * unique, no kept comparisons.  Needed after each d2g_load_uniques. */
int d2g_store_reset(d2g_ctx* ctx);

/* The compare of centers[0 .. n_centers) (as d2g_compare_batch) followed by
 * the store, the centers acting as successive clusters first_cluster,
 * first_cluster + 1, ... (one b_compare_parallel per center, in order).
 * total_reads is B.reads as f64.  *out_nstored receives the number kept by
 * this call; they are appended to the context's store and, when out_stored
 * is not null, copied to out_stored (capacity records; D2G_ERR_ARG when
 * more were kept, which stay on the device for d2g_store_fetch).
 * out_self_lambda (nullable, n_centers) receives each center's lambda
 * against itself (Bi.self_, cluster.rs:372-374).  B.nalign / B.nshroud grow
 * by stats->comparisons - stats->skipped and stats->shrouded. */
int d2g_compare_store(d2g_ctx* ctx, const uint32_t* centers, uint32_t n_centers,
                      uint32_t first_cluster, double total_reads, const double* err,
                      uint32_t ncol, const d2g_align_params* params, int use_quals, int greedy,
                      const uint8_t* lock, uint64_t capacity, d2g_comparison* out_stored,
                      uint64_t* out_nstored, double* out_self_lambda, d2g_stats* stats);
/* Kept comparisons so far, and a copy of records [first, first + count). */
int d2g_store_size(d2g_ctx* ctx, uint64_t* out_count);
int d2g_store_fetch(d2g_ctx* ctx, uint64_t first, uint64_t count, d2g_comparison* out);

/* ---- the partition loop --------------------------------------------------
 * Replaces run_dada (dada.rs:749-1150) plus the final per-raw p-value pass
 * and read map (dada.rs:502-533), over the loaded uniques (raw index = load
 * order).  Each cluster's compare runs on the device (d2g_compare_center).
 * B, the store, b_p_update, b_bud and the shuffle's moves run on the host;
 * the shuffle's best-cluster scan runs on the device with device_shuffle,
 * else on the host's threads (the same moves either way).
 * DadaParams fields (dada.rs:54-90; defaults main.rs:4393-4433): */
typedef struct {
  double omega_a;        /* 1e-40: bud when min p * nraw < omega_a */
  double omega_p;        /* 1e-4: prior bud when min prior p < omega_p */
  double omega_c;        /* 1e-40: a raw is corrected when its final p >= omega_c */
  double min_fold;       /* 1.0 */
  uint32_t min_hamming;  /* 1 */
  uint32_t min_abund;    /* 1 */
  uint32_t max_clust;    /* 0: up to the number of uniques */
  uint32_t max_shuffle;  /* MAX_SHUFFLE = 10 (dada.rs:39); >= 1 */
  int32_t greedy;        /* 1 */
  int32_t detect_singletons; /* 0 */
  int32_t use_quals;     /* 1 */
  int32_t device_shuffle; /* 0: the shuffle's best-cluster scan on the host's threads;
                           * 1: on the device (same moves, same order) */
} d2g_dada_params;

/* One cluster of the final partition (Bi, containers.rs:200-238). */
typedef struct {
  uint32_t center;       /* unique index of the center */
  uint32_t reads;        /* reads of the members */
  uint32_t nraw;         /* members */
  int32_t birth_type;    /* BirthType (containers.rs:59): 0 initial, 1 abundance, 2 prior */
  uint32_t birth_from;   /* cluster it budded from */
  uint32_t birth_hamming;
  double birth_pval;
  double birth_fold;
  double birth_e;
  double birth_lambda;
  double self_lambda;    /* Bi.self_: the center's lambda against itself */
  double posthoc_pval;   /* post_hoc_pvals (error.rs:145-180): calc_pA of the center's
                          * reads against the expected reads the other clusters'
                          * stored comparisons put on it (1.0 without a center) */
} d2g_cluster_info;

/* Counters and host-clock phase times of one run (the phase timers of
 * run_dada, dada.rs:752-760). */
typedef struct {
  uint64_t nalign;       /* B.nalign */
  uint64_t nshroud;      /* B.nshroud */
  uint32_t compares;     /* device compares (one per cluster) */
  uint32_t shuffle_passes;
  double ms_compare;     /* device compare calls, results on the host */
  double ms_store;       /* serial store */
  double ms_shuffle;
  double ms_bud;
  double ms_pupdate;
  double ms_total;
  uint64_t dp_cells;     /* DP cells of the device compares */
  double ms_dev_screen;  /* summed device kernel times of the compares, */
  double ms_dev_plan;    /* filled when the context's timing is on */
  double ms_dev_dp;
  double ms_dev_lambda;
} d2g_dada_stats;

/* prior: nullable per-unique flags (Raw.prior).  cluster_capacity bounds
 * out_clusters; when the partition has more clusters the call returns
 * D2G_ERR_ARG with *out_nclusters set.  Per unique (n = loaded uniques):
 *   out_cluster_of[i]  final cluster
 *   out_p[i]           final p (1.0 for centers, calc_pA otherwise)
 *   out_map[i]         cluster when corrected (p >= omega_c), else -1
 *   out_comp_lambda/hamming[i]  Raw.comp at the end (nullable)
 * stats is nullable. */
int d2g_run_dada(d2g_ctx* ctx, const double* err, uint32_t ncol, const d2g_align_params* align,
                 const d2g_dada_params* dada, const uint8_t* prior, uint32_t cluster_capacity,
                 uint32_t* out_nclusters, d2g_cluster_info* out_clusters,
                 uint32_t* out_cluster_of, double* out_p, int32_t* out_map,
                 double* out_comp_lambda, uint32_t* out_comp_hamming, d2g_dada_stats* stats);

/* Backs transition_counts (error.rs:194-236) over compute_final_subs
 * (dada.rs:646-673): the learn-errors transition-by-quality matrix of a
 * finished partition.  Per loaded unique i (caller order):
 *   center_of[i]  the unique at the center of i's cluster, or D2G_NO_CENTER
 *   weight[i]     the reads i adds (Raw.reads when Raw.correct), 0 = skip
 * Each counted unique is aligned to its center as compute_aux does
 * (dada.rs:612-616: align with use_kmers = 0 and kdist_cutoff = 1.0, the
 * rest of *align as given).  For each center position aligned to a base,
 * out_mat[(4 (center base - 1) + (base - 1)) * ncol_out + quality] += weight;
 * qualities >= ncol_out are not counted (error.rs:229).  ncol_out = ncol
 * when use_quals is set and the uniques were loaded with qualities, else 1
 * (dada.rs:630); out_mat holds 16 * ncol_out u32 (the sums wrap, as the
 * reference's u32 additions do in a release build). */
#define D2G_NO_CENTER 0xFFFFFFFFu
int d2g_transition_counts(d2g_ctx* ctx, const uint32_t* center_of, const uint32_t* weight,
                          const d2g_align_params* align, int use_quals, uint32_t ncol,
                          uint32_t* out_mat);

/* Backs build_trans_mat (learn_errors.rs:618-690), the per-sample counts
 * each learn-errors iteration sums.  center_of as above; weight[i] the
 * input's abundance when DadaResult.map[i] is set, else 0.  The pairs align
 * as in the reference, whose Raws there carry no k-mer vectors: no k-mer
 * screen (nothing counts when align->use_kmers and kdist_cutoff < 0), gapless
 * only at band 0.  Qualities >= nq count in column nq - 1, the u32 sums
 * saturate, and uniques loaded without qualities add nothing (the matrix is
 * all zero).  out_mat holds 16 * nq u32. */
int d2g_build_trans_mat(d2g_ctx* ctx, const uint32_t* center_of, const uint32_t* weight,
                        const d2g_align_params* align, uint32_t nq, uint32_t* out_mat);

/* Backs compute_aux's per-raw part (dada.rs:610-631): one pass of final
 * Subs (compute_final_subs, dada.rs:646-673, aligned as d2g_transition_counts
 * aligns) over a finished partition of ncl clusters:
 *   centers[k]     the unique at cluster k's center (D2G_NO_CENTER: none)
 *   cluster_of[i]  unique i's cluster (D2G_NO_CENTER: none)
 *   weight[i]      Raw.reads when Raw.correct, else 0 (not counted)
 * Outputs:
 *   out_trans      transition_counts (error.rs:194-236), 16 * ncol_out u32
 *                  as d2g_transition_counts
 *   out_nsubs[i]   Sub.nsubs of each counted unique against its center (the
 *                  n0 / n1 of cluster_stats, error.rs:55-140), else
 *                  D2G_NO_CENTER
 *   out_quality    cluster_quality (error.rs:251-305), ncl * maxlen f64: the
 *                  read-weighted mean quality at each center position below
 *                  min(center length, maxlen), NaN where no counted read
 *                  covers it and everywhere without qualities. */
int d2g_dada_aux(d2g_ctx* ctx, uint32_t ncl, const uint32_t* centers, const uint32_t* cluster_of,
                 const uint32_t* weight, const d2g_align_params* align, int use_quals,
                 uint32_t ncol, uint32_t maxlen, uint32_t* out_trans, uint32_t* out_nsubs,
                 double* out_quality);

/* Backs compute_birth_subs (dada.rs:679-693): each cluster k >= 1 whose
 * center and parent center (centers[birth_from[k]]) exist has its center
 * aligned to the parent's (the k-mer settings of *align, kdist_cutoff 1.0:
 * dada.rs:617-621), and the Sub gives
 *   out_nsubs[k]   its substitutions (D2G_NO_CENTER for cluster 0 and
 *                  clusters without a Sub)
 *   out_qave[k]    cluster_stats' birth_qave (error.rs:90-100): the mean
 *                  query quality of the substitutions, NaN when there are
 *                  none or no qualities are used
 *   the records of birth_sub_records (error.rs:328-352), cluster by
 *   cluster: out_cluster, out_pos (parent position), out_nt0 / out_nt1
 *   (ASCII by nt_decode), out_qual (the child's quality; 0 without
 *   qualities).
 * *out_nrec receives the record count; when it exceeds capacity the call
 * returns D2G_ERR_ARG and the records past capacity are not written. */
int d2g_birth_subs(d2g_ctx* ctx, uint32_t ncl, const uint32_t* centers, const uint32_t* birth_from,
                   const d2g_align_params* align, int use_quals, uint32_t capacity,
                   uint32_t* out_nsubs, double* out_qave, uint32_t* out_cluster, uint32_t* out_pos,
                   uint8_t* out_nt0, uint8_t* out_nt1, uint8_t* out_qual, uint32_t* out_nrec);

/* Counters of the last d2g_transition_counts / d2g_build_trans_mat / d2g_dada_aux call's
 * alignments (comparisons = centers x uniques screened, gapless, aligned_dp,
 * dp_cells) and, with the context's timing on, its kernel times (ms_lambda
 * holds the counting kernel's).  No reference counterpart: diagnostics. */
int d2g_transition_stats(d2g_ctx* ctx, d2g_stats* out);

/* ---- pairwise alignment (parity / aux mode) -------------------------------
 * Backs align_vectorized_with_buf (nwalign.rs:761-1140) for n arbitrary byte
 * pairs (any alphabet: 1..4 or ASCII).  s1/s2 are concatenated with offsets
 * (n+1 entries each).  al0/al1 receive each alignment at al_offsets[i]
 * (caller supplies capacity len1+len2 per pair: al_offsets[i] =
 * sum_{j<i} (len1_j + len2_j)); al_len[i] its length.  al_offsets may be
 * NULL, meaning exactly that packing.  The two sets together must stay
 * below 4 GiB of sequence bytes (D2G_ERR_UNSUPPORTED otherwise), as the
 * compare's unique set must. */
int d2g_align_vectorized(d2g_ctx* ctx, uint32_t n, const uint8_t* s1, const uint64_t* s1_off,
                         const uint8_t* s2, const uint64_t* s2_off, const d2g_vec_scores* scores,
                         uint8_t* al0, uint8_t* al1, const uint64_t* al_offsets,
                         uint32_t* al_len);

/* DP cell count of the vectorized loop for one pair (sum over anti-diagonals
 * of the active width; nwalign.rs:875-1088) — the GCUPS unit. */
uint64_t d2g_vectorized_cells(uint32_t len1, uint32_t len2, int32_t band);

/* 1 when a device kernel aligns this geometry with the vectorized aligner:
 * both lengths <= 3500 (nwalign.rs:1383) and a compressed band of at most
 * 4096 columns and 32 Mi pointer cells (up to 64 columns in the
 * many-pairs-per-wavefront kernel, wider in the one-pair kernel), 0
 * otherwise.  Pairs outside it make d2g_align_vectorized return
 * D2G_ERR_UNSUPPORTED; the compare routes pairs above 3500 nt to the
 * scalar ends-free aligner as raw_align_dp does. */
int d2g_vectorized_supported(uint32_t len1, uint32_t len2, int32_t band);

/* Backs align_endsfree_with_buf (nwalign.rs:336-417, when homo_gap_p ==
 * gap_p) and align_endsfree_homo_with_buf (nwalign.rs:435-528, homopolymer
 * runs of >= 3 bytes take homo_gap_p) for n pairs: i32 scores, band < 0
 * unbanded, free end gaps on the last row and column, no swap.  Buffers as
 * for d2g_align_vectorized. */
int d2g_align_endsfree(d2g_ctx* ctx, uint32_t n, const uint8_t* s1, const uint64_t* s1_off,
                       const uint8_t* s2, const uint64_t* s2_off, int32_t match, int32_t mismatch,
                       int32_t gap_p, int32_t homo_gap_p, int32_t band, uint8_t* al0,
                       uint8_t* al1, const uint64_t* al_offsets, uint32_t* al_len);

/* 1 when the row-band kernel handles this geometry (band diagonals plus the
 * two sentinel diagonals fit 512 = 4 cells x 2 parities x 64 lanes), 0
 * otherwise.  The compare
 * routes to it for homopolymer gap penalties (homo = 1), for vectorized = 0,
 * and when every unique is longer than 3500 nt (nwalign.rs:1335-1414). */
int d2g_band_supported(uint32_t len1, uint32_t len2, int32_t band, int homo);

/* ---- differentiable likelihood ---------------------------------------------
 * Over the aligned pairs of the last d2g_compare_* call (gapless and DP
 * pairs; skipped and shrouded pairs take no part), with logE a 16 x ncol
 * row-major FP32 log error matrix (row t = 4*ref + query, column = quality):
 *   log lambda_j = sum over query positions of logE[t][q] with the
 *                  transitions of compute_lambda (pval.rs:294-309),
 *   E_j = reads[center] * lambda_j,
 *   loss = sum_j reads_j log E_j - E_j - lgamma(reads_j + 1)
 * (the Poisson log-likelihood of the abundances).  d2g_loglik_backward
 * returns d loss / d logE[t][q] = sum_j (reads_j - E_j) count_j[t][q] for the
 * logE of the last forward; the counts are recomputed from the alignments.
 * Sums are FP64.  No counterpart in the reference, which fits the error
 * model outside this path (learn_errors.rs); ms (nullable) receives the
 * kernel time. */
int d2g_loglik_forward(d2g_ctx* ctx, const float* logE, uint32_t ncol, int use_quals,
                       double* out_loss, float* ms);
int d2g_loglik_backward(d2g_ctx* ctx, float* out_grad, double* out_grad64, float* ms);

/* ---- bimera parent search -------------------------------------------------
 * BimeraAlignParams (chimera.rs:183-196) with the NW backend.  Sequences are
 * ASCII; each (query, parent) pair is aligned by the vectorized aligner with
 * end_gap 0 and band = max_shift (chimera.rs:234-240, 336-342). */
typedef struct d2g_bimera_params {
  int32_t allow_one_off;
  int32_t min_one_off_par_dist;
  int16_t match_score;
  int16_t mismatch;
  int16_t gap_p;
  int16_t _pad;
  int32_t max_shift;
} d2g_bimera_params;

/* Per pair i (query q[q_off[i]..q_off[i+1]), parent p[...]): get_lr
 * (chimera.rs:99-175) into out_lr[4i..4i+3] = left, right, left_oo,
 * right_oo, and get_ham_endsfree (chimera.rs:54-86) into out_ham[i]. */
int d2g_bimera_pairs(d2g_ctx* ctx, uint32_t n, const uint8_t* q, const uint64_t* q_off,
                     const uint8_t* p, const uint64_t* p_off, const d2g_bimera_params* params,
                     uint32_t* out_lr, uint32_t* out_ham);

/* is_bimera_with_buf (chimera.rs:218-292) for nq queries: query i's parents
 * are p entries par_start[i] .. par_start[i+1]-1 (par_start has nq+1
 * entries; p_off has par_start[nq]+1).  out_flag[i] = 1 when bimera. */
int d2g_is_bimera_batch(d2g_ctx* ctx, uint32_t nq, const uint8_t* q, const uint64_t* q_off,
                        const uint8_t* p, const uint64_t* p_off, const uint32_t* par_start,
                        const d2g_bimera_params* params, uint8_t* out_flag);

/* The same search with the parents given as indices into one pool of npool
 * sequences (pool_off has npool+1 entries): query i's parents are
 * pool[par_idx[t]] for t in par_start[i] .. par_start[i+1]-1.  A parent
 * shared by many queries is copied to the device once, as the reference's
 * is_bimera borrows its parents' slices (chimera.rs:218-222). */
int d2g_is_bimera_pool(d2g_ctx* ctx, uint32_t nq, const uint8_t* q, const uint64_t* q_off,
                       uint32_t npool, const uint8_t* pool, const uint64_t* pool_off,
                       const uint32_t* par_idx, const uint32_t* par_start,
                       const d2g_bimera_params* params, uint8_t* out_flag);

/* table_bimera2 (chimera.rs:317-433): mat is column-major nrow (samples) x
 * ncol (sequences); seqs holds the ncol sequences with off[ncol+1].  Per
 * column j: out_nsam[j] samples where j is present, out_nflag[j] of them
 * where j is a bimera of more abundant parents (> min_fold x and >=
 * min_abund reads in that sample). */
int d2g_table_bimera2(d2g_ctx* ctx, const uint32_t* mat, uint32_t nrow, uint32_t ncol,
                      const uint8_t* seqs, const uint64_t* off, double min_fold,
                      uint32_t min_abund, const d2g_bimera_params* params,
                      uint32_t* out_nflag, uint32_t* out_nsam);

/* ---- mixed-precision lambda and abundance p-values ----------------------
 * A resident set of n aligned pairs: query i (bases 1..4, qualities, read
 * count reads[i]) with its al2subs list against its center (records
 * (query position << 16 | center base 1..4), ascending query position;
 * sub_offsets has n+1 entries) and the center's read count center_reads[i].
 * d2g_pairs_lambda computes, on the device:
 *   loglam32[i]  log compute_lambda (pval.rs:253-320) as an FP32
 *                compensated (Kahan) sum of FP32-rounded log err[t][q]
 *   lambda[i]    exp(loglam32[i]) in FP64, or the exact ordered FP64 product
 *                when recomputed
 *   pval[i]      get_pA(reads, prior = false, lambda, hamming = #subs,
 *                center_reads, detect_singletons = false) (pval.rs:376-398,
 *                calc_pA pval.rs:208-241) in FP64
 *   recomputed[i] 1 when the first p-value fell within a factor 10 of
 *                omega_a, so lambda was recomputed exactly in FP64 and the
 *                p-value taken again.
 * The reference computes lambda only as the FP64 product; this is the
 * reduced-precision variant with its FP64 safeguard. */
typedef struct d2g_pair_stats {
  uint64_t pairs;
  uint64_t recomputed;
  float ms_loglam;  /* FP32 log-lambda kernel */
  float ms_pval;    /* p-value kernel and the FP64 recompute of the pairs near omega_a */
  float ms_total;
  float _pad;
} d2g_pair_stats;

int d2g_pairs_load(d2g_ctx* ctx, uint32_t n, const uint8_t* seq_concat, const uint64_t* offsets,
                   const uint8_t* qual_concat, const uint32_t* subs, const uint64_t* sub_offsets,
                   const uint32_t* reads, const uint32_t* center_reads);
int d2g_pairs_lambda(d2g_ctx* ctx, const double* err, uint32_t ncol, int use_quals,
                     double omega_a, d2g_pair_stats* stats);
int d2g_pairs_fetch(d2g_ctx* ctx, float* out_loglam32, double* out_lambda, double* out_pval,
                    uint8_t* out_recomputed);

/* ---- allocation-free variant ----------------------------------------------
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  d2g_pairs_lambda_noalloc computes what
 * d2g_pairs_load + d2g_pairs_lambda + d2g_pairs_fetch compute, with the same
 * kernels and the same bits, over a pair set the caller already holds in
 * device memory, and writes the outputs into the caller's device memory.
 * It calls no hipMalloc / hipFree: its intermediates (error matrix, the
 * log(reads!) table, the list of pairs near omega_a, counters) live in one
 * library-owned device array per GPU of about 12.1 MiB, held until the
 * process ends and set up by d2g_ctx_create.  Every call on one GPU shares
 * that array: calls that overlap on one GPU race.  Limits (beyond them the
 * call returns D2G_ERR_UNSUPPORTED): */
#define D2G_NOALLOC_MAX_PAIRS (1u << 20)
#define D2G_NOALLOC_MAX_READS ((1u << 20) - 1u)
#define D2G_NOALLOC_MAX_NCOL 256u

/* The pair set in device memory, in d2g_pairs_load's resident layout:
 * pair i's bases (1..4) at seq[off[i] .. off[i] + len[i]), off[i] a
 * multiple of 16 and the row readable up to the next multiple of 16 inside
 * pool_bytes; qual (nullable) in the same layout; its al2subs records
 * subs[sub_off[i] .. sub_off[i+1]) as in d2g_pairs_load, inside n_subs
 * records.  off and sub_off have n + 1 entries, len / reads /
 * center_reads n.  Offsets, lengths, bases, qualities (< ncol when used),
 * records and reads (>= 1) are checked on the device before any row is
 * read: D2G_ERR_ARG / D2G_ERR_NONACGT as d2g_pairs_load returns them. */
typedef struct d2g_pairs_dev {
  uint32_t n;
  uint32_t _pad;
  uint64_t pool_bytes;
  uint64_t n_subs;
  const uint8_t* seq;
  const uint8_t* qual;
  const uint64_t* off;
  const uint32_t* len;
  const uint32_t* subs;
  const uint64_t* sub_off;
  const uint32_t* reads;
  const uint32_t* center_reads;
} d2g_pairs_dev;

/* err: host memory, 16 x ncol f64.  out_*: device memory, n entries each,
 * as d2g_pairs_fetch returns them.  The call returns after its stream has
 * finished. */
int d2g_pairs_lambda_noalloc(d2g_ctx* ctx, const d2g_pairs_dev* pairs, const double* err,
                             uint32_t ncol, int use_quals, double omega_a, float* out_loglam32,
                             double* out_lambda, double* out_pval, uint8_t* out_recomputed,
                             d2g_pair_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* DADA2_AMD_H */
