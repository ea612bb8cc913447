// d2g_launch.h — host-side launchers of the gfx950 kernels (one per .hip
// translation unit), used by the C ABI in api.hip.
#pragma once

#include "d2g_internal.h"

struct d2g_ctx;

namespace d2g {

// ---- prep.hip ----------------------------------------------------------
// Builds the resident screen of every unique (raw_assign_kmers,
// kmers.rs:213-221) in two layouts: row-major (centers, read with scalar
// loads) and unique-interleaved (queries, coalesced across lanes).
struct UniqueDev {
  uint32_t n = 0, npad = 0, k = 0;
  uint32_t kwords = 0;      // 4^k / 4 words of the dense kmer8 vector
  uint32_t owords = 0;      // ceil((maxlen - k + 1) / 2) words of kord
  uint32_t minlen = 0, maxlen = 0;
  uint8_t* seq = nullptr;   // sorted-order concatenation
  uint8_t* qual = nullptr;  // nullable
  uint64_t* off = nullptr;  // n + 1
  uint32_t* len = nullptr;
  uint32_t* reads = nullptr;
  uint32_t* orig = nullptr;  // sorted -> original index
  uint32_t* kmer_row = nullptr;  // [n][kwords]
  uint32_t* kmer_il = nullptr;   // [kwords][npad]
  uint32_t* kord_row = nullptr;  // [n][owords]
  uint32_t* kord_il = nullptr;   // [owords][npad]
  uint32_t* ksum = nullptr;      // sum of kmer8 bytes
  uint8_t* has255 = nullptr;
  uint8_t* bad_base = nullptr;   // per unique: a base outside 1..4
};

hipError_t launch_prep(const UniqueDev& u, hipStream_t s);

// ---- screen.hip --------------------------------------------------------
// Task slots per region of the pairwise align path's DP task list.
constexpr uint32_t kAlignRegion = 4096;

struct ScreenArgs {
  const UniqueDev* u;  // host copy of the pointers
  const uint32_t* centers;  // sorted indices, n_centers
  uint32_t n_centers;
  const uint8_t* lock;  // sorted order, nullable
  int greedy;
  int use_kmers;
  int gapless;
  double kdist_cutoff;
  int band;
  BinSpace bins;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
  GaplessTask* gapless_tasks;  // per block region (unsorted)
  DpTask* dp_unsorted;         // per block region
  uint32_t* bin_count;
  uint32_t* partials;  // per block: skipped, shrouded, gapless, DP (screen_blocks entries)
  Counters* ctr;
  uint32_t qblock = 64;   // uniques per workgroup: 64, 128 or 256
  int split = 4;          // wavefront groups per workgroup sharing its 16 centers: 1, 2 or 4
  int per_lane = 1;       // uniques per lane: 1, or 2 (qblock 128 or 256)
  const uint32_t* dstar = nullptr;  // [maxlen - k + 2]: shroud iff dot < dstar[n] (use_kmers)
  // the pairs of this pass (the compare splits at the vectorized aligner's
  // length limit): a center of length <= long_len pairs with sorted uniques
  // [sr_lo, sr_hi), a longer one with [lr_lo, lr_hi)
  uint32_t long_len = 0xFFFFFFFFu, sr_lo = 0, sr_hi = 0xFFFFFFFFu, lr_lo = 0, lr_hi = 0xFFFFFFFFu;
  // one center (screen1_fits): the single-center screen, whose task regions
  // are screen1_regions(n) of kS1Uniques slots
  bool single = false;
};
// dstar[n] for n = 0 .. nmax: the smallest dot in 0 .. n with
// 1.0 - (double)dot / (double)n <= cutoff in f64 (n + 1 when there is none);
// dstar[0] = 0.
void screen_thresholds(double cutoff, uint32_t nmax, uint32_t* out);
hipError_t launch_screen(const ScreenArgs& a, hipStream_t s);
// Workgroups of the screen grid; each owns screen_region(qblock) task slots
// in each list.
uint32_t screen_blocks(uint32_t n_uniques, uint32_t n_centers, uint32_t qblock);
constexpr uint32_t kScreenCenters = 16;  // centers per screen workgroup
inline uint32_t screen_region(uint32_t qblock) { return qblock * kScreenCenters; }
// The single-center screen: kS1Uniques uniques (task slots) per wavefront
// region, kS1Waves wavefronts per workgroup; k <= 5 (dense vectors of at
// most 256 words).
constexpr uint32_t kS1Uniques = 16;
constexpr uint32_t kS1Waves = 4;
bool screen1_fits(uint32_t n_centers, int use_kmers, uint32_t kwords);
uint32_t screen1_regions(uint32_t n_uniques);

// The pairing of the final-subs and birth-subs passes in place of the
// screen: unique q (sorted index) with the one center want[q] (0xFFFFFFFF:
// none), tasks with out_idx = q and no results.  The pair's k-mer distances
// (with use_kmers) decide the shroud and the gapless route as the screen's
// do.  Same outputs as the screen for the plan and scatter: kPairBlock
// uniques per workgroup, each owning kPairBlock task slots in each list.
constexpr uint32_t kPairBlock = 256;
inline uint32_t pair_blocks(uint32_t n_uniques) { return (n_uniques + kPairBlock - 1) / kPairBlock; }
hipError_t launch_pair_screen(const ScreenArgs& a, const uint32_t* want, hipStream_t s);

// ---- plan.hip ----------------------------------------------------------
hipError_t launch_zero(uint32_t* p, size_t n_words, hipStream_t s);
hipError_t launch_zero2(uint32_t* a, size_t na, uint32_t* b, size_t nb, hipStream_t s);
// The task lists to compact in the plan kernel itself, for a screen whose
// regions already list their DP tasks in bin order (one center: uniques are
// length-sorted, so a region's tasks and the regions follow the bins).
struct PlanCopy {
  const DpTask* dp_unsorted;
  DpTask* dp_sorted;
  const GaplessTask* gl_unsorted;
  GaplessTask* gl;
  uint32_t region;  // task slots per region
};
// Also totals the DP cells (info: per-bin cells, nullable) and, for a
// compare, the screen's per-block counts (partials: 4 words per block,
// nullable), writing each block's gapless offset to gl_off.  With copy
// (partials required), the plan kernel also moves the regions' DP and
// gapless tasks to their places, and no scatter is needed.
hipError_t launch_plan(const BinSpace& bs, const uint32_t* bin_count, uint32_t* bin_start,
                       uint32_t* wave_start, uint32_t* bin_fill, Counters* ctr,
                       const BinInfo* info, const uint32_t* partials, uint32_t n_partials,
                       uint32_t* gl_off, hipStream_t s, const PlanCopy* copy = nullptr);
// Bins the DP tasks of n_regions regions of `region` slots (counts from
// partials, or `count` tasks filling the regions in order when partials is
// null) and compacts the gapless tasks (compare only).
hipError_t launch_scatter(const DpTask* unsorted, DpTask* sorted, const uint32_t* bin_start,
                          uint32_t* bin_fill, const uint32_t* partials, uint32_t n_regions,
                          uint32_t region, uint32_t count, const GaplessTask* gl_unsorted,
                          GaplessTask* gl, const uint32_t* gl_off, hipStream_t s);
hipError_t launch_unpermute(uint32_t nc, uint32_t n, const uint32_t* orig, const double* src_lam,
                            const uint32_t* src_ham, const uint8_t* src_skp, double* dst_lam,
                            uint32_t* dst_ham, uint8_t* dst_skp, hipStream_t s,
                            const uint32_t* ctr_src = nullptr, uint32_t* ctr_dst = nullptr,
                            uint32_t ctr_words = 0);

// ---- nw_vec.hip --------------------------------------------------------
struct NwArgs {
  BinSpace bins;
  const uint32_t* bin_start;   // nbins + 1
  const uint32_t* wave_start;  // nbins + 1
  const DpTask* tasks;         // sorted by bin
  const uint8_t* pool;         // sequence bytes
  const RowRec* sched;         // row schedules (launch_sched)
  const uint64_t* sched_off;   // per bin: its first record in sched
  const BinInfo* info;         // per-bin steady range and cell count (launch_sched)
  int match, mismatch, gap_p, end_gap_p;
  int homo_gap_p;              // row-band kernel with bins.homo
  Counters* ctr;
  int emit_alignment;
  // compare mode: per sorted task ti, substitution list (query_pos << 16 |
  // center_pos) at sub_pool[ti * sub_stride]
  uint32_t* sub_pool;
  size_t sub_capacity;
  uint32_t sub_stride;     // longest sequence of the bin space + 1
  uint32_t* sub_cnt;       // per task: count, bit 31 = invalid traceback
  // pointer history scratch: grid slots of hist_stride words
  uint32_t* hist;
  size_t hist_stride;
  // align mode
  uint8_t* al0;
  uint8_t* al1;
  const uint64_t* al_off;  // per pair (out_idx)
  uint32_t* al_len;        // per pair
  uint32_t grid;           // persistent workgroups
  size_t lds_bytes;        // from nw_vec_lds_for
};
hipError_t launch_nw_vec(const NwArgs& a, hipStream_t s);
// Replays the vectorized row walk of every bin with off[bin] != kSchedAbsent
// into table[off[bin] + row - 2] (nrow records per bin); a bin nw_vec cannot
// hold gets only its cell count.
hipError_t launch_sched(const BinSpace& bs, const uint64_t* off, int end_gap, int gap,
                        RowRec* table, BinInfo* info, hipStream_t s);
size_t nw_vec_lds_for(const BinSpace& bs, const uint8_t* present);

// ---- nw_wide.hip -------------------------------------------------------
// The vectorized aligner for the bins nw_vec cannot hold (wide_fits), one
// pair per wavefront: the largest such geometry of a bin space sizes the
// LDS rows and the per-wavefront pointer slot (0: none in the space).
struct WideNeed {
  int max_ncol;
  size_t pslot;
};
WideNeed nw_wide_need(const BinSpace& bs, const uint8_t* present);
hipError_t launch_nw_wide(const NwArgs& a, const WideNeed& need, uint8_t* pbuf, uint32_t grid,
                          hipStream_t s);

// ---- nw_band.hip -------------------------------------------------------
// The scalar ends-free aligners (plain / homopolymer by bins.homo) over the
// same binned tasks and arguments as launch_nw_vec (sched unused).
hipError_t launch_nw_band(const NwArgs& a, hipStream_t s);
// BinInfo.cells of every bin for the row-band kernel.
hipError_t launch_band_info(const BinSpace& bs, BinInfo* info, hipStream_t s);
size_t nw_band_lds_for(const BinSpace& bs, const uint8_t* present);

// ---- lambda.hip --------------------------------------------------------
struct LambdaArgs {
  const UniqueDev* u;
  const GaplessTask* gapless_tasks;
  const DpTask* dp_tasks;  // sorted
  const uint32_t* sub_pool;  // task di's list at sub_pool[di * sub_stride]
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
  const double* err;
  uint32_t ncol;
  int use_quals;
  Counters* ctr;
  double* out_lambda;
  uint32_t* out_hamming;
  uint8_t* out_skipped;
  uint32_t grid;
};
hipError_t launch_lambda(const LambdaArgs& a, hipStream_t s);

// ---- transition.hip ----------------------------------------------------
// transition_counts (error.rs:194-236) over the tasks of a filtered compare
// whose DP pairs wrote their alignment strings at al_off[unique].
struct TransArgs {
  const UniqueDev* u;
  const GaplessTask* gapless_tasks;
  const DpTask* dp_tasks;  // sorted
  const Counters* ctr;
  const uint8_t* al0;
  const uint8_t* al1;
  const uint64_t* al_off;
  const uint32_t* al_len;
  const uint32_t* weight;  // per sorted unique
  int use_quals;
  uint32_t ncol;            // <= 256 (qualities are bytes)
  int clamp;                // qualities >= ncol: clamped to ncol - 1, else skipped
  unsigned long long* mat;  // [16][ncol], exact sums (zeroed by the caller)
  uint32_t grid;
};
hipError_t launch_transition(const TransArgs& a, hipStream_t s);

// compute_aux's per-raw sums over the same alignments: Sub.nsubs per
// counted unique and cluster_quality's sums per (cluster, center position).
struct AuxArgs {
  const UniqueDev* u;
  const uint8_t* al0;
  const uint8_t* al1;
  const uint64_t* al_off;
  const uint32_t* al_len;
  const uint32_t* order;    // counted sorted uniques, grouped by cluster
  uint32_t n_order;
  const uint32_t* cluster;  // per sorted unique
  const uint32_t* want;     // per sorted unique: its center
  const uint32_t* weight;   // per sorted unique
  int gapless;              // band 0
  int use_quals;
  uint32_t maxlen;          // row length of qual_reads / qual_sum (>= every center length)
  uint32_t* nsubs;          // per sorted unique
  uint32_t* qual_reads;     // [clusters][maxlen], zeroed by the caller
  unsigned long long* qual_sum;
  uint32_t grid;
};
hipError_t launch_aux(const AuxArgs& a, hipStream_t s);

// ---- shuffle.hip -------------------------------------------------------
// The best-cluster scan of a shuffle pass over the stored comparisons, the
// movers compacted into `moves`.  One stored comparison, and one move:
struct ShuffleComp {
  uint32_t raw, cl, ham, _pad;
  double lambda;
};
struct ShuffleMove {
  uint32_t r, to, ham, _pad;
  double lambda;
};
struct ShuffleArgs {
  const ShuffleComp* comps;  // appended per compare (cluster-major)
  uint32_t n_comp;
  const double* creads;      // per cluster: its reads
  const uint32_t* ccenter;   // per cluster: its center raw
  uint32_t n_cl;
  uint32_t n_raw;
  unsigned long long* best_bits;  // per raw, scratch (reset by each pass)
  unsigned long long* best_key;   // cluster << 32 | comparison index
  uint32_t* where;           // per raw: its cluster (centers set by the pass)
  ShuffleMove* moves;        // slot 0 holds the count in `r`; movers from slot 1
};
// ---- store.hip ---------------------------------------------------------
// The store of b_compare_parallel (cluster.rs:351-397) over the last
// compare's results: nc centers acting as successive clusters.
struct StoreArgs {
  uint32_t n;               // raws
  uint32_t nc;              // centers of the compare
  const uint32_t* inv;      // caller -> sorted unique index
  const uint32_t* centers;  // sorted indices of the centers
  const uint32_t* reads;    // sorted order
  const double* lam;        // [nc][n], sorted order (the compare's results)
  const uint32_t* ham;      // [nc][n], sorted order
  double total;             // B.reads as f64
  double* e_minmax;         // [n] sorted order, resident
  uint8_t* flag;            // [nc][n] sorted order
  uint32_t* bcount;         // [nc][store_blocks(n)] + 1: counts, then offsets and the total
  double* self_out;         // [nc]: each center's lambda against itself
  ShuffleComp* store;       // where this call's first kept comparison goes
  uint32_t first_cluster;   // cluster of center 0
};
uint32_t store_blocks(uint32_t n);
hipError_t launch_store(const StoreArgs& a, hipStream_t s);
hipError_t launch_store_reset(double* e_minmax, uint32_t n, hipStream_t s);
// lock[idx[j]] = 1 for the k sorted indices idx (the partition loop's new
// greedy locks).
hipError_t launch_lock_set(uint8_t* lock, const uint32_t* idx, uint32_t k, hipStream_t s);

hipError_t launch_shuffle_init(const ShuffleArgs& a, hipStream_t s);  // where = 0, scratch reset
hipError_t launch_shuffle_pass(const ShuffleArgs& a, hipStream_t s);

// ---- bimera.hip --------------------------------------------------------
// get_lr + get_ham_endsfree of n device-resident alignments (5 words each).
hipError_t launch_bimera_lr(uint32_t n, const uint8_t* al0, const uint8_t* al1,
                            const uint64_t* al_off, const uint32_t* al_len, int allow_one_off,
                            uint32_t max_shift, uint32_t* out, hipStream_t s);

// ---- loglik.hip --------------------------------------------------------
// Poisson log-likelihood of the abundances over the last compare's aligned
// pairs (gapless + DP tasks) and its gradient in the 16 x ncol log error
// matrix.
struct LoglikArgs {
  const UniqueDev* u;
  const GaplessTask* gapless_tasks;
  const DpTask* dp_tasks;  // sorted, with their substitution lists
  const uint32_t* sub_pool;
  uint32_t sub_stride;
  const uint32_t* sub_cnt;
  uint32_t ncol;
  int use_quals;
  const Counters* ctr;     // n_gapless, n_dp of that compare
  double* w;               // per task scratch: a - E
  double* block_part;      // grid * 16 * ncol doubles
  uint32_t grid;
};
hipError_t launch_loglik_fwd(const LoglikArgs& a, const float* logE, double* loss, hipStream_t s);
hipError_t launch_loglik_bwd(const LoglikArgs& a, double* grad, hipStream_t s);

// ---- lambda32.hip ------------------------------------------------------
// FP32 Kahan log-lambda and FP64 p-values (with FP64 lambda recompute near
// omega) over a resident set of aligned pairs.
struct PairArgs {
  uint32_t n;
  const uint8_t* seq;       // rows padded to 16 bytes at off[i]
  const uint8_t* qual;      // nullable
  const uint64_t* off;      // n + 1
  const uint32_t* len;
  const uint32_t* subs;     // (query pos << 16 | center base), ascending
  const uint64_t* sub_off;  // n + 1
  const uint32_t* reads;
  const uint32_t* center_reads;
  const double* err;
  uint32_t ncol;
  int use_quals;
  double omega;
  float* out_loglam;
  double* out_lambda;
  double* out_p;
  uint8_t* out_flag;
  uint32_t* n_recomputed;   // zeroed by the caller; counts `list`
  uint32_t* list;           // n: pairs whose first p-value fell near omega
  const double* log_kfact;  // [max reads + 1]: sum_{i=1..r} log(i) in running order
  uint32_t max_blocks;
  hipEvent_t ev_mid;        // nullable: recorded between the two kernels
};
hipError_t launch_pairs_lambda(const PairArgs& a, hipStream_t s);

// ---- api.hip / noalloc.hip ------------------------------------------------
// What a kernel launch needs from a context (api.hip owns the struct).
struct CtxExec {
  int device;
  hipStream_t stream;
  int cu_count;
  bool timing;
  hipEvent_t ev[3];
};
bool ctx_exec(const d2g_ctx* c, CtxExec* out);
// The context's resident store of kept comparisons (d2g_compare_store).
bool ctx_store(const d2g_ctx* c, const ShuffleComp** comps, uint64_t* count);
// Resident lock flags for the partition loop: begin clears the device lock
// array and makes compares with a null lock argument read it; add sets the
// flags of k raws (caller order; flags are only ever set); end returns the
// context to per-call lock uploads.
bool ctx_lock_begin(d2g_ctx* c);
bool ctx_lock_add(d2g_ctx* c, const uint32_t* raws, uint32_t k);
void ctx_lock_end(d2g_ctx* c);
// Once per device, with that device current: the address of the
// allocation-free entry point's device array and its constant table.
int noalloc_setup(int device);

}  // namespace d2g
