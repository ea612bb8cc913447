// api.hip — the C ABI declared in include/dada2_amd.h.
//
// A context owns one HIP stream and every device buffer.  The compare
// pipeline per call is: screen -> plan -> scatter -> vectorized DP ->
// lambda, all on the context stream; host work is limited to argument
// checks, uploading the center list / lock flags / error matrix, and the
// final copy of results (skipped by the device-resident entry point).
#include <algorithm>
#include <cmath>
#include <array>
#include <cstddef>
#include <cstring>
#include <map>
#include <numeric>
#include <limits>
#include <vector>

#include "../../include/dada2_amd.h"
#include "d2g_launch.h"

using namespace d2g;

namespace {

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e == hipSuccess) cap = std::max<size_t>(n, 1);
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

constexpr int kEvents = 6;

}  // namespace

struct d2g_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  hipEvent_t ev[kEvents] = {};
  int cu_count = 0, clock_khz = 0, mem_clock_khz = 0, mem_bus_width = 0;
  // uniques
  bool loaded = false;
  UniqueDev u;
  std::vector<uint32_t> h_len;   // sorted order
  std::vector<uint32_t> h_orig;  // sorted -> original
  std::vector<uint32_t> h_inv;   // original -> sorted
  std::vector<uint32_t> h_reads; // read counts, caller order
  uint32_t max_qual = 0;
  bool has_qual = false;
  DevBuf<uint8_t> seq, qual, has255;
  DevBuf<uint64_t> off;
  DevBuf<uint32_t> len, reads, orig, kmer_row, kmer_il, kord_row, kord_il, ksum;
  // per call
  DevBuf<uint32_t> centers;
  DevBuf<uint8_t> lock;
  DevBuf<double> err;
  DevBuf<double> out_lambda;  // length-sorted unique order
  DevBuf<uint32_t> out_hamming;
  DevBuf<uint8_t> out_skipped;
  DevBuf<double> fin_lambda;  // caller's unique order (filled by fetch)
  DevBuf<uint32_t> fin_hamming;
  DevBuf<uint8_t> fin_skipped;
  DevBuf<GaplessTask> gl, gl_unsorted;
  DevBuf<uint32_t> gl_off;
  DevBuf<DpTask> dp_unsorted, dp_sorted;
  DevBuf<uint32_t> bin_count, bin_start, wave_start, bin_fill;
  DevBuf<uint32_t> sub_pool, sub_cnt, screen_partials;
  DevBuf<Counters> ctr;
  size_t sub_capacity = 0;
  uint32_t last_nc = 0;
  // cached per bin space: LDS sizing and the row schedule of the DP kernel
  BinSpace lds_key = {-1, -1, -1, -1, -1, -1, -1};
  int sched_gap = 0, sched_end_gap = 0;
  size_t lds_bytes = 0;
  int sched_rows = 0;
  std::vector<uint8_t> sched_present;  // lengths the cached schedule covers
  DevBuf<uint64_t> sched_off;          // per bin: first record, or kSchedAbsent
  DevBuf<RowRec> sched;
  DevBuf<BinInfo> sched_info;
  // pointer-history scratch of the DP: nw_grid slots of hist_stride words
  uint32_t nw_grid = 0;
  size_t hist_stride = 0;
  DevBuf<uint32_t> hist;
  // the same state for the other DP kernel (a compare over a unique set on
  // both sides of the vectorized length limit alternates the two)
  BinSpace st_key = {-1, -1, -1, -1, -1, -1, -1};
  int st_gap = 0, st_end_gap = 0;
  size_t st_lds_bytes = 0;
  int st_rows = 0;
  std::vector<uint8_t> st_present;
  DevBuf<uint64_t> st_off;
  DevBuf<RowRec> st_sched;
  DevBuf<BinInfo> st_info;
  uint32_t st_grid = 0;
  size_t st_hist_stride = 0;
  DevBuf<uint32_t> st_hist;
  // the one-pair kernel's pointer slots (vectorized bins nw_vec cannot hold)
  WideNeed wide = {0, 0};
  uint32_t wide_grid = 0;
  DevBuf<uint8_t> wide_p;
  WideNeed st_wide = {0, 0};
  uint32_t st_wide_grid = 0;
  DevBuf<uint8_t> st_wide_p;
  // resident aligned-pair set (mixed-precision lambda / p-values)
  uint32_t p_n = 0, p_max_qual = 0;
  bool p_has_qual = false;
  DevBuf<uint8_t> p_seq, p_qual, p_flag;
  DevBuf<uint64_t> p_off, p_suboff;
  DevBuf<uint32_t> p_len, p_subs, p_reads, p_creads, p_nrec, p_list;
  DevBuf<double> p_lkf;  // sum_{i=1..r} log(i) for r = 0 .. max reads
  DevBuf<float> p_loglam;
  DevBuf<uint32_t> b_out;  // bimera: 5 words per pair
  // the compare's launch sequence as a replayable HIP graph, per shape key
  bool graphs = true;
  // screen_kernel's launch shape: uniques per workgroup, and wavefront
  // groups sharing its 16 centers
  uint32_t screen_qblock = 64;
  int screen_split = 4;
  int screen_per_lane = 1;
  // autotuned screen shape per workload class (qblock, split, per_lane)
  bool autotune = false;
  std::map<uint64_t, std::array<int, 3>> tuned;
  // autotuned DP tile (BinSpace.tile) per workload class; forced_tile >= 0
  // overrides it (d2g_ctx_set_dp_tile)
  std::map<uint64_t, int> tuned_tile;
  int forced_tile = -1;
  // the screen's shroud thresholds, for one (cutoff, k, longest unique)
  DevBuf<uint32_t> dstar;
  uint64_t dstar_cut = 0;
  uint32_t dstar_n = 0;
  bool dstar_valid = false;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint64_t> graph_key;
  // pinned host staging for per-call uploads (centers, lock flags) and the
  // last error matrix uploaded (skipped when unchanged)
  uint8_t* h_pin = nullptr;
  size_t h_pin_cap = 0;
  std::vector<double> h_err;
  // the center list (sorted indices) last uploaded into `centers`, and the
  // buffer it went to (skipped when both are unchanged)
  std::vector<uint32_t> h_centers;
  const uint32_t* h_centers_buf = nullptr;
  // pinned landing place of the compare's counters (a pageable target costs
  // a staged copy per call)
  Counters* h_ctr = nullptr;
  // the last compare's aligned pairs (for the log-likelihood and gradient);
  // its counters go up to last_ctr when a likelihood entry first needs them
  bool cmp_valid = false;
  uint32_t cmp_sub_stride = 0;
  uint64_t cmp_pairs = 0;
  Counters cmp_ctr{};
  bool cmp_ctr_uploaded = false;
  DevBuf<Counters> last_ctr;
  // small result sets come back as one packed block (lambda, hamming,
  // skipped, counters) through pinned memory
  DevBuf<uint8_t> fetch_dev;
  uint8_t* fetch_pin = nullptr;
  size_t fetch_pin_cap = 0;
  DevBuf<double> ll_w, ll_part, ll_out;
  DevBuf<float> ll_logE;
  uint32_t ll_ncol = 0;
  int ll_use_quals = 0;
  bool ll_have_w = false;
  DevBuf<double> p_lambda, p_pval, p_err;
  // transition counts: per-unique filter and weights, the 16 x ncol matrix
  // host copies of the packed uniques (sorted order, 16-byte rows)
  std::vector<uint8_t> h_seq, h_qual;
  std::vector<uint64_t> h_soff;
  DevBuf<uint32_t> tr_want, tr_weight, tr_order, tr_cluster, tr_nsubs, tr_qreads;
  DevBuf<unsigned long long> tr_qsum;
  DevBuf<unsigned long long> tr_mat;
  d2g_stats tr_stats{};  // the last transition pass's compare
  // pinned staging of the per-unique centers and weights (and the slot
  // offsets, uploaded again only when another entry reused a_aloff)
  uint8_t* tr_pin = nullptr;
  size_t tr_pin_cap = 0;
  const uint64_t* tr_aoff_for = nullptr;
  uint64_t tr_aoff_n = 0, tr_aoff_stride = 0;
  // the device store of b_compare_parallel (d2g_store_reset / d2g_compare_store):
  // e_minmax per raw (sorted order) and the kept comparisons, appended
  bool store_valid = false;
  uint64_t store_n = 0;
  DevBuf<uint32_t> inv;  // caller -> sorted
  DevBuf<double> e_minmax, store_self;
  ShuffleComp* store = nullptr;
  size_t store_cap = 0;
  DevBuf<uint8_t> store_flag;
  DevBuf<uint32_t> store_bcount;
  uint8_t* store_pin = nullptr;  // count, self lambdas, the first kept comparisons
  size_t store_pin_cap = 0;
  // resident lock flags (the partition loop): compares with a null lock
  // argument read `lock`, which ctx_lock_add updates from a short index list
  bool lock_resident = false;
  DevBuf<uint32_t> lock_idx;
  uint32_t* lock_pin = nullptr;
  size_t lock_pin_cap = 0;
  hipEvent_t lock_ev = nullptr;  // the last index upload (staging reuse)
  bool lock_ev_pending = false;
  // align mode buffers
  DevBuf<uint8_t> a_pool, a_al0, a_al1;
  DevBuf<uint64_t> a_aloff;
  DevBuf<uint32_t> a_allen;
};

#define HIP_TRY(x)                    \
  do {                                \
    hipError_t e_ = (x);              \
    if (e_ != hipSuccess) return D2G_ERR_HIP; \
  } while (0)

extern "C" {

const char* d2g_status_string(int s) {
  switch (s) {
    case D2G_OK: return "ok";
    case D2G_ERR_ARG: return "invalid argument";
    case D2G_ERR_HIP: return "HIP runtime error";
    case D2G_ERR_NONACGT: return "non-ACGT base in the compare path";
    case D2G_ERR_TRACEBACK: return "invalid traceback pointer";
    case D2G_ERR_UNSUPPORTED: return "geometry or mode not supported by the kernels";
    case D2G_ERR_STATE: return "call order (uniques not loaded)";
    default: return "unknown status";
  }
}

int d2g_ctx_create(int device, d2g_ctx** out) {
  if (!out) return D2G_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return D2G_ERR_HIP;
  d2g_ctx* c = new d2g_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return D2G_ERR_HIP;
  }
  for (int i = 0; i < kEvents; ++i) (void)hipEventCreate(&c->ev[i]);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->cu_count = prop.multiProcessorCount;
    c->clock_khz = prop.clockRate;
    c->mem_clock_khz = prop.memoryClockRate;
    c->mem_bus_width = prop.memoryBusWidth;
  }
  if (c->cu_count <= 0) c->cu_count = 256;
  if (c->ctr.ensure(1) != hipSuccess) {
    d2g_ctx_destroy(c);
    return D2G_ERR_HIP;
  }
  const int na = noalloc_setup(device);
  if (na != D2G_OK) {
    d2g_ctx_destroy(c);
    return na;
  }
  *out = c;
  return D2G_OK;
}

int d2g_ctx_destroy(d2g_ctx* c) {
  if (!c) return D2G_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto* b : {&c->seq, &c->qual, &c->has255, &c->lock, &c->out_skipped, &c->fin_skipped,
                  &c->a_pool, &c->a_al0, &c->a_al1})
    b->release();
  c->fin_lambda.release();
  c->fin_hamming.release();
  for (auto* b : {&c->len, &c->reads, &c->orig, &c->kmer_row, &c->kmer_il, &c->kord_row, &c->kord_il,
                  &c->ksum, &c->centers, &c->out_hamming, &c->bin_count, &c->bin_start, &c->wave_start,
                  &c->bin_fill, &c->sub_pool, &c->sub_cnt, &c->screen_partials, &c->a_allen,
                  &c->hist})
    b->release();
  c->off.release();
  c->a_aloff.release();
  if (c->tr_pin) (void)hipHostFree(c->tr_pin);
  c->tr_pin = nullptr;
  c->err.release();
  c->out_lambda.release();
  c->gl.release();
  c->gl_unsorted.release();
  c->gl_off.release();
  c->dp_unsorted.release();
  c->dp_sorted.release();
  c->ctr.release();
  c->sched.release();
  c->sched_off.release();
  c->st_off.release();
  c->sched_info.release();
  c->st_sched.release();
  c->st_info.release();
  c->st_hist.release();
  c->wide_p.release();
  c->st_wide_p.release();
  for (auto* b : {&c->p_seq, &c->p_qual, &c->p_flag}) b->release();
  for (auto* b : {&c->p_off, &c->p_suboff}) b->release();
  for (auto* b : {&c->p_len, &c->p_subs, &c->p_reads, &c->p_creads, &c->p_nrec, &c->p_list}) b->release();
  c->p_loglam.release();
  c->p_lkf.release();
  c->b_out.release();
  c->last_ctr.release();
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  c->fetch_dev.release();
  if (c->fetch_pin) (void)hipHostFree(c->fetch_pin);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  for (auto* b : {&c->ll_w, &c->ll_part, &c->ll_out}) b->release();
  c->ll_logE.release();
  for (auto* b : {&c->p_lambda, &c->p_pval, &c->p_err}) b->release();
  c->inv.release();
  c->e_minmax.release();
  c->store_self.release();
  c->store_flag.release();
  c->store_bcount.release();
  if (c->store) (void)hipFree(c->store);
  if (c->store_pin) (void)hipHostFree(c->store_pin);
  c->lock_idx.release();
  if (c->lock_pin) (void)hipHostFree(c->lock_pin);
  if (c->lock_ev) (void)hipEventDestroy(c->lock_ev);
  for (int i = 0; i < kEvents; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return D2G_OK;
}

int d2g_ctx_set_timing(d2g_ctx* c, int enable) {
  if (!c) return D2G_ERR_ARG;
  c->timing = enable != 0;
  return D2G_OK;
}

int d2g_ctx_set_graphs(d2g_ctx* c, int enable) {
  if (!c) return D2G_ERR_ARG;
  c->graphs = enable != 0;
  return D2G_OK;
}

int d2g_ctx_set_autotune(d2g_ctx* c, int enable) {
  if (!c) return D2G_ERR_ARG;
  c->autotune = enable != 0;
  return D2G_OK;
}

int d2g_ctx_set_dp_tile(d2g_ctx* c, int tile) {
  if (!c || tile < -1 || tile > 16) return D2G_ERR_ARG;
  c->forced_tile = tile;
  return D2G_OK;
}

int d2g_ctx_set_screen_shape(d2g_ctx* c, int qblock, int split, int per_lane) {
  if (!c || (split != 1 && split != 2 && split != 4)) return D2G_ERR_ARG;
  if (qblock != 64 && qblock != 128 && qblock != 256) return D2G_ERR_ARG;
  if (per_lane != 1 && !(per_lane == 2 && qblock >= 128)) return D2G_ERR_ARG;
  c->screen_qblock = (uint32_t)qblock;
  c->screen_split = split;
  c->screen_per_lane = per_lane;
  return D2G_OK;
}

void* d2g_ctx_stream(d2g_ctx* c) { return c ? (void*)c->stream : nullptr; }

int d2g_device_info(d2g_ctx* c, int* cu, int* clk, int* mclk, int* bus) {
  if (!c) return D2G_ERR_ARG;
  if (cu) *cu = c->cu_count;
  if (clk) *clk = c->clock_khz;
  if (mclk) *mclk = c->mem_clock_khz;
  if (bus) *bus = c->mem_bus_width;
  return D2G_OK;
}

uint64_t d2g_vectorized_cells(uint32_t l1_in, uint32_t l2_in, int32_t band) {
  return vec_cells((int)l1_in, (int)l2_in, band);
}

int d2g_vectorized_supported(uint32_t l1, uint32_t l2, int32_t band) {
  if (l1 > kVectorizedMaxLen || l2 > kVectorizedMaxLen) return 0;
  const VecGeom g = vec_geom((int)l1, (int)l2, band);
  return (nw_groups(g) > 0 || wide_fits(g)) ? 1 : 0;
}

int d2g_load_uniques(d2g_ctx* c, uint32_t n, const uint8_t* seq_concat, const uint64_t* offsets,
                     const uint8_t* qual_concat, const uint32_t* reads, uint32_t k) {
  if (!c || !seq_concat || !offsets || !reads || n == 0) return D2G_ERR_ARG;
  if (k < 3 || k > 8) return D2G_ERR_ARG;  // dada.rs:259-263
  (void)hipSetDevice(c->device);
  c->loaded = false;
  std::vector<uint32_t> len(n);
  uint32_t maxq = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (offsets[i + 1] < offsets[i]) return D2G_ERR_ARG;
    uint64_t L = offsets[i + 1] - offsets[i];
    if (L <= k || L >= 9999) return D2G_ERR_ARG;  // dada.rs:265-275 (minlen > k, SEQLEN)
    len[i] = (uint32_t)L;
    for (uint64_t p = offsets[i]; p < offsets[i + 1]; ++p)
      if (seq_concat[p] < 1 || seq_concat[p] > 4) return D2G_ERR_NONACGT;  // dada.rs:279-294
    if (qual_concat)
      for (uint64_t p = offsets[i]; p < offsets[i + 1]; ++p) maxq = std::max<uint32_t>(maxq, qual_concat[p]);
  }
  // length-sorted packing (stable, so equal lengths keep the input order)
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return len[a] < len[b]; });
  std::vector<uint64_t> soff(n + 1, 0);
  std::vector<uint32_t> slen(n), sreads(n), inv(n);
  for (uint32_t s = 0; s < n; ++s) {
    slen[s] = len[order[s]];
    sreads[s] = reads[order[s]];
    inv[order[s]] = s;
    soff[s + 1] = soff[s] + ((slen[s] + 15u) & ~15u);  // 16-byte aligned rows
  }
  const uint64_t total = soff[n];
  // kernels address the padded pool with 32-bit byte offsets
  if (total >= ((uint64_t)1 << 32)) return D2G_ERR_UNSUPPORTED;
  std::vector<uint8_t> sseq(total, 0), squal(qual_concat ? total : 0, 0);
  for (uint32_t s = 0; s < n; ++s) {
    std::memcpy(&sseq[soff[s]], seq_concat + offsets[order[s]], slen[s]);
    if (qual_concat) std::memcpy(&squal[soff[s]], qual_concat + offsets[order[s]], slen[s]);
  }
  c->h_reads.assign(reads, reads + n);
  UniqueDev& u = c->u;
  u.n = n;
  u.npad = (n + 63) & ~63u;
  u.k = k;
  u.kwords = (1u << (2 * k)) / 4;
  u.minlen = slen[0];
  u.maxlen = slen[n - 1];
  u.owords = (u.maxlen - k + 1 + 1) / 2;
  HIP_TRY(c->seq.ensure(total));
  HIP_TRY(c->off.ensure(n + 1));
  HIP_TRY(c->len.ensure(n));
  HIP_TRY(c->reads.ensure(n));
  HIP_TRY(c->orig.ensure(n));
  HIP_TRY(c->inv.ensure(n));
  HIP_TRY(c->kmer_row.ensure((size_t)n * u.kwords));
  HIP_TRY(c->kmer_il.ensure((size_t)u.npad * u.kwords));
  HIP_TRY(c->kord_row.ensure((size_t)n * u.owords));
  HIP_TRY(c->kord_il.ensure((size_t)u.npad * u.owords));
  HIP_TRY(c->ksum.ensure(n));
  HIP_TRY(c->has255.ensure(n));
  if (qual_concat) HIP_TRY(c->qual.ensure(total));
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(c->seq.p, sseq.data(), total, hipMemcpyHostToDevice, s));
  if (qual_concat) HIP_TRY(hipMemcpyAsync(c->qual.p, squal.data(), total, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->off.p, soff.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->len.p, slen.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->reads.p, sreads.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->orig.p, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->inv.p, inv.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(launch_zero(c->kord_row.p, (size_t)n * u.owords, s));
  HIP_TRY(launch_zero(c->kord_il.p, (size_t)u.npad * u.owords, s));
  HIP_TRY(launch_zero(c->kmer_il.p, (size_t)u.npad * u.kwords, s));
  u.seq = c->seq.p;
  u.qual = qual_concat ? c->qual.p : nullptr;
  u.off = c->off.p;
  u.len = c->len.p;
  u.reads = c->reads.p;
  u.orig = c->orig.p;
  u.kmer_row = c->kmer_row.p;
  u.kmer_il = c->kmer_il.p;
  u.kord_row = c->kord_row.p;
  u.kord_il = c->kord_il.p;
  u.ksum = c->ksum.p;
  u.has255 = c->has255.p;
  HIP_TRY(launch_prep(u, s));
  HIP_TRY(hipStreamSynchronize(s));
  c->h_len = std::move(slen);
  c->h_seq = std::move(sseq);
  c->h_qual = std::move(squal);
  c->h_soff = std::move(soff);
  c->h_orig = std::move(order);
  c->h_inv = std::move(inv);
  c->max_qual = maxq;
  c->has_qual = qual_concat != nullptr;
  c->loaded = true;
  c->cmp_valid = false;
  c->store_valid = false;  // the store belongs to the previous uniques
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);  // unique buffers changed
  c->graph_exec = nullptr;
  c->graph_key.clear();
  return D2G_OK;
}

#define HIP_TRY_INT(x)        \
  do {                        \
    int r_ = (x);             \
    if (r_ != D2G_OK) return r_; \
  } while (0)

// LDS size, per-bin info (cells; steady rows) and, for the vectorized
// kernel, the row schedule of a bin space, cached across calls with the same
// space and scores; also sizes the persistent grid and its pointer-history
// scratch.
static void swap_schedule(d2g_ctx* c) {
  std::swap(c->lds_key, c->st_key);
  std::swap(c->sched_gap, c->st_gap);
  std::swap(c->sched_end_gap, c->st_end_gap);
  std::swap(c->lds_bytes, c->st_lds_bytes);
  std::swap(c->sched_rows, c->st_rows);
  std::swap(c->sched_present, c->st_present);
  std::swap(c->sched_off, c->st_off);
  std::swap(c->sched, c->st_sched);
  std::swap(c->sched_info, c->st_info);
  std::swap(c->nw_grid, c->st_grid);
  std::swap(c->hist_stride, c->st_hist_stride);
  std::swap(c->hist, c->st_hist);
  std::swap(c->wide, c->st_wide);
  std::swap(c->wide_grid, c->st_wide_grid);
  std::swap(c->wide_p, c->st_wide_p);
}

static int prepare_schedule(d2g_ctx* c, const BinSpace& bs, int end_gap, int gap,
                            const std::vector<uint8_t>& present) {
  auto hit = [&]() {
    return bs.lmin == c->lds_key.lmin && bs.range == c->lds_key.range &&
           bs.band == c->lds_key.band && bs.kind == c->lds_key.kind && bs.homo == c->lds_key.homo &&
           bs.tile == c->lds_key.tile &&
           end_gap == c->sched_end_gap && gap == c->sched_gap && c->sched_info.p &&
           present == c->sched_present;
  };
  if (hit()) return D2G_OK;
  // keep one schedule per DP kernel: the other kernel's goes to the stash
  if (bs.kind != c->lds_key.kind) {
    swap_schedule(c);
    if (hit()) return D2G_OK;
  }
  const bool vec = bs.kind == kDpVectorized;
  const uint8_t* pr = present.data();
  c->lds_bytes = vec ? nw_vec_lds_for(bs, pr) : nw_band_lds_for(bs, pr);
  const int lmax = bs.lmin + bs.range - 1;
  if (c->sched_info.ensure((size_t)bs.nbins) != hipSuccess) return D2G_ERR_HIP;
  std::vector<uint64_t> off;
  if (vec) {
    // row schedules only for the length pairs the call's sequences form
    c->sched_rows = nw_sched_rows(lmax);
    off.assign((size_t)bs.nbins, kSchedAbsent);
    uint64_t recs = 0;
    for (int b = 0; b < bs.nbins; ++b) {
      int l1, l2;
      bin_lengths(bs, (uint32_t)b, &l1, &l2);
      if (!bin_present(bs, pr, l1, l2)) continue;
      const VecGeom g = vec_geom(l1, l2, bs.band);
      off[(size_t)b] = recs;
      if (nw_groups(g) > 0) recs += (uint64_t)g.nrow;
    }
    if (c->sched.ensure(std::max<uint64_t>(recs, 1)) != hipSuccess) return D2G_ERR_HIP;
    if (c->sched_off.ensure((size_t)bs.nbins) != hipSuccess) return D2G_ERR_HIP;
    if (hipMemcpyAsync(c->sched_off.p, off.data(), off.size() * sizeof(uint64_t),
                       hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return D2G_ERR_HIP;
  }
  // persistent DP workgroups: 8 one-wave groups per SIMD where LDS allows,
  // and at most 1 GiB of pointer-history scratch
  int per_cu = 32;
  if (c->lds_bytes > 0) per_cu = std::min<int>(per_cu, (int)((160 * 1024) / c->lds_bytes));
  per_cu = std::max(per_cu, 1);
  c->nw_grid = (uint32_t)c->cu_count * (uint32_t)per_cu;
  // history words per slot: K per 16 anti-diagonals for the row-band kernel
  int kmax = 1;
  if (!vec)
    for (int b = 0; b < bs.nbins; ++b) {
      int l1, l2;
      bin_lengths(bs, (uint32_t)b, &l1, &l2);
      kmax = std::max(kmax, band_geom(l1, l2, bs.band, bs.tile).k);
    }
  c->hist_stride = nw_hist_p_words(2 * lmax + 1) * (size_t)kmax;
  while (c->nw_grid > (uint32_t)c->cu_count && (size_t)c->nw_grid * c->hist_stride * 4 > ((size_t)1 << 30))
    c->nw_grid /= 2;
  if (c->hist.ensure((size_t)c->nw_grid * c->hist_stride) != hipSuccess) return D2G_ERR_HIP;
  // the one-pair kernel: a pointer slot per wavefront, at most 1 GiB
  c->wide = nw_wide_need(bs, pr);
  c->wide_grid = 0;
  if (c->wide.max_ncol > 0) {
    c->wide_grid = (uint32_t)std::max<size_t>(
        1, std::min<size_t>((size_t)c->cu_count * 2, ((size_t)1 << 30) / c->wide.pslot));
    if (c->wide_p.ensure((size_t)c->wide_grid * c->wide.pslot) != hipSuccess) return D2G_ERR_HIP;
  }
  const hipError_t e = vec ? launch_sched(bs, c->sched_off.p, end_gap, gap, c->sched.p, c->sched_info.p, c->stream)
                           : launch_band_info(bs, c->sched_info.p, c->stream);
  if (e != hipSuccess) return D2G_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return D2G_ERR_HIP;  // off is pageable
  c->lds_key = bs;
  c->sched_present = present;
  c->sched_end_gap = end_gap;
  c->sched_gap = gap;
  return D2G_OK;
}

// The DP kernel of the compare for these parameters (raw_align_dp,
// nwalign.rs:1335-1414): the vectorized aligner when vectorized, no
// homopolymer penalty and both lengths <= 3500; the homopolymer aligner when
// homo_gap_p differs from gap_p and is <= 0; the plain ends-free aligner
// otherwise.  Returns -1 when one call would need both kernels (vectorized
// parameters with some uniques above 3500 nt and some not).
static int dp_kind(const d2g_align_params* p, uint32_t minlen, uint32_t maxlen, int* homo) {
  const bool use_homo = p->homo_gap_p != p->gap_p && p->homo_gap_p <= 0;
  *homo = use_homo ? 1 : 0;
  if (p->vectorized && !use_homo) {
    if (maxlen <= kVectorizedMaxLen) return kDpVectorized;
    if (minlen > kVectorizedMaxLen) return kDpRowBand;
    return -1;
  }
  return kDpRowBand;
}

// Workload class of the autotuner: uniques and centers rounded to powers of
// two, k, and the DP bin space's band and kind.
static uint64_t tune_class(uint32_t n, uint32_t nc, uint32_t k, const BinSpace& bs) {
  auto lg = [](uint32_t v) { uint64_t b = 0; while ((1u << b) < v && b < 31) ++b; return b; };
  return lg(n) | lg(nc) << 6 | (uint64_t)k << 12 | (uint64_t)(uint32_t)bs.band << 16 |
         (uint64_t)bs.kind << 40 | (uint64_t)bs.homo << 44;
}

static float ev_ms(d2g_ctx* c, int a, int b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0.f;
  return ms;
}

// Copies the last compare's results to the host in the caller's unique
// order (unpermuted on the device); enqueued only.
static hipError_t enqueue_fetch(d2g_ctx* c, uint32_t nc, double* out_lambda,
                                uint32_t* out_hamming, uint8_t* out_skipped) {
  const size_t np = (size_t)nc * c->u.n;
  hipStream_t s = c->stream;
  hipError_t e;
  if ((e = c->fin_lambda.ensure(np)) != hipSuccess) return e;
  if ((e = c->fin_hamming.ensure(np)) != hipSuccess) return e;
  if ((e = c->fin_skipped.ensure(np)) != hipSuccess) return e;
  if ((e = launch_unpermute(nc, c->u.n, c->u.orig, c->out_lambda.p, c->out_hamming.p,
                            c->out_skipped.p, c->fin_lambda.p, c->fin_hamming.p,
                            c->fin_skipped.p, s)) != hipSuccess)
    return e;
  if (out_lambda &&
      (e = hipMemcpyAsync(out_lambda, c->fin_lambda.p, np * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if (out_hamming &&
      (e = hipMemcpyAsync(out_hamming, c->fin_hamming.p, np * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if (out_skipped &&
      (e = hipMemcpyAsync(out_skipped, c->fin_skipped.p, np, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  return hipSuccess;
}

// A result set of at most kPackedFetchMax bytes returns as one packed block:
// the unpermute writes lambda, hamming, skipped and the counters into
// fetch_dev, one copy lands it in pinned memory, and the host spreads it
// into the caller's arrays after the synchronisation (one short copy instead
// of three staged pageable ones and the counters' own).
constexpr size_t kPackedFetchMax = (size_t)4 << 20;
struct PackedLayout {
  size_t ham, skp, ctr, total;
};
static PackedLayout packed_layout(size_t np) {
  PackedLayout L;
  L.ham = np * sizeof(double);
  L.skp = L.ham + np * sizeof(uint32_t);
  L.ctr = (L.skp + np + 15) & ~(size_t)15;
  L.total = L.ctr + sizeof(Counters);
  return L;
}

static hipError_t enqueue_fetch_packed(d2g_ctx* c, uint32_t nc) {
  const size_t np = (size_t)nc * c->u.n;
  const PackedLayout L = packed_layout(np);
  hipError_t e;
  if ((e = c->fetch_dev.ensure(L.total)) != hipSuccess) return e;
  if (c->fetch_pin_cap < L.total) {
    if (c->fetch_pin) (void)hipHostFree(c->fetch_pin);  // no copy into it is pending
    c->fetch_pin = nullptr;
    c->fetch_pin_cap = 0;
    if ((e = hipHostMalloc((void**)&c->fetch_pin, L.total, hipHostMallocDefault)) != hipSuccess)
      return e;
    c->fetch_pin_cap = L.total;
  }
  uint8_t* d = c->fetch_dev.p;
  if ((e = launch_unpermute(nc, c->u.n, c->u.orig, c->out_lambda.p, c->out_hamming.p,
                            c->out_skipped.p, reinterpret_cast<double*>(d),
                            reinterpret_cast<uint32_t*>(d + L.ham), d + L.skp, c->stream,
                            reinterpret_cast<const uint32_t*>(c->ctr.p),
                            reinterpret_cast<uint32_t*>(d + L.ctr),
                            (uint32_t)(sizeof(Counters) / 4))) != hipSuccess)
    return e;
  return hipMemcpyAsync(c->fetch_pin, d, L.total, hipMemcpyDeviceToHost, c->stream);
}

// After the synchronisation: the packed block into the caller's arrays and
// the counters into h_ctr.
static void finish_fetch_packed(d2g_ctx* c, uint32_t nc, double* out_lambda,
                                uint32_t* out_hamming, uint8_t* out_skipped) {
  const size_t np = (size_t)nc * c->u.n;
  const PackedLayout L = packed_layout(np);
  const uint8_t* h = c->fetch_pin;
  if (out_lambda) std::memcpy(out_lambda, h, np * sizeof(double));
  if (out_hamming) std::memcpy(out_hamming, h + L.ham, np * sizeof(uint32_t));
  if (out_skipped) std::memcpy(out_skipped, h + L.skp, np);
  std::memcpy(c->h_ctr, h + L.ctr, sizeof(Counters));
}

// The transition pass's variant of the compare (transition_counts,
// error.rs:194-236): the screen keeps each unique's pair with its own center
// (want, sorted order), the DP writes alignment strings per unique, and the
// transition kernel replaces lambda.  Device pointers.
struct TransMode {
  const uint32_t* want;
  const uint32_t* weight;
  uint8_t* al0;
  uint8_t* al1;
  const uint64_t* al_off;
  uint32_t* al_len;
  int use_quals;
  uint32_t ncol;
  int clamp;
  unsigned long long* mat;
  bool count = true;  // the transition kernel (off: alignments only)
  // compute_aux's per-raw sums (optional): order / cluster per sorted unique
  bool aux = false;
  const uint32_t* order = nullptr;
  uint32_t n_order = 0;
  const uint32_t* cluster = nullptr;
  uint32_t maxlen = 0;
  uint32_t* nsubs = nullptr;
  uint32_t* qual_reads = nullptr;
  unsigned long long* qual_sum = nullptr;
};

// One pass of the compare: the DP kernel and its bin space, and which
// (center, unique) pairs the pass takes (see ScreenArgs).
struct Route {
  int kind, homo;
  uint32_t lmin, lmax;  // lengths the bin space covers
  uint32_t long_len, sr_lo, sr_hi, lr_lo, lr_hi;
  bool graphs;          // replay from a captured graph (single-pass calls)
};

// One pass of the compare pipeline; with fetch, the results are copied to
// the host before the pass's single synchronisation.
static int compare_pass(d2g_ctx* c, const uint32_t* centers, uint32_t nc, const double* err,
                        uint32_t ncol, const d2g_align_params* p, int use_quals, int greedy,
                        const uint8_t* lock, d2g_stats* stats, bool fetch, double* out_lambda,
                        uint32_t* out_hamming, uint8_t* out_skipped, const TransMode* tm,
                        const Route& rt) {
  UniqueDev& u = c->u;
  const uint32_t n = u.n;
  const size_t npairs = (size_t)nc * n;
  BinSpace bs;
  bs.lmin = (int)rt.lmin;
  bs.range = (int)(rt.lmax - rt.lmin + 1);
  bs.band = p->band;
  bs.kind = rt.kind;
  bs.homo = rt.homo;
  bs.tile = c->forced_tile >= 0 ? c->forced_tile : 0;
  if ((size_t)bs.range * bs.range > (1u << 22)) return D2G_ERR_UNSUPPORTED;
  if (c->autotune && !tm && c->forced_tile < 0) {
    bs.nbins = bs.range * bs.range;
    const auto it = c->tuned_tile.find(tune_class(n, nc, u.k, bs));
    if (it != c->tuned_tile.end()) bs.tile = it->second;
  }
  bs.nbins = bs.range * bs.range;
  const bool vec = bs.kind == kDpVectorized;
  // the lengths of the uniques in the bin space
  std::vector<uint8_t> present((size_t)bs.range, 0);
  for (uint32_t L : c->h_len)
    if (L >= rt.lmin && L <= rt.lmax) present[L - rt.lmin] = 1;
  HIP_TRY_INT(prepare_schedule(c, bs, 0, p->gap_p, present));
  HIP_TRY(c->centers.ensure(nc));
  if (c->err.cap < (size_t)16 * ncol) c->h_err.clear();  // reallocation loses the upload
  HIP_TRY(c->err.ensure((size_t)16 * ncol));
  // results and task lists: one per pair, or one per unique with the
  // transition filter (which writes no results)
  const size_t ntask = tm ? n : npairs;
  HIP_TRY(c->out_lambda.ensure(tm ? 1 : npairs));
  HIP_TRY(c->out_hamming.ensure(tm ? 1 : npairs));
  HIP_TRY(c->out_skipped.ensure(tm ? 1 : npairs));
  // per screen block: a region of screen_region(qblock) task slots in each
  // list, sized for every screen shape (the autotuner tries several); the
  // transition pass's pairing: kPairBlock slots per block
  size_t max_part = 0, max_slots = 0;
  if (tm) {
    max_part = pair_blocks(n);
    max_slots = max_part * kPairBlock;
  }
  for (uint32_t qb : {64u, 128u, 256u}) {
    if (tm) break;
    max_part = std::max<size_t>(max_part, screen_blocks(n, nc, qb));
    max_slots = std::max<size_t>(max_slots, (size_t)screen_blocks(n, nc, qb) * screen_region(qb));
  }
  // one center: the single-center screen's regions
  const bool single = !tm && screen1_fits(nc, p->use_kmers, u.kwords);
  if (single) {
    max_part = std::max<size_t>(max_part, screen1_regions(n));
    max_slots = std::max<size_t>(max_slots, (size_t)screen1_regions(n) * kS1Uniques);
  }
  HIP_TRY(c->gl.ensure(ntask));
  HIP_TRY(c->gl_unsorted.ensure(max_slots));
  HIP_TRY(c->gl_off.ensure(max_part));
  HIP_TRY(c->dp_unsorted.ensure(max_slots));
  HIP_TRY(c->dp_sorted.ensure(ntask));
  HIP_TRY(c->sub_cnt.ensure(ntask));
  HIP_TRY(c->bin_count.ensure(bs.nbins + 1));
  HIP_TRY(c->bin_start.ensure(bs.nbins + 1));
  HIP_TRY(c->wave_start.ensure(bs.nbins + 1));
  HIP_TRY(c->bin_fill.ensure(bs.nbins + 1));
  HIP_TRY(c->screen_partials.ensure((size_t)4 * max_part));
  const uint32_t sub_stride = (uint32_t)(bs.lmin + bs.range);  // longest length + 1
  if (c->sub_capacity < (size_t)1 << 20) c->sub_capacity = (size_t)1 << 20;
  HIP_TRY(c->sub_pool.ensure(c->sub_capacity));
  hipStream_t s = c->stream;
  // uploads through pinned staging (the previous call ended with a stream
  // synchronisation, so the staging area is free)
  const size_t pin_need = (size_t)nc * 4 + (lock ? n : 0) + 16;
  if (c->h_pin_cap < pin_need) {
    if (c->h_pin) (void)hipHostFree(c->h_pin);  // the graphs never read the staging area
    c->h_pin = nullptr;
    c->h_pin_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&c->h_pin, pin_need, hipHostMallocDefault));
    c->h_pin_cap = pin_need;
  }
  uint32_t* hc = reinterpret_cast<uint32_t*>(c->h_pin);
  for (uint32_t i = 0; i < nc; ++i) hc[i] = c->h_inv[centers[i]];
  if (c->h_centers_buf != c->centers.p || c->h_centers.size() != nc ||
      std::memcmp(c->h_centers.data(), hc, (size_t)nc * sizeof(uint32_t)) != 0) {
    HIP_TRY(hipMemcpyAsync(c->centers.p, hc, nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    c->h_centers.assign(hc, hc + nc);
    c->h_centers_buf = c->centers.p;
  }
  const uint8_t* dlock = nullptr;
  if (lock) {
    uint8_t* sl = c->h_pin + (size_t)nc * 4;
    for (uint32_t i = 0; i < n; ++i) sl[i] = lock[c->h_orig[i]];
    HIP_TRY(c->lock.ensure(n));
    HIP_TRY(hipMemcpyAsync(c->lock.p, sl, n, hipMemcpyHostToDevice, s));
    dlock = c->lock.p;
  } else if (c->lock_resident) {
    dlock = c->lock.p;  // kept current by ctx_lock_add
  }
  const size_t ne = (size_t)16 * ncol;
  if (c->h_err.size() != ne || std::memcmp(c->h_err.data(), err, ne * sizeof(double)) != 0) {
    c->h_err.assign(err, err + ne);
    HIP_TRY(hipMemcpyAsync(c->err.p, c->h_err.data(), ne * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // h_err is pageable: finish before it can change
  }
  // shroud thresholds of the screen (cached per cutoff and longest unique)
  if (p->use_kmers) {
    uint64_t cut;
    std::memcpy(&cut, &p->kdist_cutoff, 8);
    const uint32_t nmax = u.maxlen - u.k + 1;
    if (!c->dstar_valid || c->dstar_cut != cut || c->dstar_n != nmax) {
      std::vector<uint32_t> t(nmax + 1);
      screen_thresholds(p->kdist_cutoff, nmax, t.data());
      HIP_TRY(c->dstar.ensure(nmax + 1));
      HIP_TRY(hipMemcpyAsync(c->dstar.p, t.data(), (nmax + 1) * sizeof(uint32_t),
                             hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));  // t is pageable
      c->dstar_cut = cut;
      c->dstar_n = nmax;
      c->dstar_valid = true;
    }
  }

  // The launch sequence of one compare (screen -> plan -> scatter -> DP ->
  // lambda).  Timed runs bracket the kernels with events; otherwise the
  // sequence is captured once per shape into a HIP graph and replayed.
  auto enqueue = [&](bool timed) -> int {
    const uint32_t n_part = tm ? pair_blocks(n)
                            : single ? screen1_regions(n) : screen_blocks(n, nc, c->screen_qblock);
    const uint32_t region = tm ? kPairBlock : single ? kS1Uniques : screen_region(c->screen_qblock);
    if (timed) HIP_TRY(hipEventRecord(c->ev[0], s));
    HIP_TRY(launch_zero2(reinterpret_cast<uint32_t*>(c->ctr.p), sizeof(Counters) / 4,
                         c->bin_count.p, bs.nbins + 1, s));
    ScreenArgs sa;
    sa.u = &u;
    sa.centers = c->centers.p;
    sa.n_centers = nc;
    sa.lock = dlock;
    sa.greedy = greedy;
    sa.use_kmers = p->use_kmers;
    sa.gapless = p->gapless;
    sa.kdist_cutoff = p->kdist_cutoff;
    sa.band = p->band;
    sa.bins = bs;
    sa.out_lambda = c->out_lambda.p;
    sa.out_hamming = c->out_hamming.p;
    sa.out_skipped = c->out_skipped.p;
    sa.gapless_tasks = c->gl_unsorted.p;
    sa.dp_unsorted = c->dp_unsorted.p;
    sa.bin_count = c->bin_count.p;
    sa.partials = c->screen_partials.p;
    sa.ctr = c->ctr.p;
    sa.qblock = c->screen_qblock;
    sa.split = c->screen_split;
    sa.per_lane = c->screen_per_lane;
    sa.dstar = c->dstar.p;
    sa.long_len = rt.long_len;
    sa.sr_lo = rt.sr_lo;
    sa.sr_hi = rt.sr_hi;
    sa.lr_lo = rt.lr_lo;
    sa.lr_hi = rt.lr_hi;
    sa.single = single;
    HIP_TRY(tm ? launch_pair_screen(sa, tm->want, s) : launch_screen(sa, s));
    if (timed) HIP_TRY(hipEventRecord(c->ev[1], s));
    if (single) {
      // one center: the plan kernel compacts the length-ordered regions
      const PlanCopy cp{c->dp_unsorted.p, c->dp_sorted.p, c->gl_unsorted.p, c->gl.p, region};
      HIP_TRY(launch_plan(bs, c->bin_count.p, c->bin_start.p, c->wave_start.p, c->bin_fill.p,
                          c->ctr.p, c->sched_info.p, c->screen_partials.p, n_part, c->gl_off.p, s,
                          &cp));
    } else {
      HIP_TRY(launch_plan(bs, c->bin_count.p, c->bin_start.p, c->wave_start.p, c->bin_fill.p,
                          c->ctr.p, c->sched_info.p, c->screen_partials.p, n_part, c->gl_off.p, s));
      HIP_TRY(launch_scatter(c->dp_unsorted.p, c->dp_sorted.p, c->bin_start.p, c->bin_fill.p,
                             c->screen_partials.p, n_part, region, 0, c->gl_unsorted.p, c->gl.p,
                             c->gl_off.p, s));
    }
    if (timed) HIP_TRY(hipEventRecord(c->ev[2], s));
    NwArgs na;
    std::memset(&na, 0, sizeof(na));
    na.bins = bs;
    na.bin_start = c->bin_start.p;
    na.wave_start = c->wave_start.p;
    na.tasks = c->dp_sorted.p;
    na.pool = u.seq;
    // the vectorized aligner takes i16 scores (nwalign.rs:1388-1394), the
    // scalar ones i32
    na.match = vec ? (int)(int16_t)p->match_score : p->match_score;
    na.mismatch = vec ? (int)(int16_t)p->mismatch : p->mismatch;
    na.gap_p = vec ? (int)(int16_t)p->gap_p : p->gap_p;
    na.homo_gap_p = p->homo_gap_p;
    na.end_gap_p = 0;  // raw_align_dp always aligns ends-free (nwalign.rs:1393)
    na.ctr = c->ctr.p;
    na.emit_alignment = 0;
    na.sub_pool = c->sub_pool.p;
    na.sub_capacity = c->sub_capacity;
    na.sub_stride = sub_stride;
    na.sub_cnt = c->sub_cnt.p;
    // no more wavefronts than pairs can exist: small calls launch small grids
    na.grid = (uint32_t)std::min<size_t>(c->nw_grid, std::max<size_t>(npairs, 1));
    na.hist = c->hist.p;
    na.hist_stride = c->hist_stride;
    na.lds_bytes = c->lds_bytes;
    na.sched = c->sched.p;
    na.sched_off = c->sched_off.p;
    na.info = c->sched_info.p;
    if (tm) {
      na.emit_alignment = 1;
      na.al0 = tm->al0;
      na.al1 = tm->al1;
      na.al_off = tm->al_off;
      na.al_len = tm->al_len;
    }
    HIP_TRY(vec ? launch_nw_vec(na, s) : launch_nw_band(na, s));
    if (vec) HIP_TRY(launch_nw_wide(na, c->wide, c->wide_p.p, c->wide_grid, s));
    if (timed) HIP_TRY(hipEventRecord(c->ev[3], s));
    if (tm) {
      TransArgs ta;
      ta.u = &u;
      ta.gapless_tasks = c->gl.p;
      ta.dp_tasks = c->dp_sorted.p;
      ta.ctr = c->ctr.p;
      ta.al0 = tm->al0;
      ta.al1 = tm->al1;
      ta.al_off = tm->al_off;
      ta.al_len = tm->al_len;
      ta.weight = tm->weight;
      ta.use_quals = tm->use_quals;
      ta.ncol = tm->ncol;
      ta.clamp = tm->clamp;
      ta.mat = tm->mat;
      // eight waves per block, one pair per wave at a time; about 16 waves
      // per SIMD keep the loads of the chunks in flight
      ta.grid = (uint32_t)std::min<size_t>((size_t)c->cu_count * 8, (n + 7) / 8);
      if (tm->count) HIP_TRY(launch_transition(ta, s));
      if (tm->aux) {
        AuxArgs xa;
        xa.u = &u;
        xa.al0 = tm->al0;
        xa.al1 = tm->al1;
        xa.al_off = tm->al_off;
        xa.al_len = tm->al_len;
        xa.order = tm->order;
        xa.n_order = tm->n_order;
        xa.cluster = tm->cluster;
        xa.want = tm->want;
        xa.weight = tm->weight;
        xa.gapless = p->band == 0;
        xa.use_quals = tm->use_quals;
        xa.maxlen = tm->maxlen;
        xa.nsubs = tm->nsubs;
        xa.qual_reads = tm->qual_reads;
        xa.qual_sum = tm->qual_sum;
        xa.grid = (uint32_t)c->cu_count * 8;
        HIP_TRY(launch_aux(xa, s));
      }
      if (timed) HIP_TRY(hipEventRecord(c->ev[4], s));
      return D2G_OK;
    }
    LambdaArgs la;
    la.u = &u;
    la.gapless_tasks = c->gl.p;
    la.dp_tasks = c->dp_sorted.p;
    la.sub_pool = c->sub_pool.p;
    la.sub_stride = sub_stride;
    la.sub_cnt = c->sub_cnt.p;
    la.err = c->err.p;
    la.ncol = ncol;
    la.use_quals = use_quals;
    la.ctr = c->ctr.p;
    la.out_lambda = c->out_lambda.p;
    la.out_hamming = c->out_hamming.p;
    la.out_skipped = c->out_skipped.p;
    la.grid = (uint32_t)std::min<size_t>((size_t)c->cu_count * 3, (npairs + 63) / 64);  // 64 pairs each
    HIP_TRY(launch_lambda(la, s));
    if (timed) HIP_TRY(hipEventRecord(c->ev[4], s));
    return D2G_OK;
  };

  // Autotuning: the first compare of a workload class times each screen
  // shape on this call's inputs (results do not depend on the shape) and
  // keeps the fastest for the class.
  if (c->autotune && !tm) {
    const uint64_t cls = tune_class(n, nc, u.k, bs);
    auto it = c->tuned.find(cls);
    if (it == c->tuned.end()) {
      static const int kShapes[][3] = {{64, 4, 1}, {64, 2, 1}, {128, 2, 1}, {256, 2, 1},
                                       {128, 4, 2}, {256, 4, 2}};
      float best = 0.f;
      int best_i = 0;
      for (int i = 0; i < (int)(sizeof(kShapes) / sizeof(kShapes[0])); ++i) {
        c->screen_qblock = (uint32_t)kShapes[i][0];
        c->screen_split = kShapes[i][1];
        c->screen_per_lane = kShapes[i][2];
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the second run is timed
          HIP_TRY_INT(enqueue(true));
          HIP_TRY(hipEventSynchronize(c->ev[4]));
          ms = ev_ms(c, 0, 1);
        }
        if (i == 0 || ms < best) {
          best = ms;
          best_i = i;
        }
      }
      it = c->tuned.emplace(cls, std::array<int, 3>{kShapes[best_i][0], kShapes[best_i][1],
                                                     kShapes[best_i][2]}).first;
    }
    c->screen_qblock = (uint32_t)it->second[0];
    c->screen_split = it->second[1];
    c->screen_per_lane = it->second[2];
    // then the DP tile: pairs per wavefront (vectorized: as many as fit, at
    // most 2, 1) or cells per lane (row band: 1, 2, 4), the DP kernel timed
    // on this call's pairs
    if (c->forced_tile < 0 && c->tuned_tile.find(cls) == c->tuned_tile.end()) {
      static const int kVecTiles[] = {0, 2, 1};
      static const int kBandTiles[] = {1, 2, 4};
      const int* tiles = vec ? kVecTiles : kBandTiles;
      float best = 0.f;
      int best_tile = tiles[0];
      for (int i = 0; i < 3; ++i) {
        bs.tile = tiles[i];
        HIP_TRY_INT(prepare_schedule(c, bs, 0, p->gap_p, present));
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the second run is timed
          HIP_TRY_INT(enqueue(true));
          HIP_TRY(hipEventSynchronize(c->ev[4]));
          ms = ev_ms(c, 2, 3);
        }
        if (i == 0 || ms < best) {
          best = ms;
          best_tile = tiles[i];
        }
      }
      c->tuned_tile.emplace(cls, best_tile);
      bs.tile = best_tile;
      HIP_TRY_INT(prepare_schedule(c, bs, 0, p->gap_p, present));
    }
  }

  for (int attempt = 0; attempt < 6; ++attempt) {
    const bool graph = c->graphs && !c->timing && !tm && rt.graphs;
    if (graph) {
      // everything a kernel argument depends on
      std::vector<uint64_t> key = {
          nc, n, (uint64_t)(uintptr_t)u.seq, (uint64_t)(uintptr_t)dlock, (uint64_t)greedy,
          (uint64_t)use_quals, ncol, (uint64_t)bs.lmin, (uint64_t)bs.range, (uint64_t)bs.band,
          (uint64_t)(uint32_t)bs.tile,
          (uint64_t)bs.kind, (uint64_t)bs.homo, (uint64_t)p->use_kmers, (uint64_t)p->gapless,
          (uint64_t)p->match_score, (uint64_t)(uint32_t)p->mismatch, (uint64_t)(uint32_t)p->gap_p,
          (uint64_t)(uint32_t)p->homo_gap_p, (uint64_t)p->kmer_size, sub_stride, c->sub_capacity,
          rt.long_len, rt.sr_lo, rt.sr_hi, rt.lr_lo, rt.lr_hi,
          c->nw_grid, c->lds_bytes, c->hist_stride, (uint64_t)c->sched_rows,
          (uint64_t)c->screen_split, (uint64_t)c->screen_qblock, (uint64_t)c->screen_per_lane,
          (uint64_t)(uintptr_t)c->dstar.p};
      uint64_t cut;
      std::memcpy(&cut, &p->kdist_cutoff, 8);
      key.push_back(cut);
      for (const void* q : {(const void*)c->centers.p, (const void*)c->err.p,
                            (const void*)c->out_lambda.p, (const void*)c->out_hamming.p,
                            (const void*)c->out_skipped.p, (const void*)c->gl.p,
                            (const void*)c->gl_unsorted.p, (const void*)c->gl_off.p,
                            (const void*)c->dp_unsorted.p, (const void*)c->dp_sorted.p,
                            (const void*)c->sub_cnt.p, (const void*)c->sub_pool.p,
                            (const void*)c->bin_count.p, (const void*)c->bin_start.p,
                            (const void*)c->wave_start.p, (const void*)c->bin_fill.p,
                            (const void*)c->screen_partials.p, (const void*)c->hist.p,
                            (const void*)c->sched.p, (const void*)c->sched_info.p,
                            (const void*)c->sched_off.p, (const void*)c->wide_p.p,
                            (const void*)c->ctr.p})
        key.push_back((uint64_t)(uintptr_t)q);
      if (!c->graph_exec || key != c->graph_key) {
        if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
        c->graph_key.clear();
        hipGraph_t gr = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(false);
        const hipError_t ec = hipStreamEndCapture(s, &gr);
        if (rc != D2G_OK) return rc;
        HIP_TRY(ec);
        const hipError_t ei = hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIP_TRY(ei);
        c->graph_key = key;
      }
      HIP_TRY(hipGraphLaunch(c->graph_exec, s));
    } else {
      HIP_TRY_INT(enqueue(c->timing));
    }
    if (!c->h_ctr) HIP_TRY(hipHostMalloc((void**)&c->h_ctr, sizeof(Counters), hipHostMallocDefault));
    const bool packed = fetch && packed_layout(npairs).total <= kPackedFetchMax;
    if (packed) {
      HIP_TRY(enqueue_fetch_packed(c, nc));
    } else {
      if (fetch) HIP_TRY(enqueue_fetch(c, nc, out_lambda, out_hamming, out_skipped));
      HIP_TRY(hipMemcpyAsync(c->h_ctr, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (packed) finish_fetch_packed(c, nc, out_lambda, out_hamming, out_skipped);
    const Counters h = *c->h_ctr;
    if (h.err_flags & kErrSubOverflow) {
      // every DP pair owns sub_stride records
      c->sub_capacity = std::max(c->sub_capacity * 2, (size_t)h.n_dp * sub_stride);
      // the lambda kernel addresses the pool with 32-bit word offsets
      if ((size_t)h.n_dp * sub_stride >= ((size_t)1 << 32)) return D2G_ERR_UNSUPPORTED;
      c->sub_capacity = std::min(c->sub_capacity, ((size_t)1 << 32) - 1);
      HIP_TRY(c->sub_pool.ensure(c->sub_capacity));
      continue;
    }
    if (tm) {  // the task lists now hold the filtered pairs only
      c->cmp_valid = false;
      if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->comparisons = (uint64_t)h.n_gapless + h.n_dp;  // one pair per counted unique
        stats->gapless = h.n_gapless;
        stats->aligned_dp = h.n_dp;
        stats->dp_cells = h.dp_cells;
        if (c->timing) {
          stats->ms_screen = ev_ms(c, 0, 1);
          stats->ms_plan = ev_ms(c, 1, 2);
          stats->ms_dp = ev_ms(c, 2, 3);
          stats->ms_lambda = ev_ms(c, 3, 4);  // the counting kernel
          stats->ms_total = ev_ms(c, 0, 4);
        }
      }
      if (h.err_flags & kErrTraceback) return D2G_ERR_TRACEBACK;
      if (h.err_flags & kErrGeometry) return D2G_ERR_UNSUPPORTED;
      return D2G_OK;
    }
    c->last_nc = nc;
    // the aligned pairs of this call stay resident for d2g_loglik_*
    c->cmp_ctr = h;
    c->cmp_ctr_uploaded = false;
    c->cmp_valid = true;
    c->cmp_sub_stride = sub_stride;
    c->cmp_pairs = (uint64_t)h.n_gapless + h.n_dp;
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->comparisons = npairs;
      stats->skipped = h.n_skipped;
      stats->shrouded = h.n_shrouded;
      stats->gapless = h.n_gapless;
      stats->aligned_dp = h.n_dp;
      stats->dp_cells = h.dp_cells;
      if (c->timing) {
        stats->ms_screen = ev_ms(c, 0, 1);
        stats->ms_plan = ev_ms(c, 1, 2);
        stats->ms_dp = ev_ms(c, 2, 3);
        stats->ms_lambda = ev_ms(c, 3, 4);
        stats->ms_total = ev_ms(c, 0, 4);
      }
    }
    if (h.err_flags & kErrTraceback) return D2G_ERR_TRACEBACK;
    if (h.err_flags & kErrGeometry) return D2G_ERR_UNSUPPORTED;
    return D2G_OK;
  }
  return D2G_ERR_HIP;
}

// The compare: raw_align_dp's routing (nwalign.rs:1315-1414) per pair.
// With vectorized parameters, pairs whose longer side exceeds 3500 nt take
// the scalar ends-free aligner (nwalign.rs:1379-1385): a unique set on both
// sides of that length runs two passes, the vectorized one over the pairs
// of short centers with short uniques, the row-band one over the rest.
static int compare_device(d2g_ctx* c, const uint32_t* centers, uint32_t nc, const double* err,
                          uint32_t ncol, const d2g_align_params* p, int use_quals, int greedy,
                          const uint8_t* lock, d2g_stats* stats, bool fetch = false,
                          double* out_lambda = nullptr, uint32_t* out_hamming = nullptr,
                          uint8_t* out_skipped = nullptr, const TransMode* tm = nullptr) {
  if (!c || !centers || !err || !p || nc == 0 || ncol == 0) return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  const UniqueDev& u = c->u;
  const uint32_t n = u.n;
  if ((uint32_t)p->kmer_size != u.k && p->use_kmers) return D2G_ERR_ARG;
  if (use_quals && c->has_qual && c->max_qual >= ncol) return D2G_ERR_ARG;  // pval.rs:289, err extension dada.rs:337-384
  for (uint32_t i = 0; i < nc; ++i)
    if (centers[i] >= n) return D2G_ERR_ARG;
  const size_t npairs = (size_t)nc * n;
  // (the pairing passes form at most one pair per unique)
  if (!tm && npairs > 0xFFFFFFFFull) return D2G_ERR_UNSUPPORTED;
  int homo = 0;
  const int kind = dp_kind(p, u.minlen, u.maxlen, &homo);
  if (kind >= 0 || p->band == 0) {  // band 0: every pair is gapless
    const Route rt{kind < 0 ? kDpVectorized : kind, homo, u.minlen, u.maxlen, 0xFFFFFFFFu, 0, n,
                   0, n, true};
    return compare_pass(c, centers, nc, err, ncol, p, use_quals, greedy, lock, stats, fetch,
                        out_lambda, out_hamming, out_skipped, tm, rt);
  }
  const uint32_t n_short =
      (uint32_t)(std::upper_bound(c->h_len.begin(), c->h_len.end(), kVectorizedMaxLen) - c->h_len.begin());
  const Route rv{kDpVectorized, 0, u.minlen, c->h_len[n_short - 1], kVectorizedMaxLen, 0, n_short,
                 0, 0, false};
  const Route rr{kDpRowBand, 0, u.minlen, u.maxlen, kVectorizedMaxLen, n_short, n, 0, n, false};
  d2g_stats s1{}, s2{};
  TransMode tm1;
  if (tm) {
    tm1 = *tm;
    tm1.aux = false;  // compute_aux's per-raw pass runs once, after both
  }
  int rc = compare_pass(c, centers, nc, err, ncol, p, use_quals, greedy, lock, &s1, false, nullptr,
                        nullptr, nullptr, tm ? &tm1 : nullptr, rv);
  if (rc != D2G_OK) return rc;
  rc = compare_pass(c, centers, nc, err, ncol, p, use_quals, greedy, lock, &s2, fetch, out_lambda,
                    out_hamming, out_skipped, tm, rr);
  c->cmp_valid = false;  // the likelihood entries take a single-pass compare
  if (stats) {
    *stats = s2;
    stats->comparisons = tm ? s1.comparisons + s2.comparisons : npairs;
    stats->skipped += s1.skipped;
    stats->shrouded += s1.shrouded;
    stats->gapless += s1.gapless;
    stats->aligned_dp += s1.aligned_dp;
    stats->dp_cells += s1.dp_cells;
    stats->ms_screen += s1.ms_screen;
    stats->ms_plan += s1.ms_plan;
    stats->ms_dp += s1.ms_dp;
    stats->ms_lambda += s1.ms_lambda;
    stats->ms_total += s1.ms_total;
  }
  return rc;
}

int d2g_compare_batch_device(d2g_ctx* c, const uint32_t* centers, uint32_t nc, const double* err,
                             uint32_t ncol, const d2g_align_params* p, int use_quals, int greedy,
                             const uint8_t* lock, d2g_stats* stats) {
  return compare_device(c, centers, nc, err, ncol, p, use_quals, greedy, lock, stats);
}

int d2g_fetch_results(d2g_ctx* c, double* out_lambda, uint32_t* out_hamming, uint8_t* out_skipped) {
  if (!c || !c->loaded || c->last_nc == 0) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  HIP_TRY(enqueue_fetch(c, c->last_nc, out_lambda, out_hamming, out_skipped));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return D2G_OK;
}

int d2g_compare_batch(d2g_ctx* c, const uint32_t* centers, uint32_t nc, const double* err,
                      uint32_t ncol, const d2g_align_params* p, int use_quals, int greedy,
                      const uint8_t* lock, double* out_lambda, uint32_t* out_hamming,
                      uint8_t* out_skipped, d2g_stats* stats) {
  return compare_device(c, centers, nc, err, ncol, p, use_quals, greedy, lock, stats, true,
                        out_lambda, out_hamming, out_skipped);
}

int d2g_compare_center(d2g_ctx* c, uint32_t center, const double* err, uint32_t ncol,
                       const d2g_align_params* p, int use_quals, int greedy, const uint8_t* lock,
                       double* out_lambda, uint32_t* out_hamming, uint8_t* out_skipped,
                       d2g_stats* stats) {
  return d2g_compare_batch(c, &center, 1, err, ncol, p, use_quals, greedy, lock, out_lambda,
                           out_hamming, out_skipped, stats);
}

// ---- the device store ------------------------------------------------------
static_assert(sizeof(d2g_comparison) == sizeof(ShuffleComp), "store record layout");
static_assert(offsetof(d2g_comparison, index) == offsetof(ShuffleComp, raw), "store record layout");
static_assert(offsetof(d2g_comparison, cluster) == offsetof(ShuffleComp, cl), "store record layout");
static_assert(offsetof(d2g_comparison, hamming) == offsetof(ShuffleComp, ham), "store record layout");
static_assert(offsetof(d2g_comparison, lambda) == offsetof(ShuffleComp, lambda), "store record layout");

int d2g_store_reset(d2g_ctx* c) {
  if (!c) return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  HIP_TRY(c->e_minmax.ensure(c->u.n));
  HIP_TRY(launch_store_reset(c->e_minmax.p, c->u.n, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->store_n = 0;
  c->store_valid = true;
  return D2G_OK;
}

// Room for `need` records, keeping the ones stored so far.
static hipError_t store_reserve(d2g_ctx* c, size_t need) {
  if (c->store && need <= c->store_cap) return hipSuccess;
  const size_t cap = std::max(need, 2 * c->store_cap + 4096);
  ShuffleComp* q = nullptr;
  hipError_t e = hipMalloc((void**)&q, cap * sizeof(ShuffleComp));
  if (e != hipSuccess) return e;
  if (c->store_n) {
    if ((e = hipMemcpyAsync(q, c->store, c->store_n * sizeof(ShuffleComp), hipMemcpyDeviceToDevice,
                            c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
      (void)hipFree(q);
      return e;
    }
  }
  if (c->store) (void)hipFree(c->store);
  c->store = q;
  c->store_cap = cap;
  return hipSuccess;
}

int d2g_compare_store(d2g_ctx* c, const uint32_t* centers, uint32_t nc, uint32_t first_cluster,
                      double total_reads, const double* err, uint32_t ncol,
                      const d2g_align_params* p, int use_quals, int greedy, const uint8_t* lock,
                      uint64_t capacity, d2g_comparison* out_stored, uint64_t* out_nstored,
                      double* out_self_lambda, d2g_stats* stats) {
  if (!c || !out_nstored || nc == 0) return D2G_ERR_ARG;
  *out_nstored = 0;
  if (!c->loaded || !c->store_valid) return D2G_ERR_STATE;
  const uint32_t n = c->u.n;
  const size_t npairs = (size_t)nc * n;
  // grid rows per center; the shuffle addresses the store with 32-bit indices
  if (nc > 65535u || c->store_n + npairs >= ((uint64_t)1 << 32)) return D2G_ERR_UNSUPPORTED;
  const int rc = compare_device(c, centers, nc, err, ncol, p, use_quals, greedy, lock, stats);
  if (rc != D2G_OK) return rc;
  hipStream_t s = c->stream;
  const uint32_t nrb = store_blocks(n);
  const size_t m = (size_t)nrb * nc;
  HIP_TRY(store_reserve(c, c->store_n + npairs));
  HIP_TRY(c->store_flag.ensure(npairs));
  HIP_TRY(c->store_bcount.ensure(m + 1));
  HIP_TRY(c->store_self.ensure(nc));
  StoreArgs sa;
  sa.n = n;
  sa.nc = nc;
  sa.inv = c->inv.p;
  sa.centers = c->centers.p;
  sa.reads = c->u.reads;
  sa.lam = c->out_lambda.p;
  sa.ham = c->out_hamming.p;
  sa.total = total_reads;
  sa.e_minmax = c->e_minmax.p;
  sa.flag = c->store_flag.p;
  sa.bcount = c->store_bcount.p;
  sa.self_out = c->store_self.p;
  sa.store = c->store + c->store_n;
  sa.first_cluster = first_cluster;
  HIP_TRY(launch_store(sa, s));
  // one synchronisation for the count, the self lambdas and the first
  // kFirst kept comparisons; a longer list takes a second copy
  constexpr size_t kFirst = 4096;
  const size_t first = out_stored ? std::min<size_t>({(size_t)capacity, kFirst, npairs}) : 0;
  const size_t self_off = 8, rec_off = self_off + (size_t)nc * 8;
  const size_t pin_need = rec_off + first * sizeof(ShuffleComp);
  if (c->store_pin_cap < pin_need) {
    if (c->store_pin) (void)hipHostFree(c->store_pin);
    c->store_pin = nullptr;
    c->store_pin_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&c->store_pin, std::max<size_t>(pin_need, 1 << 16),
                          hipHostMallocDefault));
    c->store_pin_cap = std::max<size_t>(pin_need, 1 << 16);
  }
  HIP_TRY(hipMemcpyAsync(c->store_pin, c->store_bcount.p + m, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(c->store_pin + self_off, c->store_self.p, (size_t)nc * 8,
                         hipMemcpyDeviceToHost, s));
  if (first)
    HIP_TRY(hipMemcpyAsync(c->store_pin + rec_off, c->store + c->store_n,
                           first * sizeof(ShuffleComp), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  uint32_t cnt;
  std::memcpy(&cnt, c->store_pin, 4);
  if (out_self_lambda) std::memcpy(out_self_lambda, c->store_pin + self_off, (size_t)nc * 8);
  const uint64_t at = c->store_n;
  c->store_n += cnt;
  *out_nstored = cnt;
  if (!out_stored) return D2G_OK;
  if (cnt > capacity) return D2G_ERR_ARG;  // kept on the device: d2g_store_fetch
  if (cnt > first) {
    // the rest through the pinned area too (grown as needed), one copy
    const size_t need = rec_off + (size_t)cnt * sizeof(ShuffleComp);
    if (c->store_pin_cap < need) {
      (void)hipHostFree(c->store_pin);
      c->store_pin = nullptr;
      c->store_pin_cap = 0;
      const size_t cap = std::max(need, 2 * pin_need);
      HIP_TRY(hipHostMalloc((void**)&c->store_pin, cap, hipHostMallocDefault));
      c->store_pin_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(c->store_pin + rec_off, c->store + at, (size_t)cnt * sizeof(ShuffleComp),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  std::memcpy(out_stored, c->store_pin + rec_off, (size_t)cnt * sizeof(ShuffleComp));
  return D2G_OK;
}

int d2g_store_size(d2g_ctx* c, uint64_t* out_count) {
  if (!c || !out_count) return D2G_ERR_ARG;
  if (!c->store_valid) return D2G_ERR_STATE;
  *out_count = c->store_n;
  return D2G_OK;
}

int d2g_store_fetch(d2g_ctx* c, uint64_t first, uint64_t count, d2g_comparison* out) {
  if (!c || (count && !out)) return D2G_ERR_ARG;
  if (!c->store_valid) return D2G_ERR_STATE;
  if (first > c->store_n || count > c->store_n - first) return D2G_ERR_ARG;
  if (count == 0) return D2G_OK;
  (void)hipSetDevice(c->device);
  HIP_TRY(hipMemcpyAsync(out, c->store + first, count * sizeof(ShuffleComp), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return D2G_OK;
}

// The transition-by-quality counts of a finished partition: each counted
// unique aligned to its center with *fp, counts per (transition, quality)
// summed exactly on the device; ncol_out columns, qualities >= ncol_out
// clamped (clamp) or skipped; u32 results saturated (saturate) or wrapped.
// compute_aux's per-raw outputs besides the transitions (caller order):
// Sub.nsubs per counted unique (D2G_NO_CENTER otherwise) and cluster_quality
// (error.rs:251-305) as read-weighted means per (cluster, center position).
struct AuxSpec {
  uint32_t ncl;
  const uint32_t* cluster_of;
  uint32_t maxlen;
  uint32_t* out_nsubs;
  double* out_quality;
};

static int transitions_run(d2g_ctx* c, const uint32_t* center_of, const uint32_t* weight,
                           const d2g_align_params& fp, bool quals, uint32_t ncol_out, bool clamp,
                           bool saturate, uint32_t* out_mat, const AuxSpec* aux = nullptr) {
  (void)hipSetDevice(c->device);
  const uint32_t n = c->u.n;
  // qualities are bytes: with skip, columns from 256 on never count; with
  // clamp, min(q, ncol - 1) = q there too
  const uint32_t ncol_dev = std::min(ncol_out, 256u);
  std::vector<uint32_t> centers, slot(n, D2G_NO_CENTER);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t ci = center_of[i];
    if (ci == D2G_NO_CENTER || weight[i] == 0) continue;
    if (ci >= n) return D2G_ERR_ARG;
    if (slot[ci] == D2G_NO_CENTER) {
      slot[ci] = (uint32_t)centers.size();
      centers.push_back(ci);
    }
  }
  std::fill(out_mat, out_mat + (size_t)16 * ncol_out, 0u);
  c->tr_stats = d2g_stats{};
  if (aux) {
    std::fill(aux->out_nsubs, aux->out_nsubs + n, D2G_NO_CENTER);
    std::fill(aux->out_quality, aux->out_quality + (size_t)aux->ncl * aux->maxlen,
              std::numeric_limits<double>::quiet_NaN());
  }
  if (centers.empty()) return D2G_OK;
  // pinned staging (free: the previous call ended with a synchronisation)
  const size_t pin_need = (size_t)n * 16;
  if (c->tr_pin_cap < pin_need) {
    if (c->tr_pin) (void)hipHostFree(c->tr_pin);
    c->tr_pin = nullptr;
    c->tr_pin_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&c->tr_pin, pin_need, hipHostMallocDefault));
    c->tr_pin_cap = pin_need;
  }
  uint64_t* aoff = reinterpret_cast<uint64_t*>(c->tr_pin);
  uint32_t* want = reinterpret_cast<uint32_t*>(c->tr_pin + (size_t)n * 8);
  uint32_t* w = want + n;
  // per sorted unique: its center (sorted index) or none, and its weight
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t i = c->h_orig[s];
    const bool counted = center_of[i] != D2G_NO_CENTER && weight[i] != 0;
    want[s] = counted ? c->h_inv[center_of[i]] : D2G_NO_CENTER;
    w[s] = counted ? weight[i] : 0u;
  }
  // one alignment slot of 2 * longest length per unique (the strings of an
  // ends-free alignment are at most len1 + len2 long)
  const uint64_t stride = 2ull * c->u.maxlen;
  hipStream_t st = c->stream;
  HIP_TRY(c->tr_want.ensure(n));
  HIP_TRY(c->tr_weight.ensure(n));
  HIP_TRY(c->tr_mat.ensure((size_t)16 * ncol_dev));
  HIP_TRY(c->a_al0.ensure(n * stride));
  HIP_TRY(c->a_al1.ensure(n * stride));
  HIP_TRY(c->a_aloff.ensure(n));
  HIP_TRY(c->a_allen.ensure(n));
  HIP_TRY(hipMemcpyAsync(c->tr_want.p, want, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->tr_weight.p, w, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  if (c->tr_aoff_for != c->a_aloff.p || c->tr_aoff_n != n || c->tr_aoff_stride != stride) {
    for (uint32_t s = 0; s < n; ++s) aoff[s] = s * stride;
    HIP_TRY(hipMemcpyAsync(c->a_aloff.p, aoff, n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    c->tr_aoff_for = c->a_aloff.p;
    c->tr_aoff_n = n;
    c->tr_aoff_stride = stride;
  }
  HIP_TRY(launch_zero(reinterpret_cast<uint32_t*>(c->tr_mat.p), (size_t)32 * ncol_dev, st));
  TransMode tm;
  tm.want = c->tr_want.p;
  tm.weight = c->tr_weight.p;
  tm.al0 = c->a_al0.p;
  tm.al1 = c->a_al1.p;
  tm.al_off = c->a_aloff.p;
  tm.al_len = c->a_allen.p;
  tm.use_quals = quals ? 1 : 0;
  tm.ncol = ncol_dev;
  tm.clamp = clamp ? 1 : 0;
  tm.mat = c->tr_mat.p;
  std::vector<uint32_t> order, cl_s;
  if (aux) {
    // the counted uniques grouped by cluster (a counting sort by cluster)
    cl_s.assign(n, 0u);
    std::vector<uint32_t> cnt(aux->ncl + 1, 0u);
    for (uint32_t s = 0; s < n; ++s)
      if (w[s]) {
        cl_s[s] = aux->cluster_of[c->h_orig[s]];
        ++cnt[cl_s[s] + 1];
      }
    for (uint32_t k = 0; k < aux->ncl; ++k) cnt[k + 1] += cnt[k];
    order.assign(cnt[aux->ncl], 0u);
    for (uint32_t s = 0; s < n; ++s)
      if (w[s]) order[cnt[cl_s[s]]++] = s;
    const size_t cells = (size_t)aux->ncl * aux->maxlen;
    HIP_TRY(c->tr_order.ensure(std::max<size_t>(order.size(), 1)));
    HIP_TRY(c->tr_cluster.ensure(n));
    HIP_TRY(c->tr_nsubs.ensure(n));
    HIP_TRY(c->tr_qreads.ensure(std::max<size_t>(cells, 1)));
    HIP_TRY(c->tr_qsum.ensure(std::max<size_t>(cells, 1)));
    if (!order.empty())
      HIP_TRY(hipMemcpyAsync(c->tr_order.p, order.data(), order.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->tr_cluster.p, cl_s.data(), n * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st));
    HIP_TRY(launch_zero(c->tr_qreads.p, std::max<size_t>(cells, 1), st));
    HIP_TRY(launch_zero(reinterpret_cast<uint32_t*>(c->tr_qsum.p), 2 * std::max<size_t>(cells, 1), st));
    HIP_TRY(hipStreamSynchronize(st));  // pageable host vectors
    tm.aux = true;
    tm.order = c->tr_order.p;
    tm.n_order = (uint32_t)order.size();
    tm.cluster = c->tr_cluster.p;
    tm.maxlen = aux->maxlen;
    tm.nsubs = c->tr_nsubs.p;
    tm.qual_reads = c->tr_qreads.p;
    tm.qual_sum = c->tr_qsum.p;
  }
  // the lambda stage does not run: a zero one-column error matrix stands in
  const double zero_err[16] = {0};
  HIP_TRY_INT(compare_device(c, centers.data(), (uint32_t)centers.size(), zero_err, 1, &fp, 0, 0,
                             nullptr, &c->tr_stats, false, nullptr, nullptr, nullptr, &tm));
  std::vector<unsigned long long> m((size_t)16 * ncol_dev);
  HIP_TRY(hipMemcpy(m.data(), c->tr_mat.p, m.size() * sizeof(m[0]), hipMemcpyDeviceToHost));
  for (uint32_t t = 0; t < 16; ++t)
    for (uint32_t q = 0; q < ncol_dev; ++q) {
      const unsigned long long v = m[(size_t)t * ncol_dev + q];
      out_mat[(size_t)t * ncol_out + q] =
          saturate ? (uint32_t)std::min<unsigned long long>(v, 0xFFFFFFFFull) : (uint32_t)v;
    }
  if (aux) {
    std::vector<uint32_t> ns(n);
    HIP_TRY(hipMemcpy(ns.data(), c->tr_nsubs.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t k : order) aux->out_nsubs[c->h_orig[k]] = ns[k];
    if (quals) {
      // qual_sum / nreads per position with reads (error.rs:296-302); the
      // f64 sum of the integer terms is exact, so its order does not matter
      const size_t cells = (size_t)aux->ncl * aux->maxlen;
      std::vector<uint32_t> qr(cells);
      std::vector<unsigned long long> qsum(cells);
      HIP_TRY(hipMemcpy(qr.data(), c->tr_qreads.p, cells * sizeof(uint32_t), hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(qsum.data(), c->tr_qsum.p, cells * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost));
      for (size_t i = 0; i < cells; ++i)
        if (qr[i] > 0) aux->out_quality[i] = (double)qsum[i] / (double)qr[i];
    }
  }
  return D2G_OK;
}

// learn-errors' aux transition counts (compute_aux, dada.rs:610-631:
// compute_final_subs dada.rs:646-673 with the k-mer screen off, then
// transition_counts error.rs:194-236).
int d2g_transition_counts(d2g_ctx* c, const uint32_t* center_of, const uint32_t* weight,
                          const d2g_align_params* p, int use_quals, uint32_t ncol,
                          uint32_t* out_mat) {
  if (!c || !center_of || !weight || !p || !out_mat || ncol == 0) return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  const bool quals = use_quals && c->has_qual;
  d2g_align_params fp = *p;  // dada.rs:612-616
  fp.use_kmers = 0;
  fp.kdist_cutoff = 1.0;
  return transitions_run(c, center_of, weight, fp, quals, quals ? ncol : 1u, false, false,
                         out_mat);
}

// compute_aux's per-raw part (dada.rs:610-631): one final-subs pass over a
// finished partition giving the transitions, each raw's Sub.nsubs and
// cluster_quality.
int d2g_dada_aux(d2g_ctx* c, uint32_t ncl, const uint32_t* centers, const uint32_t* cluster_of,
                 const uint32_t* weight, const d2g_align_params* p, int use_quals, uint32_t ncol,
                 uint32_t maxlen, uint32_t* out_trans, uint32_t* out_nsubs, double* out_quality) {
  if (!c || !centers || !cluster_of || !weight || !p || !out_trans || !out_nsubs || !out_quality ||
      ncol == 0 || ncl == 0 || maxlen == 0)
    return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  const uint32_t n = c->u.n;
  std::vector<uint32_t> center_of(n, D2G_NO_CENTER);
  for (uint32_t k = 0; k < ncl; ++k)
    if (centers[k] != D2G_NO_CENTER && centers[k] >= n) return D2G_ERR_ARG;
  for (uint32_t i = 0; i < n; ++i) {
    if (cluster_of[i] == D2G_NO_CENTER) continue;
    if (cluster_of[i] >= ncl) return D2G_ERR_ARG;
    center_of[i] = centers[cluster_of[i]];
  }
  const bool quals = use_quals && c->has_qual;
  d2g_align_params fp = *p;  // dada.rs:612-616
  fp.use_kmers = 0;
  fp.kdist_cutoff = 1.0;
  AuxSpec aux{ncl, cluster_of, maxlen, out_nsubs, out_quality};
  return transitions_run(c, center_of.data(), weight, fp, quals, quals ? ncol : 1u, false, false,
                         out_trans, &aux);
}

// compute_birth_subs (dada.rs:679-693) and the Sub-derived outputs of
// birth_sub_records (error.rs:328-352) and cluster_stats' birth_qave
// (error.rs:90-100).  Each child center aligned to its parent's center on
// the device (one task per child unique, alignment strings in a slot per
// cluster), the Subs read from the alignments on the host (al2subs,
// nwalign.rs:1169-1224).
int d2g_birth_subs(d2g_ctx* c, uint32_t ncl, const uint32_t* centers, const uint32_t* birth_from,
                   const d2g_align_params* p, int use_quals, uint32_t capacity, uint32_t* out_nsubs,
                   double* out_qave, uint32_t* out_cluster, uint32_t* out_pos, uint8_t* out_nt0,
                   uint8_t* out_nt1, uint8_t* out_qual, uint32_t* out_nrec) {
  if (!c || !centers || !birth_from || !p || !out_nsubs || !out_qave || !out_nrec || ncl == 0)
    return D2G_ERR_ARG;
  if (capacity && (!out_cluster || !out_pos || !out_nt0 || !out_nt1 || !out_qual)) return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  const uint32_t n = c->u.n;
  const bool quals = use_quals && c->has_qual;
  std::fill(out_nsubs, out_nsubs + ncl, D2G_NO_CENTER);
  std::fill(out_qave, out_qave + ncl, std::numeric_limits<double>::quiet_NaN());
  *out_nrec = 0;
  // the child unique of each birth pair (sorted index) and its cluster
  std::vector<uint32_t> want(n, D2G_NO_CENTER), one(n, 0u), slot_of(n, D2G_NO_CENTER);
  std::vector<uint32_t> child_of(ncl, D2G_NO_CENTER);
  uint32_t npair = 0;
  for (uint32_t ci = 0; ci < ncl; ++ci) {
    if (centers[ci] != D2G_NO_CENTER && centers[ci] >= n) return D2G_ERR_ARG;
    if (ci == 0) continue;  // cluster 0 has no parent (dada.rs:684)
    if (birth_from[ci] >= ncl) return D2G_ERR_ARG;
    const uint32_t pc = centers[birth_from[ci]], cc = centers[ci];
    if (pc == D2G_NO_CENTER || cc == D2G_NO_CENTER) continue;
    const uint32_t s = c->h_inv[cc];
    if (want[s] != D2G_NO_CENTER) return D2G_ERR_ARG;  // one cluster per center
    want[s] = c->h_inv[pc];
    one[s] = 1;
    slot_of[s] = npair++;
    child_of[ci] = s;
  }
  if (npair == 0) return D2G_OK;
  // birth alignments: the k-mer screen as given, cutoff 1.0 (dada.rs:617-621)
  d2g_align_params bp = *p;
  bp.kdist_cutoff = 1.0;
  const uint64_t stride = 2ull * c->u.maxlen;
  std::vector<uint64_t> aoff(n, 0);
  for (uint32_t s = 0; s < n; ++s)
    if (slot_of[s] != D2G_NO_CENTER) aoff[s] = slot_of[s] * stride;
  hipStream_t st = c->stream;
  HIP_TRY(c->tr_want.ensure(n));
  HIP_TRY(c->tr_weight.ensure(n));
  HIP_TRY(c->tr_mat.ensure(16));
  HIP_TRY(c->a_al0.ensure(npair * stride));
  HIP_TRY(c->a_al1.ensure(npair * stride));
  HIP_TRY(c->a_aloff.ensure(n));
  HIP_TRY(c->a_allen.ensure(n));
  c->tr_aoff_for = nullptr;  // the offsets below are the birth pass's
  HIP_TRY(hipMemcpyAsync(c->tr_want.p, want.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->tr_weight.p, one.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->a_aloff.p, aoff.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // pageable host vectors
  TransMode tm;
  tm.want = c->tr_want.p;
  tm.weight = c->tr_weight.p;
  tm.al0 = c->a_al0.p;
  tm.al1 = c->a_al1.p;
  tm.al_off = c->a_aloff.p;
  tm.al_len = c->a_allen.p;
  tm.use_quals = 0;
  tm.ncol = 1;
  tm.clamp = 0;
  tm.mat = c->tr_mat.p;
  tm.count = false;
  std::vector<uint32_t> parents;
  for (uint32_t s = 0; s < n; ++s)
    if (want[s] != D2G_NO_CENTER) parents.push_back(c->h_orig[want[s]]);
  std::sort(parents.begin(), parents.end());
  parents.erase(std::unique(parents.begin(), parents.end()), parents.end());
  const double zero_err[16] = {0};
  HIP_TRY_INT(compare_device(c, parents.data(), (uint32_t)parents.size(), zero_err, 1, &bp, 0, 0,
                             nullptr, &c->tr_stats, false, nullptr, nullptr, nullptr, &tm));
  // which pairs took the gapless alignment, and the DP alignments
  Counters h;
  HIP_TRY(hipMemcpy(&h, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost));
  std::vector<GaplessTask> glt(h.n_gapless);
  if (h.n_gapless)
    HIP_TRY(hipMemcpy(glt.data(), c->gl.p, h.n_gapless * sizeof(GaplessTask), hipMemcpyDeviceToHost));
  std::vector<uint8_t> gl_flag(n, 0);
  for (const GaplessTask& t : glt) gl_flag[t.u_idx] = 1;
  std::vector<uint32_t> alen(n);
  std::vector<uint8_t> a0(npair * stride), a1(npair * stride);
  HIP_TRY(hipMemcpy(alen.data(), c->a_allen.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(a0.data(), c->a_al0.p, a0.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(a1.data(), c->a_al1.p, a1.size(), hipMemcpyDeviceToHost));
  std::vector<uint8_t> g0, g1;
  uint32_t nrec = 0;
  for (uint32_t ci = 1; ci < ncl; ++ci) {
    const uint32_t s = child_of[ci];
    if (s == D2G_NO_CENTER) continue;
    const uint32_t ps = want[s];
    const uint8_t *x0, *x1;
    uint32_t L;
    if (gl_flag[s]) {  // align_gapless_with_buf (nwalign.rs:645-655)
      const uint32_t l0 = c->h_len[ps], l1 = c->h_len[s];
      L = std::max(l0, l1);
      g0.assign(L, (uint8_t)'-');
      g1.assign(L, (uint8_t)'-');
      std::memcpy(g0.data(), &c->h_seq[c->h_soff[ps]], l0);
      std::memcpy(g1.data(), &c->h_seq[c->h_soff[s]], l1);
      x0 = g0.data();
      x1 = g1.data();
    } else {
      x0 = &a0[slot_of[s] * stride];
      x1 = &a1[slot_of[s] * stride];
      L = alen[s];
    }
    const uint8_t* q1 = quals ? &c->h_qual[c->h_soff[s]] : nullptr;
    static const uint8_t kDecode[6] = {'?', 'A', 'C', 'G', 'T', 'N'};  // nt_decode, misc.rs:360-370
    uint32_t ns = 0;
    double qsum = 0.0;
    int64_t i0 = -1, i1 = -1;
    for (uint32_t k = 0; k < L; ++k) {
      const bool nt0 = x0[k] >= 1 && x0[k] <= 5, nt1 = x1[k] >= 1 && x1[k] <= 5;
      i0 += nt0;
      i1 += nt1;
      if (nt0 && nt1 && x0[k] != x1[k] && x0[k] != 5 && x1[k] != 5) {
        const uint8_t qv = q1 ? q1[i1] : 0;
        if (nrec < capacity) {
          out_cluster[nrec] = ci;
          out_pos[nrec] = (uint32_t)i0;
          out_nt0[nrec] = kDecode[x0[k]];
          out_nt1[nrec] = kDecode[x1[k]];
          out_qual[nrec] = qv;
        }
        ++nrec;
        ++ns;
        qsum += (double)qv;
      }
    }
    out_nsubs[ci] = ns;
    if (quals && ns > 0) out_qave[ci] = qsum / (double)ns;  // error.rs:90-100
  }
  *out_nrec = nrec;
  return nrec > capacity ? D2G_ERR_ARG : D2G_OK;
}

int d2g_transition_stats(d2g_ctx* c, d2g_stats* out) {
  if (!c || !out) return D2G_ERR_ARG;
  *out = c->tr_stats;
  return D2G_OK;
}

// build_trans_mat (learn_errors.rs:618-690): the counts of one sample's
// result that each learn-errors iteration sums.
int d2g_build_trans_mat(d2g_ctx* c, const uint32_t* center_of, const uint32_t* weight,
                        const d2g_align_params* p, uint32_t nq, uint32_t* out_mat) {
  if (!c || !center_of || !weight || !p || !out_mat || nq == 0) return D2G_ERR_ARG;
  if (!c->loaded) return D2G_ERR_STATE;
  // inputs without qualities add nothing (learn_errors.rs:648-651)
  if (!c->has_qual) {
    std::fill(out_mat, out_mat + (size_t)16 * nq, 0u);
    return D2G_OK;
  }
  // the center and query Raws carry no k-mer vectors, so kdist is 0.0 and
  // kodist -1.0 (nwalign.rs:1262-1292): the screen passes every pair unless
  // the cutoff is below 0, and only band 0 aligns gaplessly
  if (p->use_kmers && 0.0 > p->kdist_cutoff) {
    std::fill(out_mat, out_mat + (size_t)16 * nq, 0u);
    return D2G_OK;
  }
  d2g_align_params fp = *p;
  fp.use_kmers = 0;
  fp.kdist_cutoff = 1.0;
  return transitions_run(c, center_of, weight, fp, true, nq, true, true, out_mat);
}

// n independent pairs through one DP kernel with the alignment strings left
// in the context's device buffers (a_al0/a_al1 at a_aloff, lengths a_allen);
// *al_total receives the bytes spanned.  Shared by the align and bimera
// entry points.
// The sequences of n pairs: pair i is sequence i1(i) of the first set and
// i2(i) of the second (each set concatenated, offsets with n1 + 1 / n2 + 1
// entries; a null index is the identity).  Pairs that share a sequence
// share its copy on the device.
struct PairSeqs {
  const uint8_t* s1;
  const uint64_t* s1_off;
  uint32_t n1;
  const uint32_t* idx1;
  const uint8_t* s2;
  const uint64_t* s2_off;
  uint32_t n2;
  const uint32_t* idx2;
  uint32_t i1(uint32_t i) const { return idx1 ? idx1[i] : i; }
  uint32_t i2(uint32_t i) const { return idx2 ? idx2[i] : i; }
  uint64_t len1(uint32_t i) const { return s1_off[i1(i) + 1] - s1_off[i1(i)]; }
  uint64_t len2(uint32_t i) const { return s2_off[i2(i) + 1] - s2_off[i2(i)]; }
};

static int align_pairs_device(d2g_ctx* c, uint32_t n, const PairSeqs& ps, int kind, int homo,
                              int band, int match, int mismatch, int gap, int end_gap,
                              int homo_gap, const uint64_t* al_offsets, uint64_t* al_total_out) {
  if (!c || !ps.s1 || !ps.s1_off || !ps.s2 || !ps.s2_off) return D2G_ERR_ARG;
  *al_total_out = 0;
  c->cmp_valid = false;  // the task buffers are reused below
  if (n == 0) return D2G_OK;
  (void)hipSetDevice(c->device);
  for (uint32_t k = 0; k < ps.n1; ++k)
    if (ps.s1_off[k + 1] < ps.s1_off[k]) return D2G_ERR_ARG;
  for (uint32_t k = 0; k < ps.n2; ++k)
    if (ps.s2_off[k + 1] < ps.s2_off[k]) return D2G_ERR_ARG;
  // pool: the first set, then the second
  std::vector<uint32_t> l1(n), l2(n);
  uint32_t lmin = 0xFFFFFFFFu, lmax = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (ps.i1(i) >= ps.n1 || ps.i2(i) >= ps.n2) return D2G_ERR_ARG;
    l1[i] = (uint32_t)ps.len1(i);
    l2[i] = (uint32_t)ps.len2(i);
    if (l1[i] == 0 || l2[i] == 0) return D2G_ERR_ARG;
    if (kind == kDpVectorized && (l1[i] > kVectorizedMaxLen || l2[i] > kVectorizedMaxLen))
      return D2G_ERR_UNSUPPORTED;
    if (l1[i] > 65535 || l2[i] > 65535) return D2G_ERR_UNSUPPORTED;
    lmin = std::min({lmin, l1[i], l2[i]});
    lmax = std::max({lmax, l1[i], l2[i]});
  }
  const uint64_t n1 = ps.s1_off[ps.n1] - ps.s1_off[0], n2 = ps.s2_off[ps.n2] - ps.s2_off[0];
  if (n1 + n2 >= ((uint64_t)1 << 32)) return D2G_ERR_UNSUPPORTED;  // 32-bit pool offsets
  std::vector<uint64_t> aoff(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) aoff[i + 1] = aoff[i] + l1[i] + l2[i];
  if (al_offsets)
    for (uint32_t i = 0; i < n; ++i) aoff[i] = al_offsets[i];
  uint64_t al_total = 0;
  for (uint32_t i = 0; i < n; ++i) al_total = std::max<uint64_t>(al_total, aoff[i] + l1[i] + l2[i]);
  BinSpace bs;
  bs.lmin = (int)lmin;
  bs.range = (int)(lmax - lmin + 1);
  bs.band = band;
  bs.kind = kind;
  bs.homo = homo;
  bs.tile = 0;
  if ((size_t)bs.range * bs.range > (1u << 22)) return D2G_ERR_UNSUPPORTED;
  bs.nbins = bs.range * bs.range;
  std::vector<DpTask> tasks(n);
  std::vector<uint32_t> counts(bs.nbins + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    DpTask& t = tasks[i];
    t.s1_off = ps.s1_off[ps.i1(i)] - ps.s1_off[0];
    t.s2_off = n1 + (ps.s2_off[ps.i2(i)] - ps.s2_off[0]);
    t.out_idx = i;
    t.bin = (l1[i] - lmin) * (uint32_t)bs.range + (l2[i] - lmin);
    t.u_idx = 0;
    t.c_idx = 0;
    counts[t.bin] += 1;
    if (bin_groups(bs, (int)l1[i], (int)l2[i]) == 0 &&
        !(kind == kDpVectorized && wide_fits(vec_geom((int)l1[i], (int)l2[i], band))))
      return D2G_ERR_UNSUPPORTED;
  }
  hipStream_t s = c->stream;
  HIP_TRY(c->a_pool.ensure(n1 + n2 + 1));
  const uint32_t n_regions = (n + kAlignRegion - 1) / kAlignRegion;
  HIP_TRY(c->dp_unsorted.ensure((size_t)n_regions * kAlignRegion));
  HIP_TRY(c->dp_sorted.ensure(n));
  HIP_TRY(c->bin_count.ensure(bs.nbins + 1));
  HIP_TRY(c->bin_start.ensure(bs.nbins + 1));
  HIP_TRY(c->wave_start.ensure(bs.nbins + 1));
  HIP_TRY(c->bin_fill.ensure(bs.nbins + 1));
  HIP_TRY(c->a_al0.ensure(al_total + 1));
  HIP_TRY(c->a_al1.ensure(al_total + 1));
  HIP_TRY(c->a_aloff.ensure(n));
  HIP_TRY(c->a_allen.ensure(n));
  HIP_TRY(hipMemcpyAsync(c->a_pool.p, ps.s1 + ps.s1_off[0], n1, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->a_pool.p + n1, ps.s2 + ps.s2_off[0], n2, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->dp_unsorted.p, tasks.data(), n * sizeof(DpTask), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->bin_count.p, counts.data(), (bs.nbins + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  c->tr_aoff_for = nullptr;  // the transition pass's offsets are overwritten
  HIP_TRY(hipMemcpyAsync(c->a_aloff.p, aoff.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  Counters h0;
  std::memset(&h0, 0, sizeof(h0));
  h0.n_dp = n;
  HIP_TRY(hipMemcpyAsync(c->ctr.p, &h0, sizeof(Counters), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<uint8_t> present((size_t)bs.range, 0);
  for (uint32_t i = 0; i < n; ++i) {
    present[l1[i] - lmin] = 1;
    present[l2[i] - lmin] = 1;
  }
  HIP_TRY_INT(prepare_schedule(c, bs, end_gap, gap, present));
  HIP_TRY(launch_plan(bs, c->bin_count.p, c->bin_start.p, c->wave_start.p, c->bin_fill.p, c->ctr.p,
                      c->sched_info.p, nullptr, 0, nullptr, s));
  HIP_TRY(launch_scatter(c->dp_unsorted.p, c->dp_sorted.p, c->bin_start.p, c->bin_fill.p, nullptr,
                         n_regions, kAlignRegion, n, nullptr, nullptr, nullptr, s));
  NwArgs na;
  std::memset(&na, 0, sizeof(na));
  na.bins = bs;
  na.bin_start = c->bin_start.p;
  na.wave_start = c->wave_start.p;
  na.tasks = c->dp_sorted.p;
  na.pool = c->a_pool.p;
  na.match = match;
  na.mismatch = mismatch;
  na.gap_p = gap;
  na.end_gap_p = end_gap;
  na.homo_gap_p = homo_gap;
  na.ctr = c->ctr.p;
  na.emit_alignment = 1;
  na.al0 = c->a_al0.p;
  na.al1 = c->a_al1.p;
  na.al_off = c->a_aloff.p;
  na.al_len = c->a_allen.p;
  na.grid = c->nw_grid;
  na.hist = c->hist.p;
  na.hist_stride = c->hist_stride;
  na.lds_bytes = c->lds_bytes;
  na.sched = c->sched.p;
  na.sched_off = c->sched_off.p;
  na.info = c->sched_info.p;
  HIP_TRY(kind == kDpVectorized ? launch_nw_vec(na, s) : launch_nw_band(na, s));
  if (kind == kDpVectorized) HIP_TRY(launch_nw_wide(na, c->wide, c->wide_p.p, c->wide_grid, s));
  Counters h;
  HIP_TRY(hipMemcpyAsync(&h, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *al_total_out = al_total;
  if (h.err_flags & kErrTraceback) return D2G_ERR_TRACEBACK;
  if (h.err_flags & kErrGeometry) return D2G_ERR_UNSUPPORTED;
  return D2G_OK;
}

// align_pairs_device, then the strings and lengths to the caller.
static int align_pairs(d2g_ctx* c, uint32_t n, const uint8_t* s1, const uint64_t* s1_off,
                       const uint8_t* s2, const uint64_t* s2_off, int kind, int homo, int band,
                       int match, int mismatch, int gap, int end_gap, int homo_gap, uint8_t* al0,
                       uint8_t* al1, const uint64_t* al_offsets, uint32_t* al_len) {
  if (!al0 || !al1 || !al_len) return D2G_ERR_ARG;
  uint64_t al_total = 0;
  if (!s1 || !s1_off || !s2 || !s2_off) return D2G_ERR_ARG;
  const PairSeqs ps{s1, s1_off, n, nullptr, s2, s2_off, n, nullptr};
  const int rc = align_pairs_device(c, n, ps, kind, homo, band, match, mismatch, gap, end_gap,
                                    homo_gap, al_offsets, &al_total);
  if (rc != D2G_OK || n == 0) return rc;
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(al_len, c->a_allen.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(al0, c->a_al0.p, al_total, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(al1, c->a_al1.p, al_total, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return D2G_OK;
}

int d2g_align_vectorized(d2g_ctx* c, uint32_t n, const uint8_t* s1, const uint64_t* s1_off,
                         const uint8_t* s2, const uint64_t* s2_off, const d2g_vec_scores* sc,
                         uint8_t* al0, uint8_t* al1, const uint64_t* al_offsets, uint32_t* al_len) {
  if (!sc) return D2G_ERR_ARG;
  return align_pairs(c, n, s1, s1_off, s2, s2_off, kDpVectorized, 0, sc->band, sc->match_score,
                     sc->mismatch, sc->gap_p, sc->end_gap_p, sc->gap_p, al0, al1, al_offsets, al_len);
}

int d2g_align_endsfree(d2g_ctx* c, uint32_t n, const uint8_t* s1, const uint64_t* s1_off,
                       const uint8_t* s2, const uint64_t* s2_off, int32_t match, int32_t mismatch,
                       int32_t gap_p, int32_t homo_gap_p, int32_t band, uint8_t* al0, uint8_t* al1,
                       const uint64_t* al_offsets, uint32_t* al_len) {
  // with homo_gap_p == gap_p the homopolymer aligner equals the plain one
  const int homo = homo_gap_p != gap_p ? 1 : 0;
  return align_pairs(c, n, s1, s1_off, s2, s2_off, kDpRowBand, homo, band, match, mismatch, gap_p,
                     0, homo_gap_p, al0, al1, al_offsets, al_len);
}

int d2g_band_supported(uint32_t len1, uint32_t len2, int32_t band, int homo) {
  return band_groups(band_geom((int)len1, (int)len2, band), homo != 0) > 0 ? 1 : 0;
}

// ---- log-likelihood and gradient ------------------------------------------
static int loglik_args(d2g_ctx* c, uint32_t ncol, int use_quals, LoglikArgs* la) {
  if (!c->cmp_valid) return D2G_ERR_STATE;
  if (!c->cmp_ctr_uploaded) {  // the compare's counters (n_gapless, n_dp)
    HIP_TRY(c->last_ctr.ensure(1));
    HIP_TRY(hipMemcpyAsync(c->last_ctr.p, &c->cmp_ctr, sizeof(Counters), hipMemcpyHostToDevice,
                           c->stream));
    c->cmp_ctr_uploaded = true;
  }
  if (use_quals && c->has_qual && c->max_qual >= ncol) return D2G_ERR_ARG;
  // LDS of the backward workgroup: four wavefronts' gradient rows and counts
  // (12 bytes per table entry) and the 64 pairs' staged rows (33 KB)
  if ((size_t)4 * 16 * ncol * 12 + 64 * 1024 > (size_t)160 * 1024) return D2G_ERR_UNSUPPORTED;
  // one workgroup per 64 aligned pairs, at most 4 per CU
  const uint32_t grid = (uint32_t)std::max<size_t>(
      1, std::min<size_t>((size_t)c->cu_count * 4, (c->cmp_pairs + 63) / 64));
  HIP_TRY(c->ll_w.ensure(c->cmp_pairs + 1));
  HIP_TRY(c->ll_part.ensure((size_t)grid * 16 * ncol));
  HIP_TRY(c->ll_out.ensure((size_t)16 * ncol + 1));
  la->u = &c->u;
  la->gapless_tasks = c->gl.p;
  la->dp_tasks = c->dp_sorted.p;
  la->sub_pool = c->sub_pool.p;
  la->sub_stride = c->cmp_sub_stride;
  la->sub_cnt = c->sub_cnt.p;
  la->ncol = ncol;
  la->use_quals = use_quals;
  la->ctr = c->last_ctr.p;
  la->w = c->ll_w.p;
  la->block_part = c->ll_part.p;
  la->grid = grid;
  return D2G_OK;
}

int d2g_loglik_forward(d2g_ctx* c, const float* logE, uint32_t ncol, int use_quals,
                       double* out_loss, float* ms) {
  if (!c || !logE || ncol == 0 || !out_loss) return D2G_ERR_ARG;
  (void)hipSetDevice(c->device);
  LoglikArgs la;
  HIP_TRY_INT(loglik_args(c, ncol, use_quals, &la));
  hipStream_t s = c->stream;
  HIP_TRY(c->ll_logE.ensure((size_t)16 * ncol));
  HIP_TRY(hipMemcpyAsync(c->ll_logE.p, logE, (size_t)16 * ncol * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(c->ev[0], s));
  HIP_TRY(launch_loglik_fwd(la, c->ll_logE.p, c->ll_out.p, s));
  HIP_TRY(hipEventRecord(c->ev[1], s));
  HIP_TRY(hipMemcpyAsync(out_loss, c->ll_out.p, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (ms) *ms = ev_ms(c, 0, 1);
  c->ll_ncol = ncol;
  c->ll_use_quals = use_quals;
  c->ll_have_w = true;
  return D2G_OK;
}

int d2g_loglik_backward(d2g_ctx* c, float* out_grad, double* out_grad64, float* ms) {
  if (!c || (!out_grad && !out_grad64)) return D2G_ERR_ARG;
  if (!c->ll_have_w || !c->cmp_valid) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  LoglikArgs la;
  HIP_TRY_INT(loglik_args(c, c->ll_ncol, c->ll_use_quals, &la));
  hipStream_t s = c->stream;
  const size_t m = (size_t)16 * c->ll_ncol;
  HIP_TRY(hipEventRecord(c->ev[0], s));
  HIP_TRY(launch_loglik_bwd(la, c->ll_out.p, s));
  HIP_TRY(hipEventRecord(c->ev[1], s));
  std::vector<double> g(m);
  HIP_TRY(hipMemcpyAsync(g.data(), c->ll_out.p, m * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (ms) *ms = ev_ms(c, 0, 1);
  for (size_t i = 0; i < m; ++i) {
    if (out_grad) out_grad[i] = (float)g[i];
    if (out_grad64) out_grad64[i] = g[i];
  }
  return D2G_OK;
}

// ---- bimera ---------------------------------------------------------------
// (query, parent) pairs -> get_lr + get_ham_endsfree, 5 words per pair.
static int bimera_lr_pairs(d2g_ctx* c, uint32_t n, const PairSeqs& ps,
                           const d2g_bimera_params* p, uint32_t* out5) {
  if (n == 0) return D2G_OK;
  uint64_t al_total = 0;
  // the vectorized aligner with end_gap 0 and band = max_shift (chimera.rs:234-240)
  const int rc = align_pairs_device(c, n, ps, kDpVectorized, 0, p->max_shift, p->match_score,
                                    p->mismatch, p->gap_p, 0, p->gap_p, nullptr, &al_total);
  if (rc != D2G_OK) return rc;
  hipStream_t s = c->stream;
  HIP_TRY(c->b_out.ensure((size_t)5 * n));
  HIP_TRY(launch_bimera_lr(n, c->a_al0.p, c->a_al1.p, c->a_aloff.p, c->a_allen.p,
                           p->allow_one_off, (uint32_t)p->max_shift, c->b_out.p, s));
  HIP_TRY(hipMemcpyAsync(out5, c->b_out.p, (size_t)5 * n * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return D2G_OK;
}

// get_lr + get_ham_endsfree for n (query, parent) pairs in device batches
// that bound the alignment-string buffers (1 GiB of pair bytes, 2^18 pairs).
// Pair t is query qidx[t] and parent pidx[t] of the two sets; a null index
// means one sequence per pair.
static int bimera_run(d2g_ctx* c, uint32_t n, const uint8_t* q, const uint64_t* q_off,
                      uint32_t nq, const uint32_t* qidx, const uint8_t* par,
                      const uint64_t* p_off, uint32_t np, const uint32_t* pidx,
                      const d2g_bimera_params* p, uint32_t* out_lr, uint32_t* out_ham) {
  if (qidx)
    for (uint32_t t = 0; t < n; ++t)
      if (qidx[t] >= nq) return D2G_ERR_ARG;
  if (pidx)
    for (uint32_t t = 0; t < n; ++t)
      if (pidx[t] >= np) return D2G_ERR_ARG;
  std::vector<uint32_t> o5;
  for (uint32_t i0 = 0; i0 < n;) {
    PairSeqs ps;
    ps.s1 = q;
    ps.s1_off = qidx ? q_off : q_off + i0;
    ps.idx1 = qidx ? qidx + i0 : nullptr;
    ps.s2 = par;
    ps.s2_off = pidx ? p_off : p_off + i0;
    ps.idx2 = pidx ? pidx + i0 : nullptr;
    uint64_t bytes = 0;
    uint32_t k = 0;
    while (i0 + k < n && k < (1u << 18)) {
      bytes += ps.len1(k) + ps.len2(k);
      if (bytes > ((uint64_t)1 << 30) && k > 0) break;
      ++k;
    }
    ps.n1 = qidx ? nq : k;
    ps.n2 = pidx ? np : k;
    o5.resize((size_t)5 * k);
    const int rc = bimera_lr_pairs(c, k, ps, p, o5.data());
    if (rc != D2G_OK) return rc;
    for (uint32_t t = 0; t < k; ++t) {
      for (int f = 0; f < 4; ++f) out_lr[4 * (size_t)(i0 + t) + f] = o5[5 * (size_t)t + f];
      out_ham[i0 + t] = o5[5 * (size_t)t + 4];
    }
    i0 += k;
  }
  return D2G_OK;
}

int d2g_bimera_pairs(d2g_ctx* c, uint32_t n, const uint8_t* q, const uint64_t* q_off,
                     const uint8_t* par, const uint64_t* p_off, const d2g_bimera_params* p,
                     uint32_t* out_lr, uint32_t* out_ham) {
  if (!c || !q || !q_off || !par || !p_off || !p || !out_lr || !out_ham) return D2G_ERR_ARG;
  return bimera_run(c, n, q, q_off, n, nullptr, par, p_off, n, nullptr, p, out_lr, out_ham);
}

// is_bimera_with_buf (chimera.rs:218-292) over each query's parents: pair
// t = (query i, parent t) for t in [par_start[i], par_start[i+1]); pidx maps
// pairs to the parent set (null: parent t is sequence t).
static int is_bimera_run(d2g_ctx* c, uint32_t nq, const uint8_t* q, const uint64_t* q_off,
                         const uint8_t* par, const uint64_t* p_off, uint32_t np,
                         const uint32_t* pidx, const uint32_t* par_start,
                         const d2g_bimera_params* p, uint8_t* out_flag) {
  const uint32_t npairs = par_start[nq];
  std::vector<uint32_t> qidx(npairs);
  for (uint32_t i = 0; i < nq; ++i) {
    if (par_start[i + 1] < par_start[i] || q_off[i + 1] < q_off[i]) return D2G_ERR_ARG;
    for (uint32_t t = par_start[i]; t < par_start[i + 1]; ++t) qidx[t] = i;
  }
  std::vector<uint32_t> lr((size_t)4 * npairs + 4), ham(npairs + 1);
  const int rc = bimera_run(c, npairs, q, q_off, nq, qidx.data(), par, p_off,
                            pidx ? np : npairs, pidx, p, lr.data(), ham.data());
  if (rc != D2G_OK) return rc;
  // maxima over the parents; the early exits of the reference only stop a
  // search already flagged
  for (uint32_t i = 0; i < nq; ++i) {
    const uint32_t sqlen = (uint32_t)(q_off[i + 1] - q_off[i]);
    uint32_t ml = 0, mr = 0, oml = 0, omr = 0, omloo = 0, omroo = 0;
    bool flag = false;
    for (uint32_t t = par_start[i]; t < par_start[i + 1] && !flag; ++t) {
      const uint32_t* e = lr.data() + 4 * (size_t)t;
      if (e[0] + e[1] >= sqlen) continue;
      ml = std::max(ml, e[0]);
      mr = std::max(mr, e[1]);
      if (p->allow_one_off && ham[t] >= (uint32_t)p->min_one_off_par_dist) {
        oml = std::max(oml, e[0]);
        omr = std::max(omr, e[1]);
        omloo = std::max(omloo, e[2]);
        omroo = std::max(omroo, e[3]);
      }
      if (ml + mr >= sqlen) flag = true;
      if (p->allow_one_off && (oml + omroo >= sqlen || omloo + omr >= sqlen)) flag = true;
    }
    out_flag[i] = flag ? 1 : 0;
  }
  return D2G_OK;
}

int d2g_is_bimera_batch(d2g_ctx* c, uint32_t nq, const uint8_t* q, const uint64_t* q_off,
                        const uint8_t* par, const uint64_t* p_off, const uint32_t* par_start,
                        const d2g_bimera_params* p, uint8_t* out_flag) {
  if (!c || !q || !q_off || !par || !p_off || !par_start || !p || !out_flag) return D2G_ERR_ARG;
  return is_bimera_run(c, nq, q, q_off, par, p_off, 0, nullptr, par_start, p, out_flag);
}

int d2g_is_bimera_pool(d2g_ctx* c, uint32_t nq, const uint8_t* q, const uint64_t* q_off,
                       uint32_t npool, const uint8_t* pool, const uint64_t* pool_off,
                       const uint32_t* par_idx, const uint32_t* par_start,
                       const d2g_bimera_params* p, uint8_t* out_flag) {
  if (!c || !q || !q_off || !pool || !pool_off || !par_idx || !par_start || !p || !out_flag ||
      npool == 0)
    return D2G_ERR_ARG;
  return is_bimera_run(c, nq, q, q_off, pool, pool_off, npool, par_idx, par_start, p, out_flag);
}

int d2g_table_bimera2(d2g_ctx* c, const uint32_t* mat, uint32_t nrow, uint32_t ncol,
                      const uint8_t* seqs, const uint64_t* off, double min_fold,
                      uint32_t min_abund, const d2g_bimera_params* p, uint32_t* out_nflag,
                      uint32_t* out_nsam) {
  if (!c || !mat || !seqs || !off || !p || !out_nflag || !out_nsam) return D2G_ERR_ARG;
  // every (j, k) the reference would align: k qualifies as a parent of j in
  // some sample (chimera.rs:351-374)
  std::vector<std::vector<uint32_t>> need(ncol);
  std::vector<uint8_t> mark(ncol);
  for (uint32_t j = 0; j < ncol; ++j) {
    std::fill(mark.begin(), mark.end(), 0);
    for (uint32_t i = 0; i < nrow; ++i) {
      const uint32_t ja = mat[i + (size_t)j * nrow];
      if (ja == 0) continue;
      for (uint32_t k = 0; k < ncol; ++k) {
        const uint32_t ka = mat[i + (size_t)k * nrow];
        if ((double)ka <= min_fold * (double)ja || ka < min_abund) continue;
        if (!mark[k]) {
          mark[k] = 1;
          need[j].push_back(k);
        }
      }
    }
  }
  std::vector<uint32_t> pj, pk;
  for (uint32_t j = 0; j < ncol; ++j)
    for (uint32_t k : need[j]) {
      pj.push_back(j);
      pk.push_back(k);
    }
  const uint32_t npairs = (uint32_t)pj.size();
  for (uint32_t j = 0; j < ncol; ++j)
    if (off[j + 1] < off[j]) return D2G_ERR_ARG;
  // both sides index the one sequence table
  std::vector<uint32_t> lr((size_t)4 * npairs + 4), ham(npairs + 1);
  int rc = bimera_run(c, npairs, seqs, off, ncol, pj.data(), seqs, off, ncol, pk.data(), p,
                      lr.data(), ham.data());
  if (rc != D2G_OK) return rc;
  // per column: the sample loop of chimera.rs:351-431 over the cached pairs
  std::vector<int64_t> slot(ncol, -1);
  uint32_t t0 = 0;
  for (uint32_t j = 0; j < ncol; ++j) {
    for (uint32_t x = 0; x < need[j].size(); ++x) slot[need[j][x]] = (int64_t)(t0 + x);
    const uint32_t sqlen = (uint32_t)(off[j + 1] - off[j]);
    uint32_t nsam = 0, nflag = 0;
    for (uint32_t i = 0; i < nrow; ++i) {
      const uint32_t ja = mat[i + (size_t)j * nrow];
      if (ja == 0) continue;
      ++nsam;
      uint32_t ml = 0, mr = 0, oml = 0, omr = 0, omloo = 0, omroo = 0;
      for (uint32_t k = 0; k < ncol; ++k) {
        const uint32_t ka = mat[i + (size_t)k * nrow];
        if ((double)ka <= min_fold * (double)ja || ka < min_abund) continue;
        const size_t t = (size_t)slot[k];
        const uint32_t* e = lr.data() + 4 * t;
        const bool keep = e[0] + e[1] < sqlen;
        const uint32_t l = keep ? e[0] : 0, r = keep ? e[1] : 0;
        const uint32_t loo = keep ? e[2] : 0, roo = keep ? e[3] : 0;
        const bool allowed = p->allow_one_off && ham[t] >= (uint32_t)p->min_one_off_par_dist;
        ml = std::max(ml, l);
        mr = std::max(mr, r);
        if (p->allow_one_off && allowed) {
          oml = std::max(oml, l);
          omr = std::max(omr, r);
          omloo = std::max(omloo, loo);
          omroo = std::max(omroo, roo);
        }
      }
      if (ml + mr >= sqlen || (p->allow_one_off && (oml + omroo >= sqlen || omloo + omr >= sqlen)))
        ++nflag;
    }
    for (uint32_t k : need[j]) slot[k] = -1;
    t0 += (uint32_t)need[j].size();
    out_nflag[j] = nflag;
    out_nsam[j] = nsam;
  }
  return D2G_OK;
}

int d2g_pairs_load(d2g_ctx* c, uint32_t n, const uint8_t* seq_concat, const uint64_t* offsets,
                   const uint8_t* qual_concat, const uint32_t* subs, const uint64_t* sub_offsets,
                   const uint32_t* reads, const uint32_t* center_reads) {
  if (!c || !seq_concat || !offsets || !subs || !sub_offsets || !reads || !center_reads)
    return D2G_ERR_ARG;
  (void)hipSetDevice(c->device);
  c->p_n = 0;
  if (n == 0) return D2G_OK;
  // rows padded to 16 bytes (vector loads); validation as for load_uniques
  std::vector<uint64_t> off(n + 1, 0);
  std::vector<uint32_t> len(n);
  uint32_t maxq = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (offsets[i + 1] < offsets[i] || sub_offsets[i + 1] < sub_offsets[i]) return D2G_ERR_ARG;
    len[i] = (uint32_t)(offsets[i + 1] - offsets[i]);
    if (len[i] == 0 || len[i] > 65535 || reads[i] == 0) return D2G_ERR_ARG;
    off[i + 1] = off[i] + ((len[i] + 15) & ~15u);
    for (uint64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
      const uint8_t b = seq_concat[k];
      if (b < 1 || b > 4) return D2G_ERR_NONACGT;
      if (qual_concat) maxq = std::max<uint32_t>(maxq, qual_concat[k]);
    }
    uint32_t prev = 0;
    for (uint64_t k = sub_offsets[i]; k < sub_offsets[i + 1]; ++k) {
      const uint32_t pos = subs[k] >> 16, nt0 = subs[k] & 0xFFFFu;
      if (pos >= len[i] || nt0 < 1 || nt0 > 4) return D2G_ERR_ARG;
      if (k > sub_offsets[i] && pos <= prev) return D2G_ERR_ARG;  // ascending
      prev = pos;
    }
  }
  std::vector<uint8_t> sq(off[n], 1), ql;
  if (qual_concat) ql.assign(off[n], 0);
  for (uint32_t i = 0; i < n; ++i) {
    std::memcpy(sq.data() + off[i], seq_concat + offsets[i], len[i]);
    if (qual_concat) std::memcpy(ql.data() + off[i], qual_concat + offsets[i], len[i]);
  }
  const uint64_t ns = sub_offsets[n] - sub_offsets[0];
  // the kernels address both pools with 32-bit offsets
  if (off[n] + 16 >= ((uint64_t)1 << 32) || ns >= ((uint64_t)1 << 30)) return D2G_ERR_UNSUPPORTED;
  std::vector<uint64_t> so(n + 1);
  for (uint32_t i = 0; i <= n; ++i) so[i] = sub_offsets[i] - sub_offsets[0];
  hipStream_t s = c->stream;
  HIP_TRY(c->p_seq.ensure(off[n] + 16));
  if (qual_concat) HIP_TRY(c->p_qual.ensure(off[n] + 16));
  HIP_TRY(c->p_off.ensure(n + 1));
  HIP_TRY(c->p_len.ensure(n));
  HIP_TRY(c->p_subs.ensure(ns + 1));
  HIP_TRY(c->p_suboff.ensure(n + 1));
  HIP_TRY(c->p_reads.ensure(n));
  HIP_TRY(c->p_creads.ensure(n));
  HIP_TRY(c->p_loglam.ensure(n));
  HIP_TRY(c->p_lambda.ensure(n));
  HIP_TRY(c->p_pval.ensure(n));
  HIP_TRY(c->p_flag.ensure(n));
  HIP_TRY(c->p_nrec.ensure(1));
  HIP_TRY(c->p_list.ensure(n));
  // poisson_upper_tail_direct's log(reads!) as the reference's running sum
  // of logs (pval.rs:335), tabulated once in the same order
  uint32_t max_reads = 0;
  for (uint32_t i = 0; i < n; ++i) max_reads = std::max(max_reads, reads[i]);
  if (max_reads >= (1u << 24)) return D2G_ERR_UNSUPPORTED;
  std::vector<double> lkf((size_t)max_reads + 1);
  lkf[0] = 0.0;
  for (uint32_t i = 1; i <= max_reads; ++i) lkf[i] = lkf[i - 1] + std::log((double)i);
  HIP_TRY(c->p_lkf.ensure(lkf.size()));
  HIP_TRY(hipMemcpyAsync(c->p_lkf.p, lkf.data(), lkf.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_seq.p, sq.data(), off[n], hipMemcpyHostToDevice, s));
  if (qual_concat) HIP_TRY(hipMemcpyAsync(c->p_qual.p, ql.data(), off[n], hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_off.p, off.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_len.p, len.data(), n * 4, hipMemcpyHostToDevice, s));
  if (ns) HIP_TRY(hipMemcpyAsync(c->p_subs.p, subs + sub_offsets[0], ns * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_suboff.p, so.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_reads.p, reads, n * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->p_creads.p, center_reads, n * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  c->p_n = n;
  c->p_has_qual = qual_concat != nullptr;
  c->p_max_qual = maxq;
  return D2G_OK;
}

int d2g_pairs_lambda(d2g_ctx* c, const double* err, uint32_t ncol, int use_quals, double omega_a,
                     d2g_pair_stats* stats) {
  if (!c || !err || ncol == 0) return D2G_ERR_ARG;
  if (c->p_n == 0) return D2G_ERR_STATE;
  if (use_quals && c->p_has_qual && c->p_max_qual >= ncol) return D2G_ERR_ARG;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  HIP_TRY(c->p_err.ensure((size_t)16 * ncol));
  HIP_TRY(hipMemcpyAsync(c->p_err.p, err, (size_t)16 * ncol * 8, hipMemcpyHostToDevice, s));
  if (c->timing) HIP_TRY(hipEventRecord(c->ev[0], s));
  HIP_TRY(launch_zero(c->p_nrec.p, 1, s));
  PairArgs pa;
  pa.n = c->p_n;
  pa.seq = c->p_seq.p;
  pa.qual = c->p_has_qual ? c->p_qual.p : nullptr;
  pa.off = c->p_off.p;
  pa.len = c->p_len.p;
  pa.subs = c->p_subs.p;
  pa.sub_off = c->p_suboff.p;
  pa.reads = c->p_reads.p;
  pa.center_reads = c->p_creads.p;
  pa.err = c->p_err.p;
  pa.ncol = ncol;
  pa.use_quals = use_quals;
  pa.omega = omega_a;
  pa.out_loglam = c->p_loglam.p;
  pa.out_lambda = c->p_lambda.p;
  pa.out_p = c->p_pval.p;
  pa.out_flag = c->p_flag.p;
  pa.n_recomputed = c->p_nrec.p;
  pa.list = c->p_list.p;
  pa.log_kfact = c->p_lkf.p;
  pa.max_blocks = (uint32_t)c->cu_count * 8;
  pa.ev_mid = c->timing ? c->ev[1] : nullptr;
  HIP_TRY(launch_pairs_lambda(pa, s));
  if (c->timing) HIP_TRY(hipEventRecord(c->ev[2], s));
  uint32_t nrec = 0;
  HIP_TRY(hipMemcpyAsync(&nrec, c->p_nrec.p, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->pairs = c->p_n;
    stats->recomputed = nrec;
    if (c->timing) {
      stats->ms_loglam = ev_ms(c, 0, 1);
      stats->ms_pval = ev_ms(c, 1, 2);
      stats->ms_total = ev_ms(c, 0, 2);
    }
  }
  return D2G_OK;
}

int d2g_pairs_fetch(d2g_ctx* c, float* out_loglam32, double* out_lambda, double* out_pval,
                    uint8_t* out_recomputed) {
  if (!c) return D2G_ERR_ARG;
  if (c->p_n == 0) return D2G_ERR_STATE;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const uint32_t n = c->p_n;
  if (out_loglam32) HIP_TRY(hipMemcpyAsync(out_loglam32, c->p_loglam.p, n * 4, hipMemcpyDeviceToHost, s));
  if (out_lambda) HIP_TRY(hipMemcpyAsync(out_lambda, c->p_lambda.p, n * 8, hipMemcpyDeviceToHost, s));
  if (out_pval) HIP_TRY(hipMemcpyAsync(out_pval, c->p_pval.p, n * 8, hipMemcpyDeviceToHost, s));
  if (out_recomputed) HIP_TRY(hipMemcpyAsync(out_recomputed, c->p_flag.p, n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return D2G_OK;
}

}  // extern "C"

// Read counts of the loaded uniques in the caller's order (empty before a
// load); the partition loop's host state starts from them.
namespace d2g {
const std::vector<uint32_t>* ctx_host_reads(const d2g_ctx* c) {
  return (c && c->loaded) ? &c->h_reads : nullptr;
}

bool ctx_store(const d2g_ctx* c, const ShuffleComp** comps, uint64_t* count) {
  if (!c || !c->store_valid) return false;
  *comps = c->store;
  *count = c->store_n;
  return true;
}

bool ctx_lock_begin(d2g_ctx* c) {
  if (!c || !c->loaded) return false;
  if (c->lock.ensure(c->u.n) != hipSuccess) return false;
  if (hipMemsetAsync(c->lock.p, 0, c->u.n, c->stream) != hipSuccess) return false;
  c->lock_resident = true;
  return true;
}

bool ctx_lock_add(d2g_ctx* c, const uint32_t* raws, uint32_t k) {
  if (!c || !c->lock_resident) return false;
  if (k == 0) return true;
  // the staging area is free once the previous upload has run
  if (c->lock_ev_pending && hipEventSynchronize(c->lock_ev) != hipSuccess) return false;
  c->lock_ev_pending = false;
  if (!c->lock_ev && hipEventCreateWithFlags(&c->lock_ev, hipEventDisableTiming) != hipSuccess)
    return false;
  if (c->lock_pin_cap < k) {
    if (c->lock_pin) (void)hipHostFree(c->lock_pin);
    c->lock_pin = nullptr;
    c->lock_pin_cap = 0;
    const size_t cap = std::max<size_t>(k, 4096);
    if (hipHostMalloc((void**)&c->lock_pin, cap * sizeof(uint32_t), hipHostMallocDefault) !=
        hipSuccess)
      return false;
    c->lock_pin_cap = cap;
  }
  if (c->lock_idx.cap < k) {
    // a kernel still reading the old index list must finish first
    if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
    if (c->lock_idx.ensure(std::max<size_t>(k, 4096)) != hipSuccess) return false;
  }
  for (uint32_t j = 0; j < k; ++j) {
    if (raws[j] >= c->u.n) return false;
    c->lock_pin[j] = c->h_inv[raws[j]];
  }
  if (hipMemcpyAsync(c->lock_idx.p, c->lock_pin, (size_t)k * sizeof(uint32_t),
                     hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      launch_lock_set(c->lock.p, c->lock_idx.p, k, c->stream) != hipSuccess ||
      hipEventRecord(c->lock_ev, c->stream) != hipSuccess)
    return false;
  c->lock_ev_pending = true;
  return true;
}

void ctx_lock_end(d2g_ctx* c) {
  if (c) c->lock_resident = false;
}

bool ctx_exec(const d2g_ctx* c, CtxExec* out) {
  if (!c || !out) return false;
  out->device = c->device;
  out->stream = c->stream;
  out->cu_count = c->cu_count;
  out->timing = c->timing;
  for (int i = 0; i < 3; ++i) out->ev[i] = c->ev[i];
  return true;
}
}  // namespace d2g
