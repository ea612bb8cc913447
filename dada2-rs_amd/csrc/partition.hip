// partition.hip — run_dada's partition loop (dada.rs:749-1150) and the final
// per-raw p-value pass and read map (dada.rs:502-533), on the host, with the
// compare and the store of each new cluster center on the device
// (d2g_compare_store).
// Host code only; it links against the rest of the library through the C
// entry points.
//
// Semantics restated from the reference:
//   * B::new / init (containers.rs:337-391): one cluster holding every raw in
//     index order; e_minmax starts at -999.0 per raw and is never reset;
//     the center is the raw with the most reads, the last one among equals
//     (assign_center, containers.rs:450-463: Rust's max_by_key keeps the
//     last maximum).
//   * b_compare_parallel (cluster.rs:239-397): the device compares and
//     stores: a comparison is kept when lambda * B.reads > e_minmax[raw],
//     e_minmax (resident on the device) rises to lambda * center_reads, and
//     only the kept comparisons come back, in raw order; Raw.comp follows
//     them for cluster 0 and for the center itself.  Cluster 0's compare
//     uses kdist_cutoff 1.0 (dada.rs:835-838).
//   * b_p_update (pval.rs:45-130): clusters flagged update_e reprice their
//     members with get_pA (pval.rs:376-398); with greedy, a cluster's first
//     p-update locks the members its center explains and the center.
//   * b_bud (cluster.rs:1331-1480): the smallest p among non-first members
//     (reads as tie-break) over all clusters, seeded with cluster 0's
//     center; a Bonferroni test against omega_a, then the prior test against
//     omega_p.  b_bud_incremental (cluster.rs:1501) is its cached equivalent.
//   * b_shuffle_converge (cluster.rs:908) is byte-identical to repeating
//     b_shuffle2 (cluster.rs:615-700) until a pass moves nothing, with one
//     difference in counting: with the carried map (on by default,
//     cluster.rs:79-82, 957) every call after the first spends its first
//     MAX_SHUFFLE iteration on a reconcile (cluster.rs:1037, 1066-1073), so
//     the first call makes up to max_shuffle move passes and later calls up
//     to max_shuffle - 1.
//   * calc_pA (pval.rs:208-241): statrs's Poisson survival function is
//     gamma_lr(reads, e), restated from statrs 0.18 (d2g_statrs.h), with the
//     reference's own small-lambda series where it returns 0
//     (pval.rs:329-353).
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/dada2_amd.h"
#include "d2g_launch.h"
#include "d2g_statrs.h"

namespace d2g {
const std::vector<uint32_t>* ctx_host_reads(const d2g_ctx* c);
}

namespace {

// A persistent pool for the loop's per-raw work (p-values, the shuffle's
// best-cluster scan): run() hands out chunks of [0, n) to the workers and
// the caller and returns when every chunk is done.  Results are written by
// index, so they do not depend on the thread count.
class Pool {
 public:
  explicit Pool(int nthreads) {
    for (int t = 1; t < nthreads; ++t) th_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return (int)th_.size() + 1; }
  void run(size_t n, size_t chunk, const std::function<void(size_t, size_t)>& fn) {
    if (th_.empty() || n <= chunk) {
      if (n) fn(0, n);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      job_ = &fn;
      n_ = n;
      chunk_ = chunk;
      next_.store(0);
      busy_ = (int)th_.size();
      ++gen_;
    }
    cv_.notify_all();
    drain(fn);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void drain(const std::function<void(size_t, size_t)>& fn) {
    for (;;) {
      const size_t b = next_.fetch_add(chunk_);
      if (b >= n_) break;
      fn(b, std::min(n_, b + chunk_));
    }
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(size_t, size_t)>* job;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      drain(*job);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--busy_ == 0) done_.notify_one();
      }
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t, size_t)>* job_ = nullptr;
  size_t n_ = 0, chunk_ = 1;
  std::atomic<size_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Host threads for the loop: the CPUs this process may run on, at most 16.
int pool_threads() {
  cpu_set_t set;
  int n = 1;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  return std::max(1, std::min(n, 16));
}

// Comparison, containers.rs:15-20.
struct Comp {
  uint32_t i = 0, index = 0;
  double lambda = 0.0;
  uint32_t hamming = 0;
};

// One stored comparison of a raw, in the raw's candidate list (the
// CandIndex of dada.rs:895-897): clusters append in creation order, so a
// list is in ascending cluster order.
struct Cand {
  uint32_t i;
  uint32_t hamming;
  double lambda;
};

// The Raw fields the loop reads and writes (containers.rs:78-108).
struct Raw {
  uint32_t reads = 0;
  bool prior = false, lock = false, correct = false;
  double p = 0.0;
  Comp comp;
};

enum : int32_t { kBirthInitial = 0, kBirthAbundance = 1, kBirthPrior = 2 };

// Bi, containers.rs:200-265.
struct Cluster {
  std::vector<uint32_t> raws;
  uint32_t reads = 0;
  int64_t center = -1;
  bool update_e = true, check_locks = true;
  double self = 0.0;
  int32_t birth_type = kBirthInitial;
  uint32_t birth_from = 0;
  double birth_pval = 0.0, birth_fold = 1.0, birth_e = 0.0;
  Comp birth_comp;
  std::vector<Comp> comp;
};

// B, containers.rs:268-337.
struct Part {
  std::vector<Raw> raws;
  std::vector<Cluster> cl;
  uint32_t reads = 0;  // e_minmax lives on the device (d2g_compare_store)
  uint64_t nalign = 0, nshroud = 0;
  // each raw's cluster and position in it (the raw_cluster map of
  // containers.rs:302-316, plus the position)
  std::vector<uint32_t> where, at;

  // bi_add_raw, containers.rs:403-417
  void add_raw(size_t ci, uint32_t r) {
    where[r] = (uint32_t)ci;
    at[r] = (uint32_t)cl[ci].raws.size();
    cl[ci].raws.push_back(r);
    cl[ci].reads += raws[r].reads;
    cl[ci].update_e = true;
  }
  // bi_pop_raw, containers.rs:428-439: swap-remove
  uint32_t pop_raw(size_t ci, size_t pos) {
    std::vector<uint32_t>& v = cl[ci].raws;
    const uint32_t r = v[pos];
    v[pos] = v.back();
    at[v[pos]] = (uint32_t)pos;
    v.pop_back();
    cl[ci].reads -= raws[r].reads;
    cl[ci].update_e = true;
    return r;
  }
  // assign_center, containers.rs:450-463 (last raw among the most reads)
  void assign_center(size_t ci) {
    int64_t best = -1;
    for (uint32_t r : cl[ci].raws)
      if (best < 0 || raws[r].reads >= raws[(size_t)best].reads) best = r;
    if (best >= 0) cl[ci].center = best;
  }
};

// ---- p-values (pval.rs) ----------------------------------------------------

// ln(k!) as poisson_upper_tail_direct sums it (ln 1 + ln 2 + ... + ln k,
// left to right), tabulated once per run up to the run's largest read
// count (at most kMax entries; larger k continue the sum from the table's
// end in the same order), so every entry has the loop's bits.
class LogFact {
 public:
  static constexpr uint32_t kMax = 1u << 22;
  void build(uint32_t rmax) {
    const uint32_t m = std::min(rmax, kMax - 1u);
    t_.resize((size_t)m + 1);
    double s = 0.0;
    t_[0] = 0.0;
    for (uint32_t i = 1; i <= m; ++i) {
      s += std::log((double)i);
      t_[i] = s;
    }
  }
  double operator()(uint32_t k) const {
    const uint32_t m = (uint32_t)t_.size() - 1u;
    if (k <= m) return t_[k];
    double s = t_[m];
    for (uint32_t i = m + 1; i <= k; ++i) s += std::log((double)i);
    return s;
  }

 private:
  std::vector<double> t_{0.0};
};

// poisson_upper_tail_direct, pval.rs:329-353
double poisson_tail_direct(uint32_t reads, double lambda, const LogFact& lf) {
  const double k = (double)reads;
  const double log_lambda = std::log(lambda);
  const double log_kfact = lf(reads);
  const double log_leading = -lambda + k * log_lambda - log_kfact;
  double series = 1.0, term = 1.0;
  for (uint32_t j = 1; j < 10000u; ++j) {
    term *= lambda / (k + (double)j);
    if (!std::isfinite(term) || term <= 0.0) break;
    const double ns = series + term;
    if (ns == series) break;
    series = ns;
  }
  return std::exp(log_leading + std::log(series));
}

// calc_pA, pval.rs:208-241
double calc_pA(uint32_t reads, double e_reads, bool prior, const LogFact& lf) {
  if (!(e_reads > 0.0)) return 1.0;  // Poisson::new rejects e_reads <= 0 and NaN
  double pval = d2g::statrs_gamma_lr((double)reads, e_reads);  // sf(reads - 1)
  if (pval == 0.0) pval = poisson_tail_direct(reads, e_reads, lf);
  if (prior) return pval;
  double norm = 1.0 - std::exp(-e_reads);
  if (norm < 1e-7) norm = e_reads - 0.5 * e_reads * e_reads;  // TAIL_APPROX_CUTOFF, pval.rs:18
  return pval / norm;
}

// get_pA, pval.rs:376-398
double get_pA(uint32_t reads, bool prior, double lambda, uint32_t hamming, uint32_t bi_reads,
              bool detect_singletons, const LogFact& lf) {
  if (reads == 1 && !prior && !detect_singletons) return 1.0;
  if (hamming == 0) return 1.0;
  if (lambda == 0.0) return 0.0;
  return calc_pA(reads, lambda * (double)bi_reads, prior || detect_singletons, lf);
}

// ---- the loop's steps ------------------------------------------------------

// The shuffle scan's state on the device for one run: the stored
// comparisons (appended after each compare), each raw's cluster, and the
// scan's scratch; pinned staging for the per-compare and per-pass copies
// (one copy each way per pass, one per compare).
struct DevShuffle {
  hipStream_t s = nullptr;
  const d2g_ctx* ctx = nullptr;  // its store holds the kept comparisons
  uint32_t n = 0, ncl_cap = 0;
  uint8_t* cl_info = nullptr;  // creads (f64 x ncl_cap), then ccenter (u32)
  unsigned long long *best_bits = nullptr, *best_key = nullptr;
  uint32_t* where = nullptr;
  d2g::ShuffleMove* moves = nullptr;  // slot 0: the count
  uint8_t* pass_pin = nullptr;
  d2g::ShuffleMove* moves_pin = nullptr;
  static constexpr uint32_t kFirst = 1024;  // moves read back with the count

  ~DevShuffle() {
    if (s) (void)hipStreamSynchronize(s);
    for (void* q : {(void*)cl_info, (void*)best_bits, (void*)best_key, (void*)where, (void*)moves})
      if (q) (void)hipFree(q);
    for (void* q : {(void*)pass_pin, (void*)moves_pin})
      if (q) (void)hipHostFree(q);
  }

  static bool ok(hipError_t e) { return e == hipSuccess; }

  bool init(const d2g_ctx* c, hipStream_t stream, uint32_t nraw) {
    ctx = c;
    s = stream;
    n = nraw;
    if (!ok(hipMalloc((void**)&best_bits, (size_t)n * 8)) ||
        !ok(hipMalloc((void**)&best_key, (size_t)n * 8)) ||
        !ok(hipMalloc((void**)&where, (size_t)n * 4)) ||
        !ok(hipMalloc((void**)&moves, (size_t)(n + 1) * sizeof(d2g::ShuffleMove))) ||
        !ok(hipHostMalloc((void**)&moves_pin, (kFirst + 1) * sizeof(d2g::ShuffleMove),
                          hipHostMallocDefault)))
      return false;
    return ok(d2g::launch_shuffle_init(args(0, nullptr, 0), s));
  }

  d2g::ShuffleArgs args(uint32_t ncl, const d2g::ShuffleComp* comps, uint64_t n_comp) const {
    d2g::ShuffleArgs a;
    a.comps = comps;
    a.n_comp = (uint32_t)n_comp;  // d2g_compare_store keeps the store below 2^32 records
    a.creads = reinterpret_cast<const double*>(cl_info);
    a.ccenter = cl_info ? reinterpret_cast<const uint32_t*>(cl_info + (size_t)ncl_cap * 8) : nullptr;
    a.n_cl = ncl;
    a.n_raw = n;
    a.best_bits = best_bits;
    a.best_key = best_key;
    a.where = where;
    a.moves = moves;
    return a;
  }

  // One pass: the scan on the device, the moves read back (unordered).
  bool pass(const Part& B, std::vector<d2g::ShuffleMove>& out) {
    const uint32_t ncl = (uint32_t)B.cl.size();
    if (ncl > ncl_cap) {
      const uint32_t nc = std::max(ncl, 2 * ncl_cap + 64);
      if (cl_info) (void)hipFree(cl_info);
      if (pass_pin) (void)hipHostFree(pass_pin);
      cl_info = nullptr;
      pass_pin = nullptr;
      if (!ok(hipMalloc((void**)&cl_info, (size_t)nc * 12)) ||
          !ok(hipHostMalloc((void**)&pass_pin, (size_t)nc * 12, hipHostMallocDefault)))
        return false;
      ncl_cap = nc;
    }
    double* hr = reinterpret_cast<double*>(pass_pin);
    uint32_t* hc = reinterpret_cast<uint32_t*>(pass_pin + (size_t)ncl_cap * 8);
    for (uint32_t ci = 0; ci < ncl; ++ci) {
      hr[ci] = (double)B.cl[ci].reads;
      hc[ci] = B.cl[ci].center >= 0 ? (uint32_t)B.cl[ci].center : 0xFFFFFFFFu;
    }
    const uint32_t first = std::min(n, kFirst) + 1;
    const d2g::ShuffleComp* comps = nullptr;
    uint64_t n_comp = 0;
    if (!d2g::ctx_store(ctx, &comps, &n_comp)) return false;
    if (!ok(hipMemcpyAsync(cl_info, pass_pin, (size_t)ncl_cap * 8 + (size_t)ncl * 4,
                           hipMemcpyHostToDevice, s)) ||
        !ok(d2g::launch_shuffle_pass(args(ncl, comps, n_comp), s)) ||
        !ok(hipMemcpyAsync(moves_pin, moves, (size_t)first * sizeof(d2g::ShuffleMove),
                           hipMemcpyDeviceToHost, s)) ||
        !ok(hipStreamSynchronize(s)))
      return false;
    const uint32_t k = moves_pin[0].r;
    out.assign(moves_pin + 1, moves_pin + 1 + std::min(k, kFirst));
    if (k > kFirst) {
      out.resize(k);
      if (!ok(hipMemcpy(out.data() + kFirst, moves + 1 + kFirst,
                        (size_t)(k - kFirst) * sizeof(d2g::ShuffleMove), hipMemcpyDeviceToHost)))
        return false;
    }
    return true;
  }
};

// A pinned host array: the compare's results are copied straight into it
// (a pageable destination goes through a staging copy).
template <class T>
struct Pinned {
  T* p = nullptr;
  size_t n = 0;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  bool resize(size_t k) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    if (hipHostMalloc((void**)&p, std::max<size_t>(k, 1) * sizeof(T), hipHostMallocDefault) !=
        hipSuccess)
      return false;
    n = k;
    return true;
  }
  T* data() { return p; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
};

struct Run {
  d2g_ctx* ctx;
  const double* err;
  uint32_t ncol;
  d2g_align_params align;
  d2g_dada_params dp;
  Part B;
  Pinned<d2g_comparison> stored;  // one compare's kept comparisons (at most one per raw)
  // raws locked since the last compare (the device's lock flags follow them)
  std::vector<uint32_t> new_locks;
  std::vector<std::vector<Cand>> cand;  // per raw
  Pool* pool = nullptr;
  // A bud candidate; `better` is the scan's replacement test (smaller p,
  // then more reads).
  struct Best {
    double p;
    uint32_t reads;
    size_t ci, pos;
    bool have;
  };
  static bool better(double p, uint32_t reads, const Best& b) {
    return p < b.p || (p == b.p && reads > b.reads);
  }
  // a p-update's work: members [b, e) of cluster ci
  struct Piece {
    uint32_t ci, b, e;
  };
  std::vector<Piece> pieces;
  std::vector<Best> piece_best;  // per piece: abundance and prior candidates
  std::vector<Best> cl_best;     // per cluster: the same, kept between p-updates
  LogFact lf;                    // ln(k!) up to the largest read count
  d2g_dada_stats st{};
  DevShuffle* dev = nullptr;  // the device scan, when asked for
  bool dev_failed = false;

  static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // b_compare_parallel, cluster.rs:239-397: the map and the store on the
  // device (d2g_compare_store); the kept comparisons come back in raw order
  int compare(size_t i) {
    const uint32_t center = (uint32_t)B.cl[i].center;
    d2g_align_params p = align;
    if (i == 0) p.kdist_cutoff = 1.0;  // dada.rs:835-838
    auto t0 = std::chrono::steady_clock::now();
    // the lock flags live on the device: only the raws locked since the
    // last compare go up
    if (!d2g::ctx_lock_add(ctx, new_locks.data(), (uint32_t)new_locks.size())) return D2G_ERR_HIP;
    new_locks.clear();
    d2g_stats cs{};
    uint64_t nst = 0;
    double self = 0.0;
    const int rc = d2g_compare_store(ctx, &center, 1, (uint32_t)i, (double)B.reads, err, ncol, &p,
                                     dp.use_quals, dp.greedy, nullptr, stored.n, stored.data(),
                                     &nst, &self, &cs);
    st.ms_compare += ms_since(t0);
    ++st.compares;
    if (rc != D2G_OK) return rc;
    st.dp_cells += cs.dp_cells;
    st.ms_dev_screen += cs.ms_screen;
    st.ms_dev_plan += cs.ms_plan;
    st.ms_dev_dp += cs.ms_dp;
    st.ms_dev_lambda += cs.ms_lambda;
    t0 = std::chrono::steady_clock::now();
    B.nalign += cs.comparisons - cs.skipped;
    B.nshroud += cs.shrouded;
    Cluster& C = B.cl[i];
    C.self = self;
    C.comp.resize((size_t)nst);
    for (size_t k = 0; k < (size_t)nst; ++k) {
      const d2g_comparison& d = stored[k];
      Comp& c = C.comp[k];
      c.i = d.cluster;
      c.index = d.index;
      c.lambda = d.lambda;
      c.hamming = d.hamming;
      if (!dev) cand[d.index].push_back(Cand{c.i, c.hamming, c.lambda});
      if (i == 0 || d.index == center) B.raws[d.index].comp = c;
    }
    st.ms_store += ms_since(t0);
    return D2G_OK;
  }

  // b_p_update, pval.rs:45-130, fused with the candidate scan of b_bud.
  //   * Each member's p depends only on its own comparison and its
  //     cluster's reads, so the members of every cluster flagged update_e
  //     are repriced in one pass over the pool, in pieces of at most kPiece
  //     members of one cluster (most clusters are far smaller).
  //   * The same pass keeps each repriced cluster's bud candidates (see
  //     bud()): a cluster's members, their order, p and comparisons and its
  //     reads change only through add_raw / pop_raw, which flag update_e,
  //     so a cluster that is not repriced keeps its candidates.
  //   * Locks are only ever set; the newly set ones are queued for the
  //     device.
  void p_update() {
    constexpr uint32_t kPiece = 2048;
    const auto t0 = std::chrono::steady_clock::now();
    pieces.clear();
    cl_best.resize(2 * B.cl.size(), Best{0.0, 0u, 0, 0, false});
    for (size_t ci = 0; ci < B.cl.size(); ++ci) {
      Cluster& C = B.cl[ci];
      if (C.update_e) {
        const uint32_t m = (uint32_t)C.raws.size();
        if (m == 0) cl_best[2 * ci] = cl_best[2 * ci + 1] = Best{0.0, 0u, 0, 0, false};
        for (uint32_t b = 0; b < m; b += kPiece)
          pieces.push_back(Piece{(uint32_t)ci, b, std::min(m, b + kPiece)});
        C.update_e = false;
      }
      if (dp.greedy && C.check_locks) {
        const uint32_t center_reads = C.center >= 0 ? B.raws[(size_t)C.center].reads : 0u;
        for (uint32_t r : C.raws) {
          Raw& R = B.raws[r];
          const bool lk = (double)center_reads * R.comp.lambda > (double)R.reads ||
                          (int64_t)r == C.center;
          if (lk && !R.lock) {
            R.lock = true;
            new_locks.push_back(r);
          }
        }
        C.check_locks = false;
      }
    }
    const bool ds = dp.detect_singletons != 0;
    piece_best.resize(2 * pieces.size());
    pool->run(pieces.size(), 1, [&](size_t b, size_t e) {
      for (size_t k = b; k < e; ++k) {
        const Piece& pc = pieces[k];
        const Cluster& C = B.cl[pc.ci];
        Best ba{0.0, 0u, 0, 0, false}, bp{0.0, 0u, 0, 0, false};
        for (uint32_t pos = pc.b; pos < pc.e; ++pos) {
          Raw& R = B.raws[C.raws[pos]];
          R.p = get_pA(R.reads, R.prior, R.comp.lambda, R.comp.hamming, C.reads, ds, lf);
          if (pos == 0) continue;  // the first member is never a candidate
          if (R.reads < dp.min_abund) continue;
          if (R.comp.hamming < dp.min_hamming) continue;
          const bool fold_ok = dp.min_fold <= 1.0 ||
                               (double)R.reads >= dp.min_fold * R.comp.lambda * (double)C.reads;
          if (!fold_ok) continue;
          if (!ba.have || better(R.p, R.reads, ba)) ba = Best{R.p, R.reads, pc.ci, pos, true};
          if (R.prior && (!bp.have || better(R.p, R.reads, bp)))
            bp = Best{R.p, R.reads, pc.ci, pos, true};
        }
        piece_best[2 * k] = ba;
        piece_best[2 * k + 1] = bp;
      }
    });
    // each repriced cluster's candidates: its pieces folded in order
    for (size_t k = 0; k < pieces.size(); ++k) {
      const size_t ci = pieces[k].ci;
      Best& ca = cl_best[2 * ci];
      Best& cp = cl_best[2 * ci + 1];
      if (pieces[k].b == 0) ca = cp = Best{0.0, 0u, 0, 0, false};
      const Best& a = piece_best[2 * k];
      const Best& q = piece_best[2 * k + 1];
      if (a.have && (!ca.have || better(a.p, a.reads, ca))) ca = a;
      if (q.have && (!cp.have || better(q.p, q.reads, cp))) cp = q;
    }
    st.ms_pupdate += ms_since(t0);
  }

  // The tail of b_bud (cluster.rs:1419-1449): pop raw r of cluster ci into a
  // new cluster born with `type`.
  size_t bud_into(size_t ci, size_t pos, uint32_t r, int32_t type, double pval) {
    const double expected = B.raws[r].comp.lambda * (double)B.cl[ci].reads;
    Cluster nb;
    nb.birth_type = type;
    nb.birth_from = type == kBirthAbundance ? (uint32_t)ci : 0u;
    nb.birth_pval = pval;
    nb.birth_fold = (double)B.raws[r].reads / std::max(expected, DBL_MIN);
    nb.birth_e = expected;
    nb.birth_comp = B.raws[r].comp;
    B.pop_raw(ci, pos);
    B.cl.push_back(std::move(nb));
    const size_t ni = B.cl.size() - 1;
    B.add_raw(ni, r);
    B.assign_center(ni);
    return ni;
  }

  // b_bud, cluster.rs:1331-1480; returns the new cluster or -1.
  // The scan visits every cluster's members at positions >= 1 in order and
  // keeps the first candidate of the smallest key (p, then more reads).
  // That is an associative fold, so it runs over each cluster's kept
  // candidates (p_update) in cluster order and gives the whole scan's pick.
  int64_t bud() {
    const auto t0 = std::chrono::steady_clock::now();
    const double nraw = (double)B.raws.size();
    const uint32_t c0 = (uint32_t)B.cl[0].center;
    // seeded with cluster 0's center, which is never chosen itself
    Best ma{B.raws[c0].p, B.raws[c0].reads, 0, 0, false}, mp = ma;
    for (size_t ci = 0; ci < B.cl.size(); ++ci) {
      const Best& a = cl_best[2 * ci];
      const Best& q = cl_best[2 * ci + 1];
      if (a.have && better(a.p, a.reads, ma)) ma = a;
      if (q.have && better(q.p, q.reads, mp)) mp = q;
    }
    const double p_a = ma.p * nraw;
    int64_t out = -1;
    if (p_a < dp.omega_a && ma.have)
      out = (int64_t)bud_into(ma.ci, ma.pos, B.cl[ma.ci].raws[ma.pos], kBirthAbundance, p_a);
    else if (mp.p < dp.omega_p && mp.have)
      out = (int64_t)bud_into(mp.ci, mp.pos, B.cl[mp.ci].raws[mp.pos], kBirthPrior, mp.p);
    st.ms_bud += ms_since(t0);
    return out;
  }

  // b_shuffle2, cluster.rs:615-700: one move pass; returns the moves.
  //   * The best-cluster scan runs per raw over its candidate list
  //     (ascending cluster order, strict >: the lowest cluster keeps ties),
  //     which is b_shuffle2's cluster-major scan regrouped by raw.
  //   * The move pass walks every cluster from its end and swap-removes the
  //     raws whose best cluster is another (centers stay).  Raws appended
  //     during the pass never move again, and a swap-remove only disturbs
  //     positions already walked, so visiting just the movers, by cluster
  //     and then by descending position, makes the same moves in the same
  //     order.
  size_t shuffle_pass(std::vector<Comp>& compmax, std::vector<double>& creads,
                      std::vector<std::vector<uint64_t>>& found) {
    constexpr size_t kChunk = 2048;
    const size_t n = B.raws.size();
    compmax.resize(n);
    creads.resize(B.cl.size());
    for (size_t ci = 0; ci < B.cl.size(); ++ci) creads[ci] = (double)B.cl[ci].reads;
    found.resize((n + kChunk - 1) / kChunk);
    pool->run(n, kChunk, [&](size_t b, size_t e) {
      std::vector<uint64_t>& mv = found[b / kChunk];
      mv.clear();
      for (size_t r = b; r < e; ++r) {
        double emax = -INFINITY;
        Comp best;
        for (const Cand& c : cand[r]) {
          const double x = c.lambda * creads[c.i];
          if (x > emax) {
            emax = x;
            best.i = c.i;
            best.index = (uint32_t)r;
            best.lambda = c.lambda;
            best.hamming = c.hamming;
          }
        }
        compmax[r] = best;
        const uint32_t ci = B.where[r];
        if (best.i != ci && B.cl[ci].center != (int64_t)r)
          mv.push_back((uint64_t)ci << 32 | (uint64_t)(0xFFFFFFFFu - B.at[r]));
      }
    });
    std::vector<uint64_t> movers;
    for (const auto& f : found) movers.insert(movers.end(), f.begin(), f.end());
    std::sort(movers.begin(), movers.end());  // cluster ascending, position descending
    for (uint64_t key : movers) {
      const size_t ci = (size_t)(key >> 32);
      const size_t pos = 0xFFFFFFFFu - (uint32_t)key;
      const uint32_t r = B.cl[ci].raws[pos];
      B.pop_raw(ci, pos);
      B.add_raw(compmax[r].i, r);
      B.raws[r].comp = compmax[r];
    }
    return movers.size();
  }

  // The same pass with the scan on the device: the movers come back
  // unordered and are applied in the host pass's order.
  size_t shuffle_pass_dev(std::vector<d2g::ShuffleMove>& mv, std::vector<uint64_t>& keys) {
    if (!dev->pass(B, mv)) {
      dev_failed = true;
      return 0;
    }
    keys.resize(mv.size());
    for (size_t j = 0; j < mv.size(); ++j) {
      const uint32_t r = mv[j].r;
      // cluster ascending, position descending; the index rides in the low bits
      keys[j] = ((uint64_t)B.where[r] << 32 | (uint64_t)(0xFFFFFFFFu - B.at[r]));
    }
    std::vector<size_t> ord(mv.size());
    for (size_t j = 0; j < ord.size(); ++j) ord[j] = j;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return keys[x] < keys[y]; });
    for (size_t j : ord) {
      const d2g::ShuffleMove& m = mv[j];
      const size_t ci = (size_t)(keys[j] >> 32);
      const size_t pos = 0xFFFFFFFFu - (uint32_t)keys[j];
      B.pop_raw(ci, pos);
      B.add_raw(m.to, m.r);
      Comp c;
      c.i = m.to;
      c.index = m.r;
      c.lambda = m.lambda;
      c.hamming = m.ham;
      B.raws[m.r].comp = c;
    }
    return mv.size();
  }

  // b_shuffle_converge, cluster.rs:908-1320 (see the counting note above)
  void shuffle(bool first) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t passes = first ? dp.max_shuffle : std::max(dp.max_shuffle, 2u) - 1u;
    std::vector<Comp> compmax;
    std::vector<double> creads;
    std::vector<std::vector<uint64_t>> found;
    std::vector<d2g::ShuffleMove> mv;
    std::vector<uint64_t> keys;
    for (uint32_t k = 0; k < passes && !dev_failed; ++k) {
      ++st.shuffle_passes;
      const size_t moved = dev ? shuffle_pass_dev(mv, keys) : shuffle_pass(compmax, creads, found);
      if (moved == 0) break;
    }
    st.ms_shuffle += ms_since(t0);
  }
};

}  // namespace

extern "C" int d2g_run_dada(d2g_ctx* ctx, const double* err, uint32_t ncol,
                            const d2g_align_params* align, const d2g_dada_params* dada,
                            const uint8_t* prior, uint32_t cluster_capacity,
                            uint32_t* out_nclusters, d2g_cluster_info* out_clusters,
                            uint32_t* out_cluster_of, double* out_p, int32_t* out_map,
                            double* out_comp_lambda, uint32_t* out_comp_hamming,
                            d2g_dada_stats* stats) {
  if (!ctx || !err || !align || !dada || !out_nclusters || !out_clusters || !out_cluster_of ||
      !out_p || !out_map)
    return D2G_ERR_ARG;
  if (dada->max_shuffle < 1 || ncol == 0) return D2G_ERR_ARG;
  const std::vector<uint32_t>* reads = d2g::ctx_host_reads(ctx);
  if (!reads) return D2G_ERR_STATE;
  const auto t_start = std::chrono::steady_clock::now();
  Run run;
  run.ctx = ctx;
  run.err = err;
  run.ncol = ncol;
  run.align = *align;
  run.dp = *dada;
  Part& B = run.B;
  const uint32_t n = (uint32_t)reads->size();
  B.raws.resize(n);
  for (uint32_t r = 0; r < n; ++r) {
    B.raws[r].reads = (*reads)[r];
    B.raws[r].prior = prior ? prior[r] != 0 : false;
    B.reads += (*reads)[r];
  }
  {
    uint32_t rmax = 0;
    for (uint32_t r = 0; r < n; ++r) rmax = std::max(rmax, (*reads)[r]);
    run.lf.build(rmax);
  }
  B.where.assign(n, 0);
  B.at.resize(n);
  for (uint32_t r = 0; r < n; ++r) B.at[r] = r;
  if (!run.stored.resize(n)) return D2G_ERR_HIP;
  int rc = d2g_store_reset(ctx);  // B::new's e_minmax (containers.rs:342)
  if (rc != D2G_OK) return rc;
  // every raw starts unlocked; the flags stay on the device for the run
  struct LockScope {
    d2g_ctx* c;
    ~LockScope() { d2g::ctx_lock_end(c); }
  } lock_scope{ctx};
  if (!d2g::ctx_lock_begin(ctx)) return D2G_ERR_HIP;
  run.cand.resize(n);
  Pool pool(pool_threads());
  run.pool = &pool;
  DevShuffle dev;
  if (dada->device_shuffle) {
    if (!dev.init(ctx, (hipStream_t)d2g_ctx_stream(ctx), n)) return D2G_ERR_HIP;
    run.dev = &dev;
  }
  // init, containers.rs:363-391
  {
    Cluster c0;
    c0.birth_type = kBirthInitial;
    c0.birth_e = (double)B.reads;
    c0.raws.reserve(n);
    B.cl.push_back(std::move(c0));
    for (uint32_t r = 0; r < n; ++r) {
      B.cl[0].raws.push_back(r);
      B.cl[0].reads += B.raws[r].reads;
    }
    B.cl[0].update_e = true;
    B.assign_center(0);
  }
  // run_dada, dada.rs:822-1000
  rc = run.compare(0);
  if (rc != D2G_OK) return rc;
  run.p_update();
  const size_t max_clust = dada->max_clust == 0 ? (size_t)n : (size_t)dada->max_clust;
  bool first_shuffle = true;
  while (B.cl.size() < max_clust) {
    const int64_t ni = run.bud();
    if (ni < 0) break;
    if ((rc = run.compare((size_t)ni)) != D2G_OK) return rc;
    run.shuffle(first_shuffle);
    if (run.dev_failed) return D2G_ERR_HIP;
    first_shuffle = false;
    run.p_update();
  }
  // final p pass and map, dada.rs:502-533
  for (size_t ci = 0; ci < B.cl.size(); ++ci) {
    const Cluster& C = B.cl[ci];
    for (uint32_t r : C.raws) {
      Raw& R = B.raws[r];
      if ((int64_t)r == C.center) {
        R.p = 1.0;
        R.correct = true;
      } else {
        R.p = calc_pA(R.reads, R.comp.lambda * (double)C.reads, true, run.lf);
        R.correct = R.p >= dada->omega_c;
      }
    }
  }
  // post_hoc_pvals (error.rs:145-180): per center, the reads the other
  // clusters' stored comparisons expect of it, summed over clusters in
  // order and each cluster's comparisons in stored order
  std::vector<int64_t> cluster_at(B.raws.size(), -1);
  for (size_t ci = 0; ci < B.cl.size(); ++ci)
    if (B.cl[ci].center >= 0) cluster_at[(size_t)B.cl[ci].center] = (int64_t)ci;
  std::vector<double> tot_e(B.cl.size(), 0.0);
  for (size_t ci = 0; ci < B.cl.size(); ++ci)
    for (const Comp& cp : B.cl[ci].comp) {
      const int64_t j = cluster_at[cp.index];
      if (j >= 0 && (size_t)j != ci) tot_e[(size_t)j] += cp.lambda * (double)B.cl[ci].reads;
    }
  *out_nclusters = (uint32_t)B.cl.size();
  if (B.cl.size() > cluster_capacity) return D2G_ERR_ARG;
  for (size_t ci = 0; ci < B.cl.size(); ++ci) {
    const Cluster& C = B.cl[ci];
    d2g_cluster_info& o = out_clusters[ci];
    o.center = (uint32_t)C.center;
    o.reads = C.reads;
    o.nraw = (uint32_t)C.raws.size();
    o.birth_type = C.birth_type;
    o.birth_from = C.birth_from;
    o.birth_hamming = C.birth_comp.hamming;
    o.birth_pval = C.birth_pval;
    o.birth_fold = C.birth_fold;
    o.birth_e = C.birth_e;
    o.birth_lambda = C.birth_comp.lambda;
    o.self_lambda = C.self;
    o.posthoc_pval =
        C.center >= 0 ? calc_pA(B.raws[(size_t)C.center].reads, tot_e[ci], true, run.lf) : 1.0;
    for (uint32_t r : C.raws) {
      out_cluster_of[r] = (uint32_t)ci;
      out_map[r] = B.raws[r].correct ? (int32_t)ci : -1;
    }
  }
  for (uint32_t r = 0; r < n; ++r) {
    out_p[r] = B.raws[r].p;
    if (out_comp_lambda) out_comp_lambda[r] = B.raws[r].comp.lambda;
    if (out_comp_hamming) out_comp_hamming[r] = B.raws[r].comp.hamming;
  }
  if (stats) {
    run.st.nalign = B.nalign;
    run.st.nshroud = B.nshroud;
    run.st.ms_total = Run::ms_since(t_start);
    *stats = run.st;
  }
  return D2G_OK;
}
