// noalloc.hip — the allocation-free entry point d2g_pairs_lambda_noalloc.
//
// It computes what d2g_pairs_load + d2g_pairs_lambda + d2g_pairs_fetch
// compute (FP32 Kahan log-lambda, FP64 p-values, the FP64 recompute of the
// pairs near omega_a), with the same kernels, for callers that cannot call
// hipMalloc: the pair set and the outputs are the caller's device memory,
// and every intermediate lives in `scratch`, one fixed array per device:
//   [0, 256)        control words: validation flags, recompute counter
//   err             16 x ncol f64 (ncol <= D2G_NOALLOC_MAX_NCOL), copied per call
//   log_kfact       sum_{i=1..r} log(i) for r = 0 .. D2G_NOALLOC_MAX_READS,
//                   written once per device at set-up in the running order of
//                   d2g_pairs_load's table (the same bits)
//   list            D2G_NOALLOC_MAX_PAIRS u32: the pairs near omega_a
// Host code never passes `scratch` itself to a kernel or a copy: it uses the
// address hipGetSymbolAddress returned for the current device at set-up.
// Calls that overlap on one device share the array and race.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/dada2_amd.h"
#include "d2g_launch.h"

namespace d2g {

constexpr size_t kNaCtlBytes = 256;
constexpr size_t kNaErrOff = kNaCtlBytes;
constexpr size_t kNaErrBytes = (size_t)16 * D2G_NOALLOC_MAX_NCOL * sizeof(double);
constexpr size_t kNaLkfOff = kNaErrOff + kNaErrBytes;
constexpr size_t kNaLkfEntries = (size_t)D2G_NOALLOC_MAX_READS + 1;
constexpr size_t kNaListOff = kNaLkfOff + ((kNaLkfEntries * sizeof(double) + 255) & ~(size_t)255);
constexpr size_t kNaBytes = kNaListOff + (size_t)D2G_NOALLOC_MAX_PAIRS * sizeof(uint32_t);

// validation flags (control word 0)
constexpr uint32_t kNaBadArg = 1u, kNaBadBase = 2u, kNaBadReads = 4u;

}  // namespace d2g

static __device__ __align__(256) uint8_t scratch[d2g::kNaBytes];

namespace d2g {

namespace {

std::mutex g_na_mu;
std::vector<uint8_t*> g_na_ptr;  // per device: the address of `scratch`, null before set-up

struct NaCheck {
  uint32_t n;
  uint64_t pool_bytes, n_subs;
  const uint8_t* seq;
  const uint8_t* qual;  // null: not checked (no qualities, or use_quals off)
  uint32_t ncol;
  const uint64_t* off;
  const uint32_t* len;
  const uint32_t* subs;
  const uint64_t* sub_off;
  const uint32_t* reads;
  uint32_t* flags;
};

// One wavefront per pair (t is uniform over the wavefront): the row and
// its substitution list must lie inside the declared sizes before any of
// their bytes is read, then every base, quality and record is checked as
// d2g_pairs_load checks them on the host.
__global__ __launch_bounds__(256) void noalloc_check_kernel(NaCheck a) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  uint32_t bad = 0;
  for (uint32_t t = wave; t < a.n; t += waves) {
    const uint64_t o = a.off[t];
    const uint32_t L = a.len[t];
    const uint64_t s0 = a.sub_off[t], s1 = a.sub_off[t + 1];
    const uint32_t r = a.reads[t];
    const uint64_t span = ((uint64_t)L + 15u) & ~(uint64_t)15u;
    uint32_t f = 0;
    if ((o & 15u) || L == 0 || L > 65535u || o > a.pool_bytes || span > a.pool_bytes - o) f |= kNaBadArg;
    if (s1 < s0 || s1 > a.n_subs) f |= kNaBadArg;
    if (r == 0) f |= kNaBadArg;
    else if (r > (uint32_t)D2G_NOALLOC_MAX_READS) f |= kNaBadReads;
    bad |= f;
    if (f & kNaBadArg) continue;
    for (uint32_t j = lane; j < L; j += 64u) {
      const uint32_t b = a.seq[o + j];
      if (b < 1u || b > 4u) bad |= kNaBadBase;
      if (a.qual && a.qual[o + j] >= a.ncol) bad |= kNaBadArg;
    }
    for (uint64_t k = s0 + lane; k < s1; k += 64u) {
      const uint32_t rec = a.subs[k];
      const uint32_t pos = rec >> 16, nt0 = rec & 0xFFFFu;
      if (pos >= L || nt0 < 1u || nt0 > 4u) bad |= kNaBadArg;
      if (k > s0 && (a.subs[k - 1] >> 16) >= pos) bad |= kNaBadArg;  // ascending
    }
  }
  if (bad) atomicOr(a.flags, bad);
}

}  // namespace

// Set-up for `device` (the current device): the address of this device's
// copy of `scratch`, and its log(r!) table.  Called by d2g_ctx_create;
// runs once per device.
int noalloc_setup(int device) {
  std::lock_guard<std::mutex> g(g_na_mu);
  if (device < 0) return D2G_ERR_ARG;
  if (g_na_ptr.size() <= (size_t)device) g_na_ptr.resize((size_t)device + 1, nullptr);
  if (g_na_ptr[(size_t)device]) return D2G_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return D2G_ERR_HIP;
  uint8_t* base = static_cast<uint8_t*>(ptr);
  std::vector<double> lkf(kNaLkfEntries);
  lkf[0] = 0.0;
  for (size_t i = 1; i < kNaLkfEntries; ++i) lkf[i] = lkf[i - 1] + std::log((double)i);
  if (hipMemcpy(base + kNaLkfOff, lkf.data(), lkf.size() * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess)
    return D2G_ERR_HIP;
  g_na_ptr[(size_t)device] = base;
  return D2G_OK;
}

static uint8_t* noalloc_ptr(int device) {
  std::lock_guard<std::mutex> g(g_na_mu);
  return (device >= 0 && (size_t)device < g_na_ptr.size()) ? g_na_ptr[(size_t)device] : nullptr;
}

}  // namespace d2g

using namespace d2g;

extern "C" int d2g_pairs_lambda_noalloc(d2g_ctx* c, const d2g_pairs_dev* pd, const double* err,
                                        uint32_t ncol, int use_quals, double omega_a,
                                        float* out_loglam32, double* out_lambda, double* out_pval,
                                        uint8_t* out_recomputed, d2g_pair_stats* stats) {
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (!c || !pd || !err || ncol == 0) return D2G_ERR_ARG;
  if (ncol > D2G_NOALLOC_MAX_NCOL || pd->n > D2G_NOALLOC_MAX_PAIRS) return D2G_ERR_UNSUPPORTED;
  if (pd->n == 0) return D2G_OK;
  if (!pd->seq || !pd->off || !pd->len || !pd->sub_off || !pd->reads || !pd->center_reads ||
      (pd->n_subs && !pd->subs) || !out_loglam32 || !out_lambda || !out_pval || !out_recomputed)
    return D2G_ERR_ARG;
  // the kernels address the pools with 32-bit offsets
  if (pd->pool_bytes >= ((uint64_t)1 << 32) || pd->n_subs >= ((uint64_t)1 << 32))
    return D2G_ERR_UNSUPPORTED;
  CtxExec x;
  if (!ctx_exec(c, &x)) return D2G_ERR_ARG;
  if (hipSetDevice(x.device) != hipSuccess) return D2G_ERR_HIP;
  uint8_t* base = noalloc_ptr(x.device);
  if (!base) return D2G_ERR_STATE;
  uint32_t* flags = reinterpret_cast<uint32_t*>(base);
  uint32_t* nrec = flags + 1;
  double* derr = reinterpret_cast<double*>(base + kNaErrOff);
  const double* lkf = reinterpret_cast<const double*>(base + kNaLkfOff);
  uint32_t* list = reinterpret_cast<uint32_t*>(base + kNaListOff);
  const hipStream_t s = x.stream;
  const bool quals = use_quals && pd->qual;
  if (launch_zero(flags, 2, s) != hipSuccess) return D2G_ERR_HIP;
  if (hipMemcpyAsync(derr, err, (size_t)16 * ncol * sizeof(double), hipMemcpyHostToDevice, s) !=
      hipSuccess)
    return D2G_ERR_HIP;
  NaCheck ck;
  ck.n = pd->n;
  ck.pool_bytes = pd->pool_bytes;
  ck.n_subs = pd->n_subs;
  ck.seq = pd->seq;
  ck.qual = quals ? pd->qual : nullptr;
  ck.ncol = ncol;
  ck.off = pd->off;
  ck.len = pd->len;
  ck.subs = pd->subs;
  ck.sub_off = pd->sub_off;
  ck.reads = pd->reads;
  ck.flags = flags;
  const uint32_t cblocks = std::min<uint32_t>((pd->n + 3u) / 4u, (uint32_t)x.cu_count * 8u);
  hipLaunchKernelGGL(noalloc_check_kernel, dim3(cblocks), dim3(256), 0, s, ck);
  if (hipGetLastError() != hipSuccess) return D2G_ERR_HIP;
  uint32_t hflags = 0;
  if (hipMemcpyAsync(&hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return D2G_ERR_HIP;
  if (hflags & kNaBadArg) return D2G_ERR_ARG;
  if (hflags & kNaBadBase) return D2G_ERR_NONACGT;
  if (hflags & kNaBadReads) return D2G_ERR_UNSUPPORTED;
  PairArgs pa;
  pa.n = pd->n;
  pa.seq = pd->seq;
  pa.qual = pd->qual;
  pa.off = pd->off;
  pa.len = pd->len;
  pa.subs = pd->n_subs ? pd->subs : list;  // an empty list is never read past its offsets
  pa.sub_off = pd->sub_off;
  pa.reads = pd->reads;
  pa.center_reads = pd->center_reads;
  pa.err = derr;
  pa.ncol = ncol;
  pa.use_quals = use_quals;
  pa.omega = omega_a;
  pa.out_loglam = out_loglam32;
  pa.out_lambda = out_lambda;
  pa.out_p = out_pval;
  pa.out_flag = out_recomputed;
  pa.n_recomputed = nrec;
  pa.list = list;
  pa.log_kfact = lkf;
  pa.max_blocks = (uint32_t)x.cu_count * 8u;
  pa.ev_mid = x.timing ? x.ev[1] : nullptr;
  if (x.timing && hipEventRecord(x.ev[0], s) != hipSuccess) return D2G_ERR_HIP;
  if (launch_pairs_lambda(pa, s) != hipSuccess) return D2G_ERR_HIP;
  if (x.timing && hipEventRecord(x.ev[2], s) != hipSuccess) return D2G_ERR_HIP;
  uint32_t hrec = 0;
  if (hipMemcpyAsync(&hrec, nrec, sizeof(hrec), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return D2G_ERR_HIP;
  if (stats) {
    stats->pairs = pd->n;
    stats->recomputed = hrec;
    if (x.timing) {
      float a = 0.f, b = 0.f, t = 0.f;
      (void)hipEventElapsedTime(&a, x.ev[0], x.ev[1]);
      (void)hipEventElapsedTime(&b, x.ev[1], x.ev[2]);
      (void)hipEventElapsedTime(&t, x.ev[0], x.ev[2]);
      stats->ms_loglam = a;
      stats->ms_pval = b;
      stats->ms_total = t;
    }
  }
  return D2G_OK;
}
