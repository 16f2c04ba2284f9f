// libcsipm: C ABI over the gfx950 kernels of the matrix-free IPM hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <deque>
#include <map>
#include <set>
#include <unordered_set>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <tuple>
#include <complex>
#include <vector>

#include "csipm.h"
#include "csipm_batch.hpp"
// the FFT pass kernels are compiled in separate translation units when the
// build provides their declarations
#if __has_include("fft_extern.inc")
#include "fft_extern.inc"
#endif

using namespace csipm;

#define CSIPM_VERSION "0.2.0"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(CSIPM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                   \
  } while (0)

// Calls that synchronise or allocate cannot run on a stream that is being
// captured into a graph (the capture would be invalidated): they return an
// error instead.
static int no_capture(hipStream_t st, const char* what) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(CSIPM_ERR_UNSUPPORTED,
                "%s synchronises or allocates and cannot run on a stream under graph capture",
                what);
  return CSIPM_OK;
}

// Stream capture of the engine's CG graphs and the device calls of handle
// creation and destruction (allocations, copies, frees, small kernels)
// exclude each other: such a call made on another thread (a serving front end
// preparing requests) while a capture is open would invalidate the capture.
// Host-side table work runs outside the lock.
static std::mutex g_capture_mu;

#define HIPL(expr)                                          \
  do {                                                      \
    std::lock_guard<std::mutex> lk_(g_capture_mu);          \
    HIPCHK(expr);                                           \
  } while (0)

#define CHK(expr)                    \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != CSIPM_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// Device blocks of operator handles and engines (plans, masks, scratch
// spectra, FFT tables, engine vectors) come from a per-device cache of freed
// blocks in power-of-two size classes: a serving loop that makes and destroys
// a handle per request reuses blocks instead of calling hipMalloc / hipFree
// (which synchronises the whole device) for each.  Blocks above
// kCacheMaxBlock bypass the cache, at most kCacheKeepBytes per device stay
// cached (a serving stream's handles churn tens of GB: a smaller cache
// overflows into a device-synchronising hipFree per block), and an
// allocation that fails empties the cache and retries.  The owners make sure
// no work that uses a block is in flight when they return it (engines
// synchronise their stream; handles release their blocks once an event on
// theirs has completed).
// ---------------------------------------------------------------------------
constexpr size_t kCacheMaxBlock = size_t(1) << 30;
constexpr size_t kCacheKeepBytes = size_t(32) << 30;

struct BlockCache {
  std::mutex mu;
  std::unordered_map<int64_t, std::vector<void*>> free;   // (device, class) -> blocks
  std::unordered_map<void*, int64_t> owner;               // cached-size block -> key
  std::unordered_map<int, size_t> cached;                 // device -> cached bytes
};
static BlockCache g_blocks;

static int block_class(size_t bytes) {
  int c = 8;
  while (((size_t)1 << c) < bytes) ++c;
  return c;
}

static hipError_t dmalloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  if (bytes > kCacheMaxBlock) return hipMalloc(p, bytes);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const int c = block_class(bytes);
  const int64_t key = ((int64_t)dev << 8) | c;
  {
    std::lock_guard<std::mutex> lk(g_blocks.mu);
    auto it = g_blocks.free.find(key);
    if (it != g_blocks.free.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      g_blocks.cached[dev] -= (size_t)1 << c;
      return hipSuccess;
    }
  }
  e = hipMalloc(p, (size_t)1 << c);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(g_blocks.mu);
      for (auto& kv : g_blocks.free)
        if ((kv.first >> 8) == dev) {
          for (void* b : kv.second) {
            drop.push_back(b);
            g_blocks.owner.erase(b);
          }
          kv.second.clear();
        }
      g_blocks.cached[dev] = 0;
    }
    for (void* b : drop) (void)hipFree(b);
    e = hipMalloc(p, (size_t)1 << c);
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_blocks.mu);
    g_blocks.owner[*p] = key;
  }
  return e;
}

// pinned host blocks of engines (flag mirrors, table staging): a smaller
// cache of the same kind, so that the batch-of-one engine of every
// csipm_solve does not pin and unpin host memory
constexpr size_t kPinnedKeepBytes = size_t(256) << 20;
struct PinnedCache {
  std::mutex mu;
  std::unordered_map<int, std::vector<void*>> free;   // class -> blocks
  std::unordered_map<void*, int> owner;
  size_t cached = 0;
};
static PinnedCache g_pinned;

static hipError_t hmalloc(void** p, size_t bytes) {
  *p = nullptr;
  const int c = block_class(bytes == 0 ? 1 : bytes);
  {
    std::lock_guard<std::mutex> lk(g_pinned.mu);
    auto it = g_pinned.free.find(c);
    if (it != g_pinned.free.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      g_pinned.cached -= (size_t)1 << c;
      return hipSuccess;
    }
  }
  const hipError_t e = hipHostMalloc(p, (size_t)1 << c, hipHostMallocDefault);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_pinned.mu);
    g_pinned.owner[*p] = c;
  }
  return e;
}

static void hfree(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_pinned.mu);
    auto it = g_pinned.owner.find(p);
    if (it != g_pinned.owner.end()) {
      const size_t sz = (size_t)1 << it->second;
      if (g_pinned.cached + sz <= kPinnedKeepBytes) {
        g_pinned.free[it->second].push_back(p);
        g_pinned.cached += sz;
        return;
      }
      g_pinned.owner.erase(it);
    }
  }
  (void)hipHostFree(p);
}

static void dfree(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_blocks.mu);
    auto it = g_blocks.owner.find(p);
    if (it != g_blocks.owner.end()) {
      const int64_t key = it->second;
      const int dev = (int)(key >> 8);
      const size_t sz = (size_t)1 << (key & 255);
      if (g_blocks.cached[dev] + sz <= kCacheKeepBytes) {
        g_blocks.free[key].push_back(p);
        g_blocks.cached[dev] += sz;
        return;
      }
      g_blocks.owner.erase(it);
    }
  }
  (void)hipFree(p);
}

// dfree for callers that hold no capture lock: only a block the cache does
// not keep is freed, and only that takes the lock (hipFree must not run
// beside another thread's graph capture)
static void dfree_release(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_blocks.mu);
    auto it = g_blocks.owner.find(p);
    if (it != g_blocks.owner.end()) {
      const int64_t key = it->second;
      const int dev = (int)(key >> 8);
      const size_t sz = (size_t)1 << (key & 255);
      if (g_blocks.cached[dev] + sz <= kCacheKeepBytes) {
        g_blocks.free[key].push_back(p);
        g_blocks.cached[dev] += sz;
        return;
      }
      g_blocks.owner.erase(it);
    }
  }
  std::lock_guard<std::mutex> lk(g_capture_mu);
  (void)hipFree(p);
}

// Blocks of destroyed handles wait here until the work queued on their
// handle's stream before the destroy has finished (an event recorded at the
// destroy), then go back to the block cache: a destroy never waits for a
// stream (on the legacy stream that would also be every other thread's
// queued work).  Ready entries are released by later destroys and creates.
struct PendingBlocks {
  hipEvent_t ev;
  std::vector<void*> blocks;
};
static std::mutex g_pend_mu;
static std::deque<PendingBlocks> g_pend;

static void release_pending() {
  std::vector<PendingBlocks> ready;
  {
    std::lock_guard<std::mutex> lk(g_pend_mu);
    while (!g_pend.empty() && hipEventQuery(g_pend.front().ev) != hipErrorNotReady) {
      ready.push_back(std::move(g_pend.front()));
      g_pend.pop_front();
    }
  }
  for (PendingBlocks& pb : ready) {
    (void)hipEventDestroy(pb.ev);
    for (void* b : pb.blocks) dfree_release(b);
  }
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
namespace {

constexpr int kNT = 256;      // threads per workgroup of the FFT passes
constexpr int kNTBig = 1024;  // dimension-1 passes with long lines (1D)
constexpr int kNTSmall = 128;  // batched passes whose tile holds <= 128*kE elements
constexpr int kE = 16;        // complex elements staged per thread and stage
constexpr size_t kLdsSmall = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;

struct LineCfg {
  int lg = 0, ntile1 = 0, grid = 0;
  size_t lds = 0;
  int nt = 256;   // workgroup size of the batched pass (kNT or kNTSmall)
};
struct Dim1Cfg {
  int nt = 0, lg = 0, ps = 0, npairs = 0, grid = 0;
  size_t lds = 0;
};

struct Engine;

// per precision: twiddles and launch shapes (the plans are shared)
struct PrecState {
  void* tw[3] = {nullptr, nullptr, nullptr};
  void* chirp[3] = {nullptr, nullptr, nullptr};   // Bluestein tables (axes with bs)
  void* bfw[3] = {nullptr, nullptr, nullptr};
  void* binv[3] = {nullptr, nullptr, nullptr};
  LineCfg lines[3];
  Dim1Cfg d1;
  // long 1D lines (four-step): twiddles of the column (B) and row (A)
  // transforms (tw[0] holds those of the full length F, chirp[0].. the outer
  // Bluestein tables) and the launch shapes of the two pass kinds
  void* twA = nullptr;
  void* twB = nullptr;
  LineCfg col;
  Dim1Cfg row;
};

}  // namespace

struct csipm_op {
  int dim = 0;
  int n[3] = {1, 1, 1};
  int prec = CSIPM_F64;
  hipStream_t stream = nullptr;
  MapGeom g{};
  int H1 = 0;
  int64_t N = 0, Nc = 0, U = 0;  // U = 2*Nc: length of the spectrum-ordered copy of beta
  FftPlan plan[3]{};             // line FFT of each axis (Bluestein: length M; Rader: L - 1)
  int bs[3] = {0, 0, 0};         // line transform of each axis: 0 direct, 1 Bluestein, 2 Rader
  int rows[3] = {1, 1, 1};       // LDS rows per line (L, M, or 2L + 1 for Rader)
  int* rperm[3] = {nullptr, nullptr, nullptr};   // Rader permutation tables (2L ints)
  // 1D lines longer than one workgroup holds: four-step over B rows x A
  // columns of the length-F transform (F = N, or Bluestein's M when obs)
  int rdft = 1;                  // as created (0: the complex path's full-spectrum maps)
  bool big = false, obs = false;
  int bA = 0, bB = 0, bF = 0;
  FftPlan planA{}, planB{};
  FftPlan* dplan = nullptr;      // device copies: plan[0..2], planA, planB
  PrecState p64, p32;
  uint32_t* miss = nullptr;
  int64_t nmiss = 0;
  uint32_t* valid = nullptr;     // bit per u slot
  void* buf = nullptr;           // half-spectrum scratch (Nc complex doubles)
  double* part = nullptr;        // per-block partial sums
  int npart = 0;
  CgState* cg = nullptr;         // device
  CgState* cg_host = nullptr;    // pinned mirror
  void* scratch = nullptr;       // N-vector scratch of the NLP entry points (op precision)
  void* gz = nullptr;            // vector scratch of the Gondzio entry points
  size_t scratch_bytes = 0, gz_bytes = 0;
  Engine* solo = nullptr;        // batch-of-one engine of csipm_solve / csipm_cg_solve
  // live engine instances of this handle (at most one: an instance's passes
  // use the handle's scratch spectrum, so two would overwrite each other)
  std::atomic<int> admitted{0};
  bool has64() const { return prec == CSIPM_F64 || prec == CSIPM_MIXED; }
  bool has32() const { return prec == CSIPM_F32 || prec == CSIPM_MIXED; }
  size_t esize() const { return prec == CSIPM_F64 ? 8 : 4; }   // op / CG precision
  template <typename T>
  PrecState& ps();
};
template <>
PrecState& csipm_op::ps<double>() { return p64; }
template <>
PrecState& csipm_op::ps<float>() { return p32; }

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
static uint32_t magic(int d) {
  return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
}

// radix sequence for length L: the odd radices 31, 29, 23, 19 first (stage 0
// has no twiddles, and these radices load theirs where they use them), then
// 16, 8, 4, 2 for the power of two, then odd radices 17, 13, 11, 7, 15 (=3*5),
// 9, 5, 3
static bool factorize(int L, FftPlan& p) {
  p.L = L;
  p.nstages = 0;
  int m = L;
  auto push = [&](int r) {
    if (p.nstages >= kMaxStages) return false;
    p.radix[p.nstages++] = r;
    return true;
  };
  for (int r : {31, 29, 23, 19})
    while (m % r == 0) { if (!push(r)) return false; m /= r; }
  while (m % 16 == 0) { if (!push(16)) return false; m /= 16; }
  while (m % 8 == 0) { if (!push(8)) return false; m /= 8; }
  while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
  while (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
  for (int r : {17, 13, 11, 7})
    while (m % r == 0) { if (!push(r)) return false; m /= r; }
  while (m % 15 == 0) { if (!push(15)) return false; m /= 15; }
  while (m % 9 == 0) { if (!push(9)) return false; m /= 9; }
  while (m % 5 == 0) { if (!push(5)) return false; m /= 5; }
  while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
  if (m != 1) return false;
  int ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    p.ns[s] = ns;
    p.nsm[s] = magic(ns);
    p.tws[s] = L / (ns * p.radix[s]);
    ns *= p.radix[s];
  }
  return true;
}

static bool is_prime(int n) {
  if (n < 2) return false;
  for (int d = 2; (int64_t)d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

// Bluestein convolution lengths chosen by measurement (csipm_tune_bluestein):
// line length -> M.  A length without an entry takes the smallest M.
static std::mutex g_blue_mu;
static std::map<int, int> g_blue_m;

// axis plan: the direct mixed-radix plan when L factors into the supported
// radices; for a prime L whose L - 1 does, Rader's form (plan of length
// L - 1); else Bluestein with the tuned convolution length of L, or the
// smallest M >= 2L - 1 that factors.  rows: LDS rows a line needs.
static bool plan_axis(int L, FftPlan& p, int& bs, int& rows) {
  bs = 0;
  rows = L;
  if (factorize(L, p)) return true;
  if (L > 2 && is_prime(L) && factorize(L - 1, p)) {
    bs = 2;
    rows = 2 * L + 1;
    return true;
  }
  bs = 1;
  int tuned = 0;
  {
    std::lock_guard<std::mutex> lk(g_blue_mu);
    const auto it = g_blue_m.find(L);
    if (it != g_blue_m.end()) tuned = it->second;
  }
  if (tuned && factorize(tuned, p)) {
    rows = tuned;
    return true;
  }
  for (int M = 2 * L - 1; M <= 4 * L; ++M)
    if (factorize(M, p)) {
      rows = M;
      return true;
    }
  return false;
}

// exp(-2 pi t i / L) for t < L (L a multiple of 4 or not), in extended
// precision rounded once: the quarter-wave symmetries fill three quarters of
// the table from the first (exact identities, fewer trig evaluations)
static void unit_roots(int64_t L, std::vector<std::complex<double>>& r, int64_t count) {
  r.resize((size_t)count);
  const long double pi = acosl(-1.0L);
  auto eval = [&](int64_t t) {
    const long double ang = -2.0L * pi * (long double)t / (long double)L;
    return std::complex<double>((double)cosl(ang), (double)sinl(ang));
  };
  if (L % 4 != 0) {
    for (int64_t t = 0; t < count; ++t)
      r[(size_t)t] = (2 * t > L && L - t < count) ? std::conj(r[(size_t)(L - t)]) : eval(t);
    return;
  }
  const int64_t q = L / 4;
  for (int64_t t = 0; t < count; ++t) {
    if (t <= q) {
      r[(size_t)t] = eval(t);
    } else {
      // w^(t) = w^(t - q) * w^q = w^(t - q) * (-i)
      const std::complex<double> b = r[(size_t)(t - q)];
      r[(size_t)t] = std::complex<double>(b.imag(), -b.real());
    }
  }
}

// exp(-2 pi i k / M), k < M/2; kept per M (M a power of two)
static std::shared_ptr<const std::vector<std::complex<double>>> host_roots(size_t M) {
  static std::mutex mu;
  static std::unordered_map<size_t, std::shared_ptr<const std::vector<std::complex<double>>>> cache;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(M);
    if (it != cache.end()) return it->second;
  }
  auto r = std::make_shared<std::vector<std::complex<double>>>();
  unit_roots((int64_t)M, *r, (int64_t)(M / 2));
  std::lock_guard<std::mutex> lock(mu);
  if (M <= (size_t(1) << 22)) cache[M] = r;
  return r;
}

// in-place forward DFT (exp(-2 pi i jk/M)) of a power-of-two length, host FP64
static void host_fft_pow2(std::vector<std::complex<double>>& a,
                          const std::vector<std::complex<double>>& roots) {
  const size_t M = a.size();
  for (size_t i = 1, j = 0; i < M; ++i) {
    size_t bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= M; len <<= 1) {
    const size_t half = len / 2, stride = M / len;
    for (size_t i = 0; i < M; i += len)
      for (size_t k = 0; k < half; ++k) {
        const std::complex<double> u = a[i + k], v = a[i + k + half] * roots[k * stride];
        a[i + k] = u + v;
        a[i + k + half] = u - v;
      }
  }
}

// in-place forward DFT (exp(-2 pi i jk/M)) of any length M, host FP64, by
// recursive decimation in time over its prime factors (M smooth: the lengths
// the planner emits); roots = exp(-2 pi i t / M), t < M
static void host_fft_rec(const std::complex<double>* in, int64_t stride, int n,
                         std::complex<double>* out, const std::vector<std::complex<double>>& roots,
                         int64_t rstep) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = 2;
  while (n % r) ++r;
  const int m = n / r;
  std::vector<std::complex<double>> sub((size_t)n);
  for (int j = 0; j < r; ++j)
    host_fft_rec(in + j * stride, stride * r, m, sub.data() + (size_t)j * m, roots, rstep * r);
  // out[k + m q] = sum_j W_n^(j (k + m q)) sub_j[k], W_n^t = roots[t * rstep]
  for (int k = 0; k < m; ++k)
    for (int q = 0; q < r; ++q) {
      std::complex<double> acc = 0.0;
      const int64_t kk = k + (int64_t)m * q;
      for (int j = 0; j < r; ++j)
        acc += sub[(size_t)j * m + k] * roots[(size_t)(((int64_t)j * kk) % n * rstep)];
      out[kk] = acc;
    }
}

static void host_fft_any(std::vector<std::complex<double>>& a) {
  const size_t M = a.size();
  if ((M & (M - 1)) == 0) {
    host_fft_pow2(a, *host_roots(M));
    return;
  }
  std::vector<std::complex<double>> roots;
  unit_roots((int64_t)M, roots, (int64_t)M);
  std::vector<std::complex<double>> out(M);
  host_fft_rec(a.data(), 1, (int)M, out.data(), roots, 1);
  a.swap(out);
}

// Bluestein tables of line length L, convolution length M: the chirp
// w_n = exp(-i pi (n^2 mod 2L) / L) and the kernel spectra FFT_M(b)/M of the
// forward (b = conj w) and inverse (b = w) convolution kernels, wrapped so
// that b_{M-m} = b_m.  Computed once in FP64 per (L, M) (the last pair is
// kept, so the two precisions of a mixed handle share one computation).
struct BluesteinHost {
  int L = 0, M = 0;
  std::vector<std::complex<double>> w, bf, bi;
};

static std::shared_ptr<const BluesteinHost> bluestein_host(int L, int M) {
  static std::mutex mu;
  static std::shared_ptr<const BluesteinHost> last;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (last && last->L == L && last->M == M) return last;
  }
  auto tp = std::make_shared<BluesteinHost>();
  BluesteinHost& t = *tp;
  t.L = L;
  t.M = M;
  // chirp: q = n^2 mod 2L exactly, angle -pi q / L reduced to (-pi, pi]
  t.w.resize((size_t)L);
  const double pi = 3.141592653589793238462643383279502884;
  for (int n = 0; n < L; ++n) {
    long long q = ((long long)n * n) % (2LL * L);
    if (q > L) q -= 2LL * L;
    const double ang = -pi * ((double)q / (double)L);
    t.w[(size_t)n] = std::complex<double>(std::cos(ang), std::sin(ang));
  }
  // kernel spectra: b is symmetric (b_{M-m} = b_m), so FFT(conj b) =
  // conj(FFT(b)) and one transform gives both
  std::vector<std::complex<double>> b((size_t)M, 0.0);
  for (int m = 0; m < L; ++m) {
    b[(size_t)m] = t.w[(size_t)m];
    if (m > 0) b[(size_t)(M - m)] = t.w[(size_t)m];
  }
  host_fft_any(b);
  for (auto& x : b) x /= (double)M;
  t.bf.resize(b.size());
  for (size_t i = 0; i < b.size(); ++i) t.bf[i] = std::conj(b[i]);
  t.bi.swap(b);
  std::lock_guard<std::mutex> lock(mu);
  last = tp;
  return last;
}

template <typename T>
static int upload_cpx(const std::vector<std::complex<double>>& v, void** out) {
  std::vector<cpx<T>> h(v.size());
  for (size_t i = 0; i < v.size(); ++i) h[i] = {(T)v[i].real(), (T)v[i].imag()};
  std::lock_guard<std::mutex> lk(g_capture_mu);
  HIPCHK(dmalloc(out, sizeof(cpx<T>) * h.size()));
  HIPCHK(hipMemcpy(*out, h.data(), sizeof(cpx<T>) * h.size(), hipMemcpyHostToDevice));
  return CSIPM_OK;
}

template <typename T>
static int upload_bluestein(int L, int M, void** chirp, void** bfw, void** binv) {
  const std::shared_ptr<const BluesteinHost> t = bluestein_host(L, M);
  CHK(upload_cpx<T>(t->w, chirp));
  CHK(upload_cpx<T>(t->bf, bfw));
  CHK(upload_cpx<T>(t->bi, binv));
  return CSIPM_OK;
}

// Rader tables of a prime length p with primitive root g: the permutation
// perm[e] = q for e = g^q mod p (e = 1..p-1), perm[0] = p - 1 (the x_0 row
// after the p - 1 convolution rows),
// perm[p + m] = g^-m mod p (m = 0..p-2), and the kernel spectra
// FFT_{p-1}(b) / (p-1) of b_j = exp(-+2 pi i g^-j / p) (forward / inverse),
// in FP64; the last length's tables are kept (both precisions of a mixed
// handle, and the axes of a cube, share one computation)
struct RaderHost {
  int p = 0;
  std::vector<int> perm;
  std::vector<std::complex<double>> bf, bi;
};

static int primitive_root(int p) {
  std::vector<int> f;
  int m = p - 1;
  for (int d = 2; (int64_t)d * d <= m; ++d)
    if (m % d == 0) {
      f.push_back(d);
      while (m % d == 0) m /= d;
    }
  if (m > 1) f.push_back(m);
  auto pw = [p](int64_t b, int64_t e) {
    int64_t r = 1;
    b %= p;
    for (; e; e >>= 1, b = b * b % p)
      if (e & 1) r = r * b % p;
    return r;
  };
  for (int g = 2; g < p; ++g) {
    bool ok = true;
    for (int q : f) ok = ok && pw(g, (p - 1) / q) != 1;
    if (ok) return g;
  }
  return 1;   // p = 2 (never planned)
}

static std::shared_ptr<const RaderHost> rader_host(int p) {
  static std::mutex mu;
  static std::shared_ptr<const RaderHost> last;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (last && last->p == p) return last;
  }
  auto tp = std::make_shared<RaderHost>();
  RaderHost& t = *tp;
  t.p = p;
  const int g = primitive_root(p);
  int64_t ginv = 1;   // g^-1 = g^(p-2)
  for (int k = 0; k < p - 2; ++k) ginv = ginv * g % p;
  t.perm.assign((size_t)2 * p, 0);
  t.perm[0] = p - 1;
  int64_t e = 1, einv = 1;
  std::vector<std::complex<double>> roots;
  unit_roots(p, roots, p);   // exp(-2 pi i t / p)
  std::vector<std::complex<double>> bf((size_t)p - 1), bi((size_t)p - 1);
  for (int q = 0; q < p - 1; ++q) {
    t.perm[(size_t)e] = q;                  // e = g^q
    t.perm[(size_t)p + q] = (int)einv;      // g^-q
    bf[(size_t)q] = roots[(size_t)einv];
    bi[(size_t)q] = std::conj(roots[(size_t)einv]);
    e = e * g % p;
    einv = einv * ginv % p;
  }
  host_fft_any(bf);
  host_fft_any(bi);
  for (auto& x : bf) x /= (double)(p - 1);
  for (auto& x : bi) x /= (double)(p - 1);
  t.bf.swap(bf);
  t.bi.swap(bi);
  std::lock_guard<std::mutex> lock(mu);
  last = tp;
  return last;
}

static int upload_rader_perm(int p, int** perm) {
  const std::shared_ptr<const RaderHost> t = rader_host(p);
  std::lock_guard<std::mutex> lk(g_capture_mu);
  HIPCHK(dmalloc((void**)perm, sizeof(int) * t->perm.size()));
  HIPCHK(hipMemcpy(*perm, t->perm.data(), sizeof(int) * t->perm.size(), hipMemcpyHostToDevice));
  return CSIPM_OK;
}

template <typename T>
static int upload_rader(int p, void** bfw, void** binv) {
  const std::shared_ptr<const RaderHost> t = rader_host(p);
  CHK(upload_cpx<T>(t->bf, bfw));
  CHK(upload_cpx<T>(t->bi, binv));
  return CSIPM_OK;
}

// exp(-2 pi i t / L), t < L, in the op precision.  Tables are shared by the
// handles on a device and counted: a table no handle holds stays cached
// (a serving process meets the same lengths again) until the unheld tables
// exceed kTwiddleKeepBytes, when the least recently used go back to the
// block cache.  Lookups by key and by pointer and the choice of a victim do
// not scan the cache (a serving stream leaves thousands of lengths in it).
constexpr size_t kTwiddleKeepBytes = size_t(256) << 20;

struct TwiddleEntry {
  void* ptr = nullptr;
  size_t bytes = 0;
  int refs = 0;
  uint64_t last = 0;
};

static std::mutex g_tw_mu;
static std::unordered_map<int64_t, TwiddleEntry> g_tw;      // key -> table
static std::unordered_map<void*, int64_t> g_tw_key;         // table -> key
static std::set<std::pair<uint64_t, int64_t>> g_tw_unheld;  // (last use, key), refs == 0
static size_t g_tw_unheld_bytes = 0;
static uint64_t g_tw_clock = 0;

template <typename T>
static int upload_twiddles(int L, void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const int64_t key = ((int64_t)dev << 40) | ((int64_t)sizeof(T) << 32) | (int64_t)L;
  // an existing table with no holder is in the unheld set (a new one is not)
  auto hold = [&](TwiddleEntry& e, bool fresh) {
    if (e.refs == 0 && !fresh) {
      g_tw_unheld.erase({e.last, key});
      g_tw_unheld_bytes -= e.bytes;
    }
    e.refs += 1;
    e.last = ++g_tw_clock;
    *out = e.ptr;
  };
  {
    std::lock_guard<std::mutex> lock(g_tw_mu);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) {
      hold(it->second, false);
      return CSIPM_OK;
    }
  }
  // a new table: computed and uploaded without holding the cache lock
  std::vector<std::complex<double>> r;
  unit_roots(L, r, L);
  std::vector<cpx<T>> h((size_t)L);
  for (int t = 0; t < L; ++t) h[(size_t)t] = {(T)r[(size_t)t].real(), (T)r[(size_t)t].imag()};
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    HIPCHK(dmalloc(&p, sizeof(cpx<T>) * (size_t)L));
    const hipError_t e = hipMemcpy(p, h.data(), sizeof(cpx<T>) * (size_t)L, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      dfree(p);
      return fail(CSIPM_ERR_HIP, "twiddle upload: %s", hipGetErrorString(e));
    }
  }
  void* dup = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_tw_mu);
    TwiddleEntry& e = g_tw[key];
    const bool fresh = e.ptr == nullptr;
    if (!fresh) {
      dup = p;            // another thread made the same table meanwhile
    } else {
      e.ptr = p;
      e.bytes = sizeof(cpx<T>) * (size_t)L;
      g_tw_key[p] = key;
    }
    hold(e, fresh);
  }
  dfree_release(dup);
  return CSIPM_OK;
}

// a handle lets go of a table it got from upload_twiddles
static void release_twiddles(void* p) {
  if (!p) return;
  std::vector<void*> victims;
  {
    std::lock_guard<std::mutex> lock(g_tw_mu);
    auto k = g_tw_key.find(p);
    if (k == g_tw_key.end()) return;
    TwiddleEntry& e = g_tw[k->second];
    if (e.refs > 0 && --e.refs == 0) {
      g_tw_unheld.insert({e.last, k->second});
      g_tw_unheld_bytes += e.bytes;
    }
    while (g_tw_unheld_bytes > kTwiddleKeepBytes && !g_tw_unheld.empty()) {
      const int64_t vk = g_tw_unheld.begin()->second;
      g_tw_unheld.erase(g_tw_unheld.begin());
      auto v = g_tw.find(vk);
      g_tw_unheld_bytes -= v->second.bytes;
      victims.push_back(v->second.ptr);
      g_tw_key.erase(v->second.ptr);
      g_tw.erase(v);
    }
  }
  for (void* v : victims) dfree_release(v);
}

// Launch-shape choices of the batched passes measured by the autotuner
// (csipm.tune, persisted in a committed table and loaded at start-up): for a
// pass kind (0: line pass, 1: dimension-1 pass), the complex element size,
// the FFT length W and the LDS rows R of a line, the largest tile (log2 of
// the lines or line pairs per workgroup) and the workgroup size (128 or
// 256).  Shapes without an entry keep the built-in rules.
struct TuneVal {
  int lg_cap = 30;
  int nt = 0;   // 0: the built-in choice
};
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, TuneVal> g_tune;

static TuneVal tune_lookup(int kind, size_t cs, int W, int R) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_tune.empty()) return TuneVal{};
  auto it = g_tune.find(std::make_tuple(kind, (int)cs, W, R));
  return it == g_tune.end() ? TuneVal{} : it->second;
}

// smallest tile (lines per workgroup) whose rows span 64 bytes: with the
// XCD-aware block order of the pass kernels, the two tiles sharing a 128-byte
// line run on one XCD and fetch it once
static int min_tile_lg(size_t cs) {
  int lg = 0;
  while ((size_t)(1 << lg) * cs < 64) ++lg;
  return lg;
}

// line FFTs along dim 2/3: 2^lg lines per workgroup, as many as the
// registers (lines * W <= NT*E for the FFT length W) and 64 KB of LDS (lines *
// R rows) allow, fewer when that
// would leave the launch with under min_grid workgroups (1024 for one
// instance; a batch shares the target among its instances)

static LineCfg line_cfg(int H1, int W, int R, int64_t others, size_t cs,
                        int min_grid = 1024, int lg_cap = 30) {
  LineCfg c{};
  int lg = 0;
  while ((2 << lg) * W <= kNT * kE && (size_t)(2 << lg) * R * cs <= kLdsSmall && (1 << lg) < H1 &&
         lg < lg_cap)
    ++lg;
  // below 64-byte rows only while the launch would not even give a
  // quarter of the workgroup target (small single instances: latency first)
  const int lg_min = std::min(lg, min_tile_lg(cs));
  auto grid_of = [&](int l) { return (int64_t)((H1 + (1 << l) - 1) >> l) * others; };
  while (lg > 0 && grid_of(lg) < min_grid && !(lg <= lg_min && 4 * grid_of(lg) >= min_grid))
    --lg;
  c.lg = lg;
  c.ntile1 = (H1 + (1 << lg) - 1) >> lg;
  c.grid = (int)(c.ntile1 * others);
  c.lds = (size_t)(1 << lg) * R * cs;
  return c;
}

// four-step column pass: 2^lg adjacent columns of length B per workgroup
static LineCfg col_cfg(int A, int B, size_t cs, int min_grid = 1024) {
  LineCfg c{};
  int lg = 0;
  while ((2 << lg) * B <= kNT * kE && (size_t)(2 << lg) * B * cs <= kLdsSmall && (2 << lg) <= A)
    ++lg;
  const int lg_min = std::min(lg, min_tile_lg(cs));
  auto grid_of = [&](int l) { return (A + (1 << l) - 1) >> l; };
  while (lg > 0 && grid_of(lg) < min_grid && !(lg <= lg_min && 4 * grid_of(lg) >= min_grid)) --lg;
  c.lg = lg;
  c.ntile1 = (A + (1 << lg) - 1) >> lg;
  c.grid = c.ntile1;
  c.lds = (size_t)(1 << lg) * B * cs;
  return c;
}

// four-step row pass: 2^lg rows of length A per workgroup (LDS stride padded)
static Dim1Cfg row_cfg(int A, int B, size_t cs, int min_grid = 1024) {
  Dim1Cfg c{};
  c.nt = kNT;
  int lg = 0;
  auto lds_of = [&](int l) { return (size_t)A * ((1 << l) + 1) * cs; };
  while ((2 << lg) * A <= kNT * kE && lds_of(lg + 1) <= kLdsSmall && (2 << lg) <= B) ++lg;
  while (lg > 0 && ((B + (1 << lg) - 1) >> lg) < min_grid) --lg;
  c.lg = lg;
  c.ps = (1 << lg) + 1;
  c.grid = (B + (1 << lg) - 1) >> lg;
  c.lds = lds_of(lg);
  return c;
}

static Dim1Cfg dim1_cfg(int W, int R, int nlines, size_t cs, int min_grid = 1024,
                        int lg_cap = 30) {
  Dim1Cfg c{};
  c.npairs = (nlines + 1) / 2;
  c.nt = (W <= kNT * kE) ? kNT : kNTBig;
  int lg = 0;
  auto lds_of = [&](int l) { return (size_t)R * ((1 << l) + (l > 0 ? 1 : 0)) * cs; };
  while ((2 << lg) * W <= c.nt * kE && lds_of(lg + 1) <= kLdsSmall && (1 << lg) < c.npairs &&
         lg < lg_cap)
    ++lg;
  while (lg > 0 && ((c.npairs + (1 << lg) - 1) >> lg) < min_grid) --lg;
  c.lg = lg;
  c.ps = (1 << lg) + (lg > 0 ? 1 : 0);
  c.grid = (c.npairs + (1 << lg) - 1) >> lg;
  c.lds = lds_of(lg);
  return c;
}

// ---------------------------------------------------------------------------
// pass launchers
// ---------------------------------------------------------------------------
namespace {

template <typename T>
Epi<T> epi_store(T* out) {
  Epi<T> e{};
  e.kind = 0;
  e.out = out;
  return e;
}

// raise a kernel's dynamic-LDS limit once (the attribute persists; graph
// capture then sees no attribute calls)
template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= kLdsSmall) return CSIPM_OK;
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> done;
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = done[(const void*)kernel];
  if (have >= bytes) return CSIPM_OK;
  HIPCHK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
  have = bytes;
  return CSIPM_OK;
}

template <typename T>
PassArgs<T> base_args(csipm_op* h, int axis) {
  PassArgs<T> a{};
  a.g = h->g;
  a.plan = h->dplan + axis;
  a.tw = (const cpx<T>*)h->ps<T>().tw[axis];
  a.H1 = h->H1;
  a.buf = (cpx<T>*)h->buf;
  a.h1m = magic(h->H1);
  a.n1m = magic(h->n[0]);
  a.miss = h->miss;
  a.len = h->n[axis];
  a.bs = h->bs[axis];
  a.rperm = h->rperm[axis];
  a.chirp = (const cpx<T>*)h->ps<T>().chirp[axis];
  a.bfw = (const cpx<T>*)h->ps<T>().bfw[axis];
  a.binv = (const cpx<T>*)h->ps<T>().binv[axis];
  return a;
}

template <typename T>
PassArgs<T> lines_args(csipm_op* h, int axis, const void* src, const Epi<T>& epi,
                       const int* done, const LineCfg* cfg = nullptr) {
  const LineCfg& c = cfg ? *cfg : h->ps<T>().lines[axis];
  PassArgs<T> a = base_args<T>(h, axis);
  a.axis = axis;
  a.lg = c.lg;
  a.ps = 1 << c.lg;
  a.ntile1 = c.ntile1;
  a.beta_in = (const T*)src;
  a.u_in = (const cpx<T>*)src;
  a.epi = epi;
  a.done = done;
  return a;
}

template <typename T>
PassArgs<T> dim1_args(csipm_op* h, const void* src, const T* real_in, T* real_out,
                      const Epi<T>& epi, T scale, const int* done,
                      const Dim1Cfg* cfg = nullptr) {
  const Dim1Cfg& c = cfg ? *cfg : h->ps<T>().d1;
  PassArgs<T> a = base_args<T>(h, 0);
  a.lg = c.lg;
  a.ps = c.ps;
  a.nlines = (int)(h->N / h->n[0]);
  a.npairs = c.npairs;
  a.beta_in = (const T*)src;
  a.u_in = (const cpx<T>*)src;
  a.real_in = real_in;
  a.real_out = real_out;
  a.epi = epi;
  a.scale = scale;
  a.done = done;
  return a;
}

template <typename T, int D, int SRC, int DST, bool FWD>
int launch_lines(csipm_op* h, int axis, const void* src, const Epi<T>& epi, const int* done) {
  const LineCfg& c = h->ps<T>().lines[axis];
  PassArgs<T> a = lines_args<T>(h, axis, src, epi, done);
  auto kern = k_lines<T, kNT, kE, D, SRC, DST, FWD>;
  CHK(allow_lds(kern, c.lds));
  hipLaunchKernelGGL(kern, dim3(c.grid), dim3(kNT), c.lds, h->stream, a);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T, int D, int SRC, int DST, int MODE>
int launch_dim1(csipm_op* h, const void* src, const T* real_in, T* real_out, const Epi<T>& epi,
                T scale, const int* done) {
  const Dim1Cfg& c = h->ps<T>().d1;
  PassArgs<T> a = dim1_args<T>(h, src, real_in, real_out, epi, scale, done);
  if (c.nt == kNT) {
    auto kern = k_dim1<T, kNT, kE, D, SRC, DST, MODE>;
    CHK(allow_lds(kern, c.lds));
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(kNT), c.lds, h->stream, a);
  } else {
    auto kern = k_dim1<T, kNTBig, kE, D, SRC, DST, MODE>;
    CHK(allow_lds(kern, c.lds));
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(kNTBig), c.lds, h->stream, a);
  }
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// long 1D lines (four-step passes)
// ---------------------------------------------------------------------------
template <typename T>
PassArgs<T> big_args(csipm_op* h, bool col, const void* src, const T* real_in, T* real_out,
                     const Epi<T>& epi, T scale, const int* done, const LineCfg* cc = nullptr,
                     const Dim1Cfg* rc = nullptr) {
  PassArgs<T> a = base_args<T>(h, 0);
  const PrecState& ps = h->ps<T>();
  a.A = h->bA;
  a.B = h->bB;
  a.F = h->bF;
  a.obs = h->obs ? 1 : 0;
  a.am = magic(h->bA);
  a.twF = (const cpx<T>*)ps.tw[0];
  a.bs = 0;
  if (col) {
    const LineCfg& c = cc ? *cc : ps.col;
    a.plan = h->dplan + 4;
    a.tw = (const cpx<T>*)ps.twB;
    a.lg = c.lg;
    a.ps = 1 << c.lg;
  } else {
    const Dim1Cfg& c = rc ? *rc : ps.row;
    a.plan = h->dplan + 3;
    a.tw = (const cpx<T>*)ps.twA;
    a.lg = c.lg;
    a.ps = c.ps;
  }
  a.beta_in = (const T*)src;
  a.u_in = (const cpx<T>*)src;
  a.real_in = real_in;
  a.real_out = real_out;
  a.epi = epi;
  a.scale = scale;
  a.done = done;
  return a;
}

enum { PK_LINES = 0, PK_DIM1 = 1, PK_COL = 2, PK_ROW = 3 };
struct PassSpec {
  int kind, D, src, dst, fm, axis;  // fm: FWD (lines), MODE (dim1, col, row)
};

// pass sequence of a long 1D line.  which: 0 M^T M (+ epilogue), 1 M^T z from
// real samples, 2 M beta to real samples
enum { BIG_MTM = 0, BIG_R2C = 1, BIG_C2R = 2 };
int big_specs(const csipm_op* h, int which, int src, int dst, PassSpec* s) {
  const bool bs = h->obs;
  if (which == BIG_MTM) {
    if (!bs) {
      s[0] = {PK_COL, 1, src, DST_BUF, C_FIRST, 0};
      s[1] = {PK_ROW, 1, SRC_BUF, DST_BUF, R_MTM, 0};
      s[2] = {PK_COL, 1, SRC_BUF, dst, C_LAST, 0};
      return 3;
    }
    s[0] = {PK_COL, 1, src, DST_BUF, C_FIRST, 0};
    s[1] = {PK_ROW, 1, SRC_BUF, DST_BUF, R_CONV_INV, 0};
    s[2] = {PK_COL, 1, SRC_BUF, DST_BUF, C_MID, 0};
    s[3] = {PK_ROW, 1, SRC_BUF, DST_BUF, R_CONV_FWD, 0};
    s[4] = {PK_COL, 1, SRC_BUF, dst, C_LAST, 0};
    return 5;
  }
  if (which == BIG_R2C) {
    if (!bs) {
      s[0] = {PK_ROW, 1, SRC_REAL, DST_BUF, R_R2C, 0};
      s[1] = {PK_COL, 1, SRC_BUF, dst, C_LAST, 0};
      return 2;
    }
    s[0] = {PK_COL, 1, SRC_REAL, DST_BUF, C_REAL_IN, 0};
    s[1] = {PK_ROW, 1, SRC_BUF, DST_BUF, R_CONV_FWD, 0};
    s[2] = {PK_COL, 1, SRC_BUF, dst, C_LAST, 0};
    return 3;
  }
  if (!bs) {
    s[0] = {PK_COL, 1, src, DST_BUF, C_FIRST, 0};
    s[1] = {PK_ROW, 1, SRC_BUF, DST_REAL, R_C2R, 0};
    return 2;
  }
  s[0] = {PK_COL, 1, src, DST_BUF, C_FIRST, 0};
  s[1] = {PK_ROW, 1, SRC_BUF, DST_BUF, R_CONV_INV, 0};
  s[2] = {PK_COL, 1, SRC_BUF, DST_REAL, C_MID_C2R, 0};
  return 3;
}

// one four-step pass of a single instance
template <typename T>
int launch_big_pass(csipm_op* h, const PassSpec& sp, const PassArgs<T>& a) {
  const PrecState& ps = h->ps<T>();
  const int grid = sp.kind == PK_COL ? ps.col.grid : ps.row.grid;
  const size_t lds = sp.kind == PK_COL ? ps.col.lds : ps.row.lds;
#define CSIPM_BIG(kern)                                                      \
  do {                                                                       \
    CHK(allow_lds(kern, lds));                                               \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kNT), lds, h->stream, a);      \
    HIPCHK(hipGetLastError());                                               \
    return CSIPM_OK;                                                         \
  } while (0)
  if (sp.kind == PK_COL) {
    switch (sp.fm) {
      case C_FIRST:
        if (sp.src == SRC_BETA) CSIPM_BIG((k_col<T, kNT, kE, SRC_BETA, DST_BUF, C_FIRST>));
        CSIPM_BIG((k_col<T, kNT, kE, SRC_U, DST_BUF, C_FIRST>));
      case C_REAL_IN: CSIPM_BIG((k_col<T, kNT, kE, SRC_REAL, DST_BUF, C_REAL_IN>));
      case C_MID: CSIPM_BIG((k_col<T, kNT, kE, SRC_BUF, DST_BUF, C_MID>));
      case C_MID_C2R: CSIPM_BIG((k_col<T, kNT, kE, SRC_BUF, DST_REAL, C_MID_C2R>));
      case C_LAST:
        if (sp.dst == DST_BETA) CSIPM_BIG((k_col<T, kNT, kE, SRC_BUF, DST_BETA, C_LAST>));
        CSIPM_BIG((k_col<T, kNT, kE, SRC_BUF, DST_U, C_LAST>));
      default: break;
    }
  } else {
    switch (sp.fm) {
      case R_MTM: CSIPM_BIG((k_row<T, kNT, kE, R_MTM>));
      case R_C2R: CSIPM_BIG((k_row<T, kNT, kE, R_C2R>));
      case R_R2C: CSIPM_BIG((k_row<T, kNT, kE, R_R2C>));
      case R_CONV_INV: CSIPM_BIG((k_row<T, kNT, kE, R_CONV_INV>));
      case R_CONV_FWD: CSIPM_BIG((k_row<T, kNT, kE, R_CONV_FWD>));
      default: break;
    }
  }
#undef CSIPM_BIG
  return fail(CSIPM_ERR_UNSUPPORTED, "no four-step kernel for pass (%d, %d)", sp.kind, sp.fm);
}

// run a four-step pass sequence on one instance; ev: an event after every pass
template <typename T>
int run_big(csipm_op* h, int which, int src, int dst, const void* x, const T* real_in,
            T* real_out, const Epi<T>& epi, T scale, const int* done, hipEvent_t* ev) {
  PassSpec sp[8];
  const int np = big_specs(h, which, src, dst, sp);
  for (int k = 0; k < np; ++k) {
    const bool col = sp[k].kind == PK_COL;
    PassArgs<T> a = big_args<T>(h, col, x, real_in, real_out, epi, scale, done);
    CHK(launch_big_pass<T>(h, sp[k], a));
    if (ev) HIPCHK(hipEventRecord(ev[k], h->stream));
  }
  return CSIPM_OK;
}

enum { SP_BETA = 0, SP_U = 1 };

// the fused M^T M (epi kind 0) or condensed-KKT apply (kind 1) on beta-layout
// (SP_BETA) or spectrum-ordered (SP_U) vectors; ev: an event after every pass
template <typename T, int D, int SP>
int apply_mtm(csipm_op* h, const T* x, const Epi<T>& epi, const int* done, hipEvent_t* ev) {
  constexpr int SRC = SP == SP_U ? SRC_U : SRC_BETA;
  constexpr int DST = SP == SP_U ? DST_U : DST_BETA;
  const T s = (T)(1.0 / (double)h->N);
  int k = 0;
  auto mark = [&]() -> int {
    if (ev) HIPCHK(hipEventRecord(ev[k++], h->stream));
    return CSIPM_OK;
  };
  if (D == 1 && h->big) return run_big<T>(h, BIG_MTM, SRC, DST, x, nullptr, nullptr, epi, s, done, ev);
  if (D == 1) {
    CHK((launch_dim1<T, 1, SRC, DST, MODE_MTM>(h, x, nullptr, nullptr, epi, s, done)));
    return mark();
  }
  if (D == 2) {
    CHK((launch_lines<T, 2, SRC, DST_BUF, false>(h, 1, x, epi, done)));
    CHK(mark());
    CHK((launch_dim1<T, 2, SRC_BUF, DST_BUF, MODE_MTM>(h, nullptr, nullptr, nullptr, epi, s, done)));
    CHK(mark());
    CHK((launch_lines<T, 2, SRC_BUF, DST, true>(h, 1, nullptr, epi, done)));
    return mark();
  }
  CHK((launch_lines<T, 3, SRC, DST_BUF, false>(h, 2, x, epi, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, false>(h, 1, nullptr, epi, done)));
  CHK(mark());
  CHK((launch_dim1<T, 3, SRC_BUF, DST_BUF, MODE_MTM>(h, nullptr, nullptr, nullptr, epi, s, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, true>(h, 1, nullptr, epi, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST, true>(h, 2, nullptr, epi, done)));
  return mark();
}

template <typename T, int D>
int apply_m_perp_beta(csipm_op* h, const T* beta, T* y) {
  const T s = (T)(1.0 / std::sqrt((double)h->N));
  Epi<T> e{};
  if (D == 1 && h->big)
    return run_big<T>(h, BIG_C2R, SRC_BETA, DST_BUF, beta, nullptr, y, e, s, nullptr, nullptr);
  if (D == 1)
    return launch_dim1<T, 1, SRC_BETA, DST_REAL, MODE_C2R>(h, beta, nullptr, y, e, s, nullptr);
  if (D == 2) {
    CHK((launch_lines<T, 2, SRC_BETA, DST_BUF, false>(h, 1, beta, e, nullptr)));
    return launch_dim1<T, 2, SRC_BUF, DST_REAL, MODE_C2R>(h, nullptr, nullptr, y, e, s, nullptr);
  }
  CHK((launch_lines<T, 3, SRC_BETA, DST_BUF, false>(h, 2, beta, e, nullptr)));
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, false>(h, 1, nullptr, e, nullptr)));
  return launch_dim1<T, 3, SRC_BUF, DST_REAL, MODE_C2R>(h, nullptr, nullptr, y, e, s, nullptr);
}

// beta (SP_BETA) or u (SP_U) = M^T z
template <typename T, int D, int SP>
int apply_m_perpt_z(csipm_op* h, const T* z, T* out) {
  constexpr int DST = SP == SP_U ? DST_U : DST_BETA;
  const T s = (T)(1.0 / std::sqrt((double)h->N));
  Epi<T> e = epi_store<T>(out);
  if (D == 1 && h->big)
    return run_big<T>(h, BIG_R2C, SRC_REAL, DST, nullptr, z, nullptr, e, s, nullptr, nullptr);
  if (D == 1)
    return launch_dim1<T, 1, SRC_REAL, DST, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr);
  if (D == 2) {
    CHK((launch_dim1<T, 2, SRC_REAL, DST_BUF, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr)));
    return launch_lines<T, 2, SRC_BUF, DST, true>(h, 1, nullptr, e, nullptr);
  }
  CHK((launch_dim1<T, 3, SRC_REAL, DST_BUF, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr)));
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, true>(h, 1, nullptr, e, nullptr)));
  return launch_lines<T, 3, SRC_BUF, DST, true>(h, 2, nullptr, e, nullptr);
}

template <typename T, int SP>
int dispatch_mtm(csipm_op* h, const T* x, const Epi<T>& epi, const int* done,
                 hipEvent_t* ev = nullptr) {
  if (h->dim == 1) return apply_mtm<T, 1, SP>(h, x, epi, done, ev);
  if (h->dim == 2) return apply_mtm<T, 2, SP>(h, x, epi, done, ev);
  return apply_mtm<T, 3, SP>(h, x, epi, done, ev);
}

template <typename T, int SP>
int dispatch_m_perpt_z(csipm_op* h, const T* z, T* out) {
  if (h->dim == 1) return apply_m_perpt_z<T, 1, SP>(h, z, out);
  if (h->dim == 2) return apply_m_perpt_z<T, 2, SP>(h, z, out);
  return apply_m_perpt_z<T, 3, SP>(h, z, out);
}

int grid_for(int64_t n, int per_thread = 1) {
  int64_t g = (n + (int64_t)kNT * per_thread - 1) / ((int64_t)kNT * per_thread);
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

// condensed K apply; n = length of each half of x and y (N or U)
template <typename T, int SP>
int kkt_apply_t(csipm_op* h, const T* x, T* y, const T* L1, const T* L2, double alpha,
                double beta, double* dot_part, const int* done, hipEvent_t* ev = nullptr) {
  const int64_t n = SP == SP_U ? h->U : h->N;
  Epi<T> e{};
  e.kind = 1;
  e.xb = x;
  e.xz = x + n;
  e.L1 = L1;
  e.L2 = L2;
  e.yb = y;
  e.yz = y + n;
  e.alpha = (T)alpha;
  e.beta = (T)beta;
  e.dot_part = dot_part;
  return dispatch_mtm<T, SP>(h, x, e, done, ev);
}

// reduced (Schur-complement) apply y = alpha (M^T M + D) x + beta y on one
// n-block, the operator of the eliminated system the mixed-precision CG solves
template <typename T, int SP>
int red_apply_t(csipm_op* h, const T* x, T* y, const T* D, double alpha, double beta,
                double* dot_part, const int* done, hipEvent_t* ev = nullptr) {
  Epi<T> e{};
  e.kind = 2;
  e.xb = x;
  e.L1 = D;
  e.yb = y;
  e.alpha = (T)alpha;
  e.beta = (T)beta;
  e.dot_part = dot_part;
  return dispatch_mtm<T, SP>(h, x, e, done, ev);
}

// names and algorithmic HBM bytes of the passes of one apply:
// layout 0 condensed K on beta-layout vectors, 1 condensed K on spectrum-
// ordered vectors, 2 reduced apply on spectrum-ordered vectors
int kkt_pass_info(csipm_op* h, int layout, const char** names, double* bytes) {
  const double s = (double)h->esize(), N = (double)h->N, Nc = (double)h->Nc;
  const double n = layout == 0 ? N : (double)h->U;
  const double mask = h->nmiss > 0 ? std::ceil(N / 8.0) : 0.0;
  const double epi = (layout == 2 ? 3 : 6) * n * s;   // x (and z), diagonals in; y out
  const bool red = layout == 2;
  if (h->big) {
    const double F = 2.0 * (double)h->bF * s;   // one pass over the length-F scratch
    int k = 0;
    names[k] = "col_first";  bytes[k++] = n * s + F;
    names[k] = h->obs ? "row_conv_inv" : "row_mtm";  bytes[k++] = 2 * F + (h->obs ? F : mask);
    if (h->obs) {
      names[k] = "col_mid";      bytes[k++] = 2 * F + mask;
      names[k] = "row_conv_fwd"; bytes[k++] = 3 * F;
    }
    names[k] = red ? "col_last_red" : "col_last_kkt";  bytes[k++] = F + epi;
    return k;
  }
  if (h->dim == 1) {
    names[0] = red ? "dim1_mtm_red" : "dim1_mtm_kkt";
    bytes[0] = epi + mask;
    return 1;
  }
  const double spec = 2 * Nc * s;   // one pass over the half spectrum
  if (h->dim == 2) {
    names[0] = "lines2_inv_from_beta"; bytes[0] = n * s + spec;
    names[1] = "dim1_c2r_mask_r2c";    bytes[1] = 2 * spec + mask;
    names[2] = red ? "lines2_fwd_red" : "lines2_fwd_kkt";
    bytes[2] = spec + epi;
    return 3;
  }
  names[0] = "lines3_inv_from_beta"; bytes[0] = n * s + spec;
  names[1] = "lines2_inv";           bytes[1] = 2 * spec;
  names[2] = "dim1_c2r_mask_r2c";    bytes[2] = 2 * spec + mask;
  names[3] = "lines2_fwd";           bytes[3] = 2 * spec;
  names[4] = red ? "lines3_fwd_red" : "lines3_fwd_kkt";
  bytes[4] = spec + epi;
  return 5;
}

#include "csipm_engine.hpp"

// the handle's batch-of-one engine, created on first use, sized for it
int solo_engine(csipm_op* h, Engine** out) {
  if (!h->solo) {
    const bool mixed = h->prec == CSIPM_MIXED;
    CHK(engine_create(&h->solo, slots_of(h), 1, (int)h->esize(), mixed, h->stream));
  }
  h->solo->user = h->stream;
  *out = h->solo;
  return engine_enter(h->solo);
}

// csipm_cg_solve: the reference-layout system converted to the spectrum
// order, solved there by the batched CG (one instance), converted back
template <typename T>
int cg_solve_solo(csipm_op* h, const T* b, T* x, const T* L1, const T* L2, const T* P11,
                  const T* P12, const T* P22, double atol, double rtol, int64_t itmax,
                  int64_t* iters, double* resid, int* status) {
  Engine* E = nullptr;
  CHK(solo_engine(h, &E));
  int id = -1;
  CHK(engine_admit(E, h, nullptr, 0.0, nullptr, itmax, &id));
  const int64_t N = h->N, U = h->U, off = E->inst[(size_t)id].off;
  hipStream_t st = E->stream;
  int rc = CSIPM_OK;
  do {
    T* ub = E->C2<T>(C2_RHS) + 2 * off;
    if ((rc = to_u<T>(h, b, ub, st)) || (rc = to_u<T>(h, b + N, ub + U, st))) break;
    const T* diag[5] = {L1, L2, P11, P12, P22};
    const int slot[5] = {C_L1, C_L2, C_P11, C_P12, C_P22};
    for (int k = 0; k < 5 && !rc; ++k) rc = to_u<T>(h, diag[k], E->C<T>(slot[k]) + off, st, 0);
    if (rc) break;
    if ((rc = engine_tables<T>(E, 2, false))) break;
    if ((rc = engine_cg<T>(E, 2, atol, rtol, 0))) break;
    T* ux = E->C2<T>(C2_D) + 2 * off;
    if ((rc = from_u<T>(h, ux, x, st)) || (rc = from_u<T>(h, ux + U, x + N, st))) break;
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(CSIPM_ERR_HIP, "CG failed");
      break;
    }
    const CgState& cs = E->cg_h[id];
    *iters = cs.iters;
    *resid = cs.rnorm;
    *status = cs.status;
  } while (false);
  engine_drop(E, id);
  if (rc) return rc;
  return engine_leave(E);
}

// csipm_admm: ADMM on a batch of one
template <typename TC>
int admm_solo(csipm_op* h, const double* z0, double lam, double rho, int64_t maxt, double tol,
              double* beta_out, double* info) {
  const auto t0 = std::chrono::steady_clock::now();
  Engine* E = nullptr;
  CHK(solo_engine(h, &E));
  int id = -1;
  CHK(engine_admit(E, h, z0, lam, beta_out, 2 * h->N, &id));   // Krylov cg: itmax = 2n
  const double eps = std::sqrt(std::numeric_limits<double>::epsilon());
  const int mt = (int)std::min<int64_t>(maxt, 1 << 30);
  int rc = CSIPM_OK;
  for (;;) {
    int running = 0;
    if ((rc = engine_admm_step<TC>(E, rho, mt, tol, eps, eps, &running))) break;
    if (running == 0) break;
  }
  if (rc) {
    engine_drop(E, id);
    return rc;
  }
  CHK(engine_harvest(E, id, info));
  info[0] = 0.0;
  info[6] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return engine_leave(E);
}

// csipm_solve: the interior-point driver on a batch of one
template <typename TC>
int solve_solo(csipm_op* h, const double* z0, double lam, double tol, int64_t max_iter,
               double cg_atol, double cg_rtol, int64_t cg_itmax, double* beta_out, double* info,
               double* c_out = nullptr) {
  const auto t0 = std::chrono::steady_clock::now();
  Engine* E = nullptr;
  CHK(solo_engine(h, &E));
  int id = -1;
  CHK(engine_admit(E, h, z0, lam, beta_out, cg_itmax, &id));
  E->inst[(size_t)id].c_out = c_out;
  const int mi = (int)std::min<int64_t>(max_iter, 1 << 30);
  int rc = engine_run<TC>(E, tol, mi, cg_atol, cg_rtol);
  if (rc) {
    engine_drop(E, id);
    return rc;
  }
  CHK(engine_harvest(E, id, info));
  info[6] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return engine_leave(E);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Gondzio formulation (src/gondzio_model.jl, src/gondzio_kkt.jl):
// x = [p; q; r], beta = p - q, c(x) = z0 - M(p - q) - r.  The operator work
// is the handle's own M / M^T / M^T M; the vector work is elementwise.
// ---------------------------------------------------------------------------
extern "C++" {

template <typename F>
__global__ __launch_bounds__(256) void k_each(int64_t n, F f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    f(i);
}

template <typename F>
static int each(csipm_op* h, int64_t n, F f) {
  hipLaunchKernelGGL((k_each<F>), dim3(grid_for(n, 4)), dim3(kNT), 0, h->stream, n, f);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gz_obj_parts(const T* __restrict__ x, int64_t n,
                                                      double* part) {
  double a[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double r = x[2 * n + i];
    a[0] += r * r;
    a[1] += (double)x[i] + (double)x[n + i];
  }
  const int ops[2] = {0, 0};
  store_partials<2>(part, a, ops);
}

// the handle's Gondzio workspace of at least nvec N-vectors of T, grown on
// demand and kept (the entry points then allocate nothing per call)
template <typename T>
static int gz_workspace(csipm_op* h, int nvec, T** out) {
  const size_t need = (size_t)nvec * sizeof(T) * (size_t)h->N;
  if (h->gz_bytes < need) {
    CHK(no_capture(h->stream, "a Gondzio entry point"));
    std::lock_guard<std::mutex> lk(g_capture_mu);
    if (h->gz) {
      HIPCHK(hipStreamSynchronize(h->stream));
      dfree(h->gz);
      h->gz = nullptr;
      h->gz_bytes = 0;
    }
    HIPCHK(dmalloc(&h->gz, need));
    h->gz_bytes = need;
  }
  *out = (T*)h->gz;
  return CSIPM_OK;
}

template <typename T>
static int gz_scratch(csipm_op* h, T** a, T** b) {
  T* w = nullptr;
  CHK(gz_workspace<T>(h, 3, &w));
  *a = w;
  *b = w + h->N;
  return CSIPM_OK;
}

template <typename T>
static int op_m(csipm_op* h, const T* v, T* y) {
  if (h->dim == 1) return apply_m_perp_beta<T, 1>(h, v, y);
  if (h->dim == 2) return apply_m_perp_beta<T, 2>(h, v, y);
  return apply_m_perp_beta<T, 3>(h, v, y);
}

template <typename T>
static int op_mt(csipm_op* h, const T* z, T* beta) {
  return dispatch_m_perpt_z<T, SP_BETA>(h, z, beta);
}

template <typename T>
static int gz_obj_t(csipm_op* h, const T* x, double lam, double* f) {
  const int gv = grid_for(h->N, 4);
  hipLaunchKernelGGL((k_gz_obj_parts<T>), dim3(gv), dim3(kNT), 0, h->stream, x, h->N, h->part);
  HIPCHK(hipGetLastError());
  std::vector<double> hp(2 * (size_t)gv);
  HIPCHK(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double rr = 0, sx = 0;
  for (int i = 0; i < gv; ++i) {
    rr += hp[(size_t)i];
    sx += hp[(size_t)gv + i];
  }
  *f = 0.5 * rr + lam * sx;
  return CSIPM_OK;
}

template <typename T>
static int gz_grad_t(csipm_op* h, const T* x, double lam, T* g) {
  const int64_t n = h->N;
  const T l = (T)lam;
  return each(h, n, [=] __device__(int64_t i) {
    g[i] = l;
    g[n + i] = l;
    g[2 * n + i] = x[2 * n + i];
  });
}

template <typename T>
static int gz_cons_t(csipm_op* h, const T* x, const T* z0, T* c) {
  const int64_t n = h->N;
  T *d, *md;
  CHK(gz_scratch<T>(h, &d, &md));
  CHK(each(h, n, [=] __device__(int64_t i) { d[i] = x[i] - x[n + i]; }));
  CHK(op_m<T>(h, d, md));
  return each(h, n, [=] __device__(int64_t i) { c[i] = z0[i] - md[i] - x[2 * n + i]; });
}

template <typename T>
static int gz_jtprod_t(csipm_op* h, const T* v, T* out) {
  const int64_t n = h->N;
  T *t, *unused;
  CHK(gz_scratch<T>(h, &t, &unused));
  CHK(op_mt<T>(h, v, t));
  return each(h, n, [=] __device__(int64_t i) {
    out[i] = -t[i];
    out[n + i] = t[i];
    out[2 * n + i] = -v[i];
  });
}

template <typename T>
static int gz_kkt_apply_t(csipm_op* h, const T* x, T* y, const T* Wp, const T* Wq, double alpha,
                          double beta) {
  const int64_t n = h->N;
  T *d, *m;
  CHK(gz_scratch<T>(h, &d, &m));
  CHK(each(h, n, [=] __device__(int64_t i) { d[i] = x[i] - x[n + i]; }));
  CHK((dispatch_mtm<T, SP_BETA>(h, d, epi_store<T>(m), nullptr)));
  const T a = (T)alpha, b = (T)beta;
  return each(h, n, [=] __device__(int64_t i) {
    T yp = a * (Wp[i] * x[i] + m[i]);
    T yq = a * (Wq[i] * x[n + i] - m[i]);
    if (b != (T)0) {
      yp += b * y[i];
      yq += b * y[n + i];
    }
    y[i] = yp;
    y[n + i] = yq;
  });
}

template <typename T>
static int gz_kkt_full_t(csipm_op* h, const T* x, T* y, double alpha, double beta) {
  const int64_t n = h->N;
  T *t, *s;
  CHK(gz_scratch<T>(h, &t, &s));
  CHK(op_mt<T>(h, x + 3 * n, t));
  // s = M (x_q - x_p), formed in y's r block's place only after t is taken
  T* dqp = s;
  CHK(each(h, n, [=] __device__(int64_t i) { dqp[i] = x[n + i] - x[i]; }));
  T* ms = s + n;   // the workspace's third vector
  int rc = op_m<T>(h, dqp, ms);
  const T a = (T)alpha, b = (T)beta;
  if (!rc)
    rc = each(h, n, [=] __device__(int64_t i) {
      const T xr = x[2 * n + i], xy = x[3 * n + i];
      T v0 = -a * t[i], v1 = a * t[i], v2 = a * (xr - xy), v3 = a * (ms[i] - xr);
      if (b != (T)0) {
        v0 += b * y[i];
        v1 += b * y[n + i];
        v2 += b * y[2 * n + i];
        v3 += b * y[3 * n + i];
      }
      y[i] = v0;
      y[n + i] = v1;
      y[2 * n + i] = v2;
      y[3 * n + i] = v3;
    });
  return rc;
}

// P = [[1 + Wp, -1], [-1, 1 + Wq]]^-1 in the closed form P11 = (1 + Wq)/det,
// P12 = 1/det, P22 = (1 + Wp)/det, det = Wp + Wq + Wp Wq: the matrix
// build_kkt! forms as inv = 1/(1 + Wp), S = 1/(1 + Wq - inv), P12 = inv S,
// P11 = inv + inv S inv, P22 = S, without the difference 1 + Wq - inv
template <typename T>
static int gz_precond_build_t(csipm_op* h, const T* Wp, const T* Wq, T* P11, T* P12, T* P22) {
  return each(h, h->N, [=] __device__(int64_t i) {
    const double wp = Wp[i], wq = Wq[i];
    const double det = wp + wq + wp * wq;
    P11[i] = (T)((1.0 + wq) / det);
    P12[i] = (T)(1.0 / det);
    P22[i] = (T)((1.0 + wp) / det);
  });
}

template <typename T>
static int gz_condense_t(csipm_op* h, const T* w, T* rhs) {
  const int64_t n = h->N;
  T *sum, *t;
  CHK(gz_scratch<T>(h, &sum, &t));
  CHK(each(h, n, [=] __device__(int64_t i) { sum[i] = w[2 * n + i] + w[3 * n + i]; }));
  CHK(op_mt<T>(h, sum, t));
  return each(h, n, [=] __device__(int64_t i) {
    rhs[i] = w[i] - t[i];
    rhs[n + i] = w[n + i] + t[i];
  });
}

template <typename T>
static int gz_unpack_t(csipm_op* h, T* w, const T* dx) {
  const int64_t n = h->N;
  T *d, *s;
  CHK(gz_scratch<T>(h, &d, &s));
  CHK(each(h, n, [=] __device__(int64_t i) { d[i] = dx[i] - dx[n + i]; }));
  CHK(op_m<T>(h, d, s));
  return each(h, n, [=] __device__(int64_t i) {
    const T wr = w[2 * n + i];
    const T dr = -w[3 * n + i] - s[i];
    w[i] = dx[i];
    w[n + i] = dx[n + i];
    w[2 * n + i] = dr;
    w[3 * n + i] = dr - wr;
  });
}

// PCG on the condensed Gondzio K by the congruence x = A y, A = 1/2 [[I, I],
// [-I, I]]: A^T K A is the condensed FFT system with L1 = (Wp + Wq)/4,
// L2 = (Wp - Wq)/4, right-hand side A^T b, preconditioner A^-1 P A^-T =
// [[P11 - 2 P12 + P22, P11 - P22], [P11 - P22, P11 + 2 P12 + P22]].  PCG is
// invariant under the change of variables (same iterates x_k = A y_k, same
// preconditioned residual norms, so the same stopping point).
template <typename T>
static int gz_cg_t(csipm_op* h, const T* b, T* x, const T* Wp, const T* Wq, const T* P11,
                   const T* P12, const T* P22, double atol, double rtol, int64_t itmax,
                   int64_t* iters, double* resid, int* status) {
  const int64_t n = h->N;
  T* buf = nullptr;
  CHK(gz_workspace<T>(h, 9, &buf));
  T *bt = buf, *yt = buf + 2 * n, *L1 = buf + 4 * n, *L2 = buf + 5 * n, *Q11 = buf + 6 * n,
    *Q12 = buf + 7 * n, *Q22 = buf + 8 * n;
  int rc = each(h, n, [=] __device__(int64_t i) {
    bt[i] = (T)0.5 * (b[i] - b[n + i]);
    bt[n + i] = (T)0.5 * (b[i] + b[n + i]);
    L1[i] = (Wp[i] + Wq[i]) * (T)0.25;
    L2[i] = (Wp[i] - Wq[i]) * (T)0.25;
    Q11[i] = P11[i] - (T)2 * P12[i] + P22[i];
    Q12[i] = P11[i] - P22[i];
    Q22[i] = P11[i] + (T)2 * P12[i] + P22[i];
  });
  if (!rc)
    rc = cg_solve_solo<T>(h, bt, yt, L1, L2, Q11, Q12, Q22, atol, rtol, itmax, iters, resid,
                          status);
  if (!rc)
    rc = each(h, n, [=] __device__(int64_t i) {
      const T yb = yt[i], yc = yt[n + i];
      x[i] = (T)0.5 * (yb + yc);
      x[n + i] = (T)0.5 * (yc - yb);
    });
  if (!rc) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(CSIPM_ERR_HIP, "gondzio cg: %s", hipGetErrorString(e));
  }
  return rc;
}

}  // extern "C++"

extern "C" {

const char* csipm_version(void) { return CSIPM_VERSION; }

const char* csipm_last_error(void) { return g_err.c_str(); }

static void destroy_impl(csipm_op_t h);

int csipm_op_destroy(csipm_op_t h) {
  if (!h) return CSIPM_OK;
  if (h->admitted.load() > 0)
    return fail(CSIPM_ERR_ARG, "operator handle is admitted to a live batch instance "
                "(harvest it or destroy the batch first)");
  CHK(no_capture(h->stream, "csipm_op_destroy"));
  destroy_impl(h);
  return CSIPM_OK;
}

static void destroy_impl(csipm_op_t h) {
  // twiddle tables (tw, twA, twB) are shared across handles: released to
  // the cache; the Bluestein tables and the rest are the handle's own
  for (PrecState* ps : {&h->p64, &h->p32}) {
    for (int i = 0; i < 3; ++i) release_twiddles(ps->tw[i]);
    release_twiddles(ps->twA);
    release_twiddles(ps->twB);
  }
  // work of this handle still in flight on its stream finishes before its
  // blocks go back to the cache (engines synchronise their own stream)
  if (h->solo) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    delete h->solo;   // synchronises its own stream
  }
  std::vector<void*> blocks;
  for (PrecState* ps : {&h->p64, &h->p32})
    for (int i = 0; i < 3; ++i)
      for (void* b : {ps->chirp[i], ps->bfw[i], ps->binv[i]})
        if (b) blocks.push_back(b);
  for (int i = 0; i < 3; ++i)
    if (h->rperm[i]) blocks.push_back(h->rperm[i]);
  for (void* b : {(void*)h->miss, (void*)h->valid, h->buf, (void*)h->part, h->scratch,
                  (void*)h->gz, (void*)h->dplan})
    if (b) blocks.push_back(b);
  // work of this handle still in flight on its stream finishes before its
  // blocks go back to the cache (engines synchronise their own stream): at
  // once when the stream is idle, else once an event recorded now completes
  hipEvent_t ev = nullptr;
  if (hipStreamQuery(h->stream) == hipSuccess) {
    for (void* b : blocks) dfree_release(b);
  } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
             hipEventRecord(ev, h->stream) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_pend_mu);
    g_pend.push_back({ev, std::move(blocks)});
  } else {
    if (ev) (void)hipEventDestroy(ev);
    (void)hipGetLastError();
    (void)hipStreamSynchronize(h->stream);
    for (void* b : blocks) dfree_release(b);
  }
  delete h;
  release_pending();
}

// 1D lines above this length run as four-step transforms
constexpr int kBig1D = 4096;

// four-step plan of a long 1D line: F = N when N factors into the supported
// radices (A x B a balanced divisor pair, both <= 4096), else Bluestein with
// the power-of-two F >= 2N - 1 (host planning only)
static int plan_big1d(csipm_op* h) {
  const int N = h->n[0];
  int F = N, A = 0, B = 0;
  FftPlan probe{};
  h->obs = !factorize(N, probe);
  if (h->obs) {
    F = 1;
    while (F < 2 * N - 1) F <<= 1;
  }
  // the most balanced divisor pair with both factors <= kNT * kE
  for (int b = 1; (int64_t)b * b <= F; ++b) {
    if (F % b) continue;
    const int a = F / b;
    if (a <= kNT * kE && b <= kNT * kE) {
      A = a;
      B = b;
    }
  }
  if (A == 0 || !factorize(A, h->planA) || !factorize(B, h->planB))
    return fail(CSIPM_ERR_UNSUPPORTED, "no four-step plan for length %d", N);
  h->big = true;
  h->bA = A;
  h->bB = B;
  h->bF = F;
  h->plan[0].L = F;
  return CSIPM_OK;
}

// device tables of a four-step plan: twiddles of A, B and F, Bluestein tables
static int upload_big1d(csipm_op* h) {
  const int N = h->n[0], A = h->bA, B = h->bB, F = h->bF;
  for (PrecState* ps : {&h->p64, &h->p32}) {
    const bool want = ps == &h->p64 ? h->has64() : h->has32();
    if (!want) continue;
    if (ps == &h->p64) {
      CHK(upload_twiddles<double>(A, &ps->twA));
      CHK(upload_twiddles<double>(B, &ps->twB));
      CHK(upload_twiddles<double>(F, &ps->tw[0]));
      if (h->obs) CHK(upload_bluestein<double>(N, F, &ps->chirp[0], &ps->bfw[0], &ps->binv[0]));
    } else {
      CHK(upload_twiddles<float>(A, &ps->twA));
      CHK(upload_twiddles<float>(B, &ps->twB));
      CHK(upload_twiddles<float>(F, &ps->tw[0]));
      if (h->obs) CHK(upload_bluestein<float>(N, F, &ps->chirp[0], &ps->bfw[0], &ps->binv[0]));
    }
  }
  return CSIPM_OK;
}

static int set_shapes(csipm_op* h, PrecState& ps, size_t cs) {
  const int dim = h->dim;
  if (h->big) {
    ps.col = col_cfg(h->bA, h->bB, cs);
    ps.row = row_cfg(h->bA, h->bB, cs);
    if (ps.col.lds > kLdsMax || ps.row.lds > kLdsMax)
      return fail(CSIPM_ERR_UNSUPPORTED, "four-step plan of length %d exceeds the LDS", h->n[0]);
    return CSIPM_OK;
  }
  // FFT length W (line length, Bluestein's M, Rader's L - 1) and LDS rows of a
  // dim-1 line (L, M, Rader's 2L + 1)
  const int F0 = h->plan[0].L;
  ps.d1 = dim1_cfg(F0, h->rows[0], (int)(h->N / h->n[0]), cs);
  if (ps.d1.nt * kE < F0 || ps.d1.lds > kLdsMax)
    return fail(CSIPM_ERR_UNSUPPORTED, "dimension-1 length %d (FFT length %d) exceeds one "
                "workgroup", h->n[0], F0);
  if (dim >= 2) ps.lines[1] = line_cfg(h->H1, h->plan[1].L, h->rows[1], h->n[2], cs);
  if (dim == 3) ps.lines[2] = line_cfg(h->H1, h->plan[2].L, h->rows[2], h->n[1], cs);
  for (int ax = 1; ax < dim; ++ax)
    if (h->plan[ax].L > kNT * kE || ps.lines[ax].lds > kLdsMax)
      return fail(CSIPM_ERR_UNSUPPORTED, "dimension-%d length %d (%d LDS rows) exceeds one "
                  "workgroup", ax + 1, h->n[ax], h->rows[ax]);
  return CSIPM_OK;
}

// host-side planning of a handle: sizes, map geometry, FFT plans, launch
// shapes and partial-sum capacity.  No device work (csipm_op_footprint runs
// it alone).
static int plan_impl(csipm_op* h, int dim, const int64_t sizes[3], int rdft, int precision) {
  if (dim < 1 || dim > 3) return fail(CSIPM_ERR_ARG, "dim must be 1, 2 or 3 (got %d)", dim);
  if (!sizes) return fail(CSIPM_ERR_ARG, "sizes is null");
  // rdft = 0 (the reference's complex path) computes the same operator: the
  // complex transforms of a Hermitian spectrum and of real samples equal the
  // C2R / R2C ones, so both run the real-to-complex kernels
  if (rdft != 0 && rdft != 1) return fail(CSIPM_ERR_ARG, "rdft must be 0 or 1");
  h->rdft = rdft;
  if (precision != CSIPM_F64 && precision != CSIPM_F32 && precision != CSIPM_MIXED)
    return fail(CSIPM_ERR_ARG, "bad precision %d", precision);
  h->dim = dim;
  h->prec = precision;
  // 1D lines up to 2^23 (four-step, Bluestein length <= 2^24); 2D / 3D axes up to 16384
  const int64_t smax = dim == 1 ? ((int64_t)1 << 23) : 16384;
  for (int i = 0; i < dim; ++i) {
    if (sizes[i] < 2 || sizes[i] > smax)
      return fail(CSIPM_ERR_SHAPE, "size %d = %lld out of range [2, %lld]", i + 1,
                  (long long)sizes[i], (long long)smax);
    h->n[i] = (int)sizes[i];
  }
  h->g = make_map_geom(h->dim, h->n);
  h->N = h->g.N;
  if (h->N > ((int64_t)1 << 31)) return fail(CSIPM_ERR_SHAPE, "N too large");
  h->H1 = h->n[0] / 2 + 1;
  h->Nc = (int64_t)h->H1 * h->n[1] * h->n[2];
  h->U = 2 * h->Nc;
  if (dim == 1 && h->n[0] > kBig1D) CHK(plan_big1d(h));
  for (int i = 0; i < dim && !h->big; ++i)
    if (!plan_axis(h->n[i], h->plan[i], h->bs[i], h->rows[i]))
      return fail(CSIPM_ERR_UNSUPPORTED, "no FFT plan for size %d", h->n[i]);
  if (h->has64()) CHK(set_shapes(h, h->p64, sizeof(cpx<double>)));
  if (h->has32()) CHK(set_shapes(h, h->p32, sizeof(cpx<float>)));
  int np = 4096;
  for (PrecState* ps : {&h->p64, &h->p32}) {
    np = std::max(np, ps->d1.grid);
    for (int ax = 1; ax < dim; ++ax) np = std::max(np, ps->lines[ax].grid);
  }
  h->npart = 8 * np;   // up to 8 reduced quantities per pass
  return CSIPM_OK;
}

// device bytes a handle holds itself: plans, masks, the scratch spectrum,
// partial sums, Bluestein tables and the lazily made entry-point scratch
// (twiddle tables are shared between handles and cached: not counted)
static size_t handle_own_bytes(const csipm_op* h) {
  size_t b = 5 * sizeof(FftPlan);
  if (h->nmiss > 0) b += sizeof(uint32_t) * (size_t)((h->N + 31) / 32);
  b += sizeof(uint32_t) * (size_t)((h->U + 31) / 32);
  b += (size_t)std::max<int64_t>(h->Nc, h->bF) * sizeof(cpx<double>);
  b += sizeof(double) * (size_t)h->npart;
  for (int p = 0; p < 2; ++p) {
    if (!(p == 0 ? h->has64() : h->has32())) continue;
    const size_t cs = p == 0 ? sizeof(cpx<double>) : sizeof(cpx<float>);
    if (h->big) {
      if (h->obs) b += cs * ((size_t)h->n[0] + 2 * (size_t)h->bF);
    } else {
      for (int i = 0; i < h->dim; ++i) {
        if (h->bs[i] == 1) b += cs * ((size_t)h->n[i] + 2 * (size_t)h->plan[i].L);
        if (h->bs[i] == 2) b += cs * 2 * (size_t)h->plan[i].L + 8 * (size_t)h->n[i];
      }
    }
  }
  return b + h->scratch_bytes + h->gz_bytes;
}

static int create_impl(csipm_op* h, int dim, const int64_t sizes[3], int rdft, int precision,
                       const int64_t* missing, int64_t n_missing, void* stream) {
  CHK(plan_impl(h, dim, sizes, rdft, precision));
  h->stream = (hipStream_t)stream;
  CHK(no_capture(h->stream, "csipm_op_create"));
  release_pending();   // blocks of destroyed handles whose work has finished
  if (n_missing < 0) return fail(CSIPM_ERR_ARG, "n_missing < 0");
  if (n_missing > 0 && !missing) return fail(CSIPM_ERR_ARG, "missing is null");
  if (h->big) CHK(upload_big1d(h));
  for (int i = 0; i < dim && !h->big; ++i) {
    const int F = h->plan[i].L;
    if (h->has64()) CHK(upload_twiddles<double>(F, &h->p64.tw[i]));
    if (h->has32()) CHK(upload_twiddles<float>(F, &h->p32.tw[i]));
    if (h->bs[i] == 2) {
      CHK(upload_rader_perm(h->n[i], &h->rperm[i]));
      if (h->has64()) CHK(upload_rader<double>(h->n[i], &h->p64.bfw[i], &h->p64.binv[i]));
      if (h->has32()) CHK(upload_rader<float>(h->n[i], &h->p32.bfw[i], &h->p32.binv[i]));
    }
    if (h->bs[i] == 1) {
      if (h->has64())
        CHK(upload_bluestein<double>(h->n[i], F, &h->p64.chirp[i], &h->p64.bfw[i],
                                     &h->p64.binv[i]));
      if (h->has32())
        CHK(upload_bluestein<float>(h->n[i], F, &h->p32.chirp[i], &h->p32.bfw[i],
                                    &h->p32.binv[i]));
    }
  }
  {
    const FftPlan plans[5] = {h->plan[0], h->plan[1], h->plan[2], h->planA, h->planB};
    HIPL(dmalloc((void**)&h->dplan, sizeof(plans)));
    HIPL(hipMemcpy(h->dplan, plans, sizeof(plans), hipMemcpyHostToDevice));
  }
  // missing-sample bitmask
  h->nmiss = n_missing;
  if (n_missing > 0) {
    const int64_t words = (h->N + 31) / 32;
    std::vector<uint32_t> bits((size_t)words, 0u);
    for (int64_t i = 0; i < n_missing; ++i) {
      const int64_t s = missing[i];
      if (s < 0 || s >= h->N)
        return fail(CSIPM_ERR_SHAPE, "missing index %lld outside [0, %lld)", (long long)s,
                    (long long)h->N);
      bits[(size_t)(s >> 5)] |= 1u << (s & 31);
    }
    HIPL(dmalloc((void**)&h->miss, sizeof(uint32_t) * (size_t)words));
    HIPL(hipMemcpy(h->miss, bits.data(), sizeof(uint32_t) * (size_t)words,
                     hipMemcpyHostToDevice));
  }
  // u-slot validity
  const int64_t vwords = (h->U + 31) / 32;
  HIPL(dmalloc((void**)&h->valid, sizeof(uint32_t) * (size_t)vwords));
  {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    const int grid = (int)((vwords + kNT - 1) / kNT);
    if (dim == 1)
      hipLaunchKernelGGL((k_u_valid<1>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    else if (dim == 2)
      hipLaunchKernelGGL((k_u_valid<2>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    else
      hipLaunchKernelGGL((k_u_valid<3>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    HIPCHK(hipGetLastError());
  }
  HIPL(dmalloc(&h->buf, (size_t)std::max<int64_t>(h->Nc, h->bF) * sizeof(cpx<double>)));
  HIPL(dmalloc((void**)&h->part, sizeof(double) * (size_t)h->npart));
  HIPL(hipStreamSynchronize(h->stream));
  return CSIPM_OK;
}

int csipm_op_create(csipm_op_t* out, int dim, const int64_t sizes[3], int rdft, int precision,
                    const int64_t* missing, int64_t n_missing, void* stream) {
  if (!out) return fail(CSIPM_ERR_ARG, "handle pointer is null");
  *out = nullptr;
  csipm_op* h = new (std::nothrow) csipm_op();
  if (!h) return fail(CSIPM_ERR_ALLOC, "out of host memory");
  const int rc = create_impl(h, dim, sizes, rdft, precision, missing, n_missing, stream);
  if (rc != CSIPM_OK) {
    std::string keep = g_err;
    destroy_impl(h);
    g_err = keep;
    return rc;
  }
  *out = h;
  return CSIPM_OK;
}

int csipm_op_set_stream(csipm_op_t h, void* stream) {
  if (!h) return fail(CSIPM_ERR_ARG, "null handle");
  h->stream = (hipStream_t)stream;
  return CSIPM_OK;
}

int csipm_op_bytes(csipm_op_t h, int64_t* bytes) {
  if (!h || !bytes) return fail(CSIPM_ERR_ARG, "null argument");
  size_t b = handle_own_bytes(h);
  if (h->solo) b += engine_bytes_per_slot(h->solo->tc_size) * (size_t)h->solo->cap;
  *bytes = (int64_t)b;
  return CSIPM_OK;
}

int csipm_op_footprint(int dim, const int64_t sizes[3], int precision, int64_t n_missing,
                       int64_t* slots, int64_t* bytes) {
  if (!slots || !bytes || n_missing < 0) return fail(CSIPM_ERR_ARG, "bad argument");
  csipm_op tmp;
  CHK(plan_impl(&tmp, dim, sizes, 1, precision));
  tmp.nmiss = n_missing;
  *slots = slots_of(&tmp);
  *bytes = (int64_t)handle_own_bytes(&tmp);
  return CSIPM_OK;
}

int csipm_op_fft_shape(csipm_op_t h, int axis, int* fft_len, int* rows, int* mode) {
  if (!h || !fft_len || !rows || !mode || axis < 0 || axis >= h->dim)
    return fail(CSIPM_ERR_ARG, "bad argument");
  if (h->big) {
    *fft_len = h->bF;
    *rows = h->bF;
    *mode = 3;
    return CSIPM_OK;
  }
  *fft_len = h->plan[axis].L;
  *rows = h->rows[axis];
  *mode = h->bs[axis];
  return CSIPM_OK;
}

int csipm_tune_set(int kind, int elem_bytes, int fft_len, int rows, int lg_cap, int nt) {
  if ((kind != 0 && kind != 1) || (elem_bytes != 8 && elem_bytes != 16) || fft_len < 1 ||
      rows < fft_len || (nt != 0 && nt != kNTSmall && nt != kNT))
    return fail(CSIPM_ERR_ARG, "bad tuning entry");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const auto key = std::make_tuple(kind, elem_bytes, fft_len, rows);
  if (lg_cap < 0) {
    g_tune.erase(key);
    return CSIPM_OK;
  }
  TuneVal v;
  v.lg_cap = lg_cap;
  v.nt = nt;
  g_tune[key] = v;
  return CSIPM_OK;
}

int csipm_tune_clear(void) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
  }
  std::lock_guard<std::mutex> lk(g_blue_mu);
  g_blue_m.clear();
  return CSIPM_OK;
}

int csipm_tune_bluestein(int line_len, int conv_len) {
  if (line_len < 3 || line_len > 16384) return fail(CSIPM_ERR_ARG, "bad line length %d", line_len);
  std::lock_guard<std::mutex> lk(g_blue_mu);
  if (conv_len == 0) {
    g_blue_m.erase(line_len);
    return CSIPM_OK;
  }
  FftPlan p;
  if (conv_len < 2 * line_len - 1 || conv_len > 4 * line_len || !factorize(conv_len, p))
    return fail(CSIPM_ERR_ARG,
                "Bluestein length %d does not serve line length %d (needs 2L - 1 <= M <= 4L "
                "and a factorable M)", conv_len, line_len);
  g_blue_m[line_len] = conv_len;
  return CSIPM_OK;
}

int csipm_op_info(csipm_op_t h, int64_t* n, int64_t* nc, int* precision) {
  if (!h) return fail(CSIPM_ERR_ARG, "null handle");
  if (n) *n = h->N;
  if (nc) *nc = h->Nc;
  if (precision) *precision = h->prec;
  return CSIPM_OK;
}

// operator entry points run in the handle's operator precision: double for
// CSIPM_F64, float for CSIPM_F32 and CSIPM_MIXED
#define OP_DOUBLE(h) ((h)->prec == CSIPM_F64)

int csipm_beta_to_dft(csipm_op_t h, const double* beta, double* v, int full) {
  if (!h || !beta || !v) return fail(CSIPM_ERR_ARG, "null argument");
  const int ld1 = full ? h->n[0] : h->H1;
  const int64_t total = (int64_t)ld1 * h->n[1] * h->n[2];
  const int grid = (int)((total + kNT - 1) / kNT);
  cpx<double>* vv = (cpx<double>*)v;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_beta_to_dft<1>), dim3(grid), dim3(kNT), 0, h->stream, beta, vv, h->g, ld1, total);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_beta_to_dft<2>), dim3(grid), dim3(kNT), 0, h->stream, beta, vv, h->g, ld1, total);
  else
    hipLaunchKernelGGL((k_beta_to_dft<3>), dim3(grid), dim3(kNT), 0, h->stream, beta, vv, h->g, ld1, total);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_dft_to_beta(csipm_op_t h, const double* v, double* beta, int full) {
  if (!h || !beta || !v) return fail(CSIPM_ERR_ARG, "null argument");
  const int ld1 = full ? h->n[0] : h->H1;
  const int grid = (int)((h->N + kNT - 1) / kNT);
  const cpx<double>* vv = (const cpx<double>*)v;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_dft_to_beta<1>), dim3(grid), dim3(kNT), 0, h->stream, vv, beta, h->g, ld1);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_dft_to_beta<2>), dim3(grid), dim3(kNT), 0, h->stream, vv, beta, h->g, ld1);
  else
    hipLaunchKernelGGL((k_dft_to_beta<3>), dim3(grid), dim3(kNT), 0, h->stream, vv, beta, h->g, ld1);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_m_perp_beta(csipm_op_t h, const void* beta, void* y) {
  if (!h || !beta || !y) return fail(CSIPM_ERR_ARG, "null argument");
  if (OP_DOUBLE(h)) {
    if (h->dim == 1) return apply_m_perp_beta<double, 1>(h, (const double*)beta, (double*)y);
    if (h->dim == 2) return apply_m_perp_beta<double, 2>(h, (const double*)beta, (double*)y);
    return apply_m_perp_beta<double, 3>(h, (const double*)beta, (double*)y);
  }
  if (h->dim == 1) return apply_m_perp_beta<float, 1>(h, (const float*)beta, (float*)y);
  if (h->dim == 2) return apply_m_perp_beta<float, 2>(h, (const float*)beta, (float*)y);
  return apply_m_perp_beta<float, 3>(h, (const float*)beta, (float*)y);
}

namespace csipm {

// ---------------------------------------------------------------------------
// index_missing from a sampling mask / weights array (zero = not observed):
// the 0-based linear indices of its zero entries in increasing order, by a
// count pass, an exclusive scan of the per-block counts and an ordered write
// (wave ballots for the order inside a chunk of 256 entries).
// ---------------------------------------------------------------------------
constexpr int kMaskChunks = 16;   // chunks of 256 entries per block
static_assert(CSIPM_MASK_NOALLOC_MAX == (int64_t)(CSIPM_MASK_NOALLOC_WORDS - 1) * 256 * kMaskChunks,
              "csipm.h's sample limit of the no-alloc mask entry point");

__global__ __launch_bounds__(256) void k_mask_count(const double* __restrict__ w, int64_t n,
                                                    int64_t* __restrict__ counts) {
  const int64_t base = (int64_t)blockIdx.x * 256 * kMaskChunks;
  double c = 0.0;
#pragma unroll
  for (int j = 0; j < kMaskChunks; ++j) {
    const int64_t e = base + j * 256 + threadIdx.x;
    if (e < n && w[e] == 0.0) c += 1.0;
  }
  const double s = block_reduce(c, OpSum());
  if (threadIdx.x == 0) counts[blockIdx.x] = (int64_t)s;
}

__global__ __launch_bounds__(256) void k_mask_write(const double* __restrict__ w, int64_t n,
                                                    const int64_t* __restrict__ offsets,
                                                    int64_t* __restrict__ out) {
  __shared__ int wave_tot[4];
  const int64_t base = (int64_t)blockIdx.x * 256 * kMaskChunks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t run = offsets[blockIdx.x];
  for (int j = 0; j < kMaskChunks; ++j) {
    const int64_t e = base + j * 256 + threadIdx.x;
    const bool z = e < n && w[e] == 0.0;
    const uint64_t bal = __ballot(z);
    const int below = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) wave_tot[wave] = __popcll(bal);
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < wave) woff += wave_tot[q];
      tot += wave_tot[q];
    }
    if (z) out[run + woff + below] = e;
    run += tot;
    __syncthreads();
  }
}

// exclusive scan of c[0, nb) in place by one workgroup of 1024 threads, the
// total into c[nb]: thread t scans a contiguous run of ceil(nb / 1024)
// entries, the runs' sums are scanned across the workgroup
__global__ __launch_bounds__(1024) void k_mask_scan(int64_t* __restrict__ c, int64_t nb) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nb + 1023) / 1024;
  const int64_t lo0 = (int64_t)t * per;
  const int64_t lo = lo0 < nb ? lo0 : nb;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  int64_t sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += c[i];
  part[t] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan of the run sums
  for (int d = 1; d < 1024; d <<= 1) {
    const int64_t v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;   // exclusive prefix of this run
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t v = c[i];
    c[i] = run;
    run += v;
  }
  if (t == 1023) c[nb] = part[1023];
}

// count, scan and ordered write of the zero entries of w[0, n) into out; ws
// holds nb + 1 int64 (nb = blocks of 256 * kMaskChunks entries); the total
// is left in ws[nb].  Asynchronous on st.
static int mask_index_launch(const double* w, int64_t n, int64_t* out, int64_t* ws, int64_t nb,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_mask_count, dim3((unsigned)nb), dim3(256), 0, st, w, n, ws);
  hipLaunchKernelGGL(k_mask_scan, dim3(1), dim3(1024), 0, st, ws, nb);
  hipLaunchKernelGGL(k_mask_write, dim3((unsigned)nb), dim3(256), 0, st, w, n, ws, out);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

}  // namespace csipm

int csipm_reconstruct(csipm_op_t h, const void* beta, void* x) {
  if (!h || !beta || !x) return fail(CSIPM_ERR_ARG, "null argument");
  // M_perp_beta of the same handle with no sample missing: the pass
  // arguments are taken at launch, so the mask is restored right after
  uint32_t* miss = h->miss;
  h->miss = nullptr;
  const int rc = csipm_m_perp_beta(h, beta, x);
  h->miss = miss;
  return rc;
}

// the count workspace of csipm_mask_index_missing: one per device, grown on
// demand and kept (calls are serialised by the capture lock)
static int mask_workspace(int64_t words, int64_t** out) {
  static std::vector<std::pair<void*, int64_t>> ws;   // by device: (pointer, words)
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if ((int)ws.size() <= dev) ws.resize((size_t)dev + 1, {nullptr, 0});
  auto& w = ws[(size_t)dev];
  if (w.second < words) {
    if (w.first) HIPCHK(hipFree(w.first));
    w = {nullptr, 0};
    const int64_t grow = std::max<int64_t>(words, 4096);
    HIPCHK(hipMalloc(&w.first, sizeof(int64_t) * (size_t)grow));
    w.second = grow;
  }
  *out = (int64_t*)w.first;
  return CSIPM_OK;
}

int csipm_mask_index_missing(const double* weights, int64_t n, int64_t* index_out,
                             int64_t* count, void* stream) {
  if (!count || n < 0 || (n > 0 && (!weights || !index_out)))
    return fail(CSIPM_ERR_ARG, "null argument or negative size");
  *count = 0;
  if (n == 0) return CSIPM_OK;
  hipStream_t st = (hipStream_t)stream;
  CHK(no_capture(st, "csipm_mask_index_missing"));
  std::lock_guard<std::mutex> lk(g_capture_mu);   // may allocate; synchronises
  const int64_t per = 256 * (int64_t)kMaskChunks;
  const int64_t nb = (n + per - 1) / per;
  if (nb > INT32_MAX) return fail(CSIPM_ERR_ARG, "mask too large");
  int64_t* d = nullptr;
  CHK(mask_workspace(nb + 1, &d));
  CHK(mask_index_launch(weights, n, index_out, d, nb, st));
  int64_t tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, d + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *count = tot;
  return CSIPM_OK;
}

// The count workspace of csipm_mask_index_missing_noalloc: a fixed array in
// each device's global memory, held until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol): the device address of this device's copy is looked up once per
// device with hipGetSymbolAddress and kept in g_mask_scratch_ptr.
static __device__ __align__(256) int64_t scratch[CSIPM_MASK_NOALLOC_WORDS];

static std::mutex g_mask_scratch_mu;
static std::vector<int64_t*> g_mask_scratch_ptr;   // by device; nullptr until looked up

static int mask_scratch_ptr(int64_t** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mask_scratch_mu);
  if ((int)g_mask_scratch_ptr.size() <= dev) g_mask_scratch_ptr.resize((size_t)dev + 1, nullptr);
  int64_t*& p = g_mask_scratch_ptr[(size_t)dev];
  if (!p) {
    void* a = nullptr;
    HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    if (!a) return fail(CSIPM_ERR_HIP, "no device address for the mask scratch array");
    p = (int64_t*)a;
  }
  *out = p;
  return CSIPM_OK;
}

int csipm_mask_index_missing_noalloc(const double* weights, int64_t n, int64_t* index_out,
                                     int64_t* count, void* stream) {
  if (!count || n < 0 || (n > 0 && (!weights || !index_out)))
    return fail(CSIPM_ERR_ARG, "null argument or negative size");
  *count = 0;
  const int64_t per = 256 * (int64_t)kMaskChunks;
  if (n > CSIPM_MASK_NOALLOC_MAX)
    return fail(CSIPM_ERR_ARG,
                "csipm_mask_index_missing_noalloc: n = %lld exceeds its maximum of %lld samples "
                "(use csipm_mask_index_missing)",
                (long long)n, (long long)CSIPM_MASK_NOALLOC_MAX);
  if (n == 0) return CSIPM_OK;
  hipStream_t st = (hipStream_t)stream;
  CHK(no_capture(st, "csipm_mask_index_missing_noalloc"));
  const int64_t nb = (n + per - 1) / per;   // nb + 1 <= CSIPM_MASK_NOALLOC_WORDS
  int64_t* d = nullptr;
  CHK(mask_scratch_ptr(&d));
  std::lock_guard<std::mutex> lk(g_capture_mu);   // synchronises
  CHK(mask_index_launch(weights, n, index_out, d, nb, st));
  int64_t tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, d + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *count = tot;
  return CSIPM_OK;
}

int csipm_m_perpt_z(csipm_op_t h, const void* z, void* beta) {
  if (!h || !beta || !z) return fail(CSIPM_ERR_ARG, "null argument");
  if (OP_DOUBLE(h)) return dispatch_m_perpt_z<double, SP_BETA>(h, (const double*)z, (double*)beta);
  return dispatch_m_perpt_z<float, SP_BETA>(h, (const float*)z, (float*)beta);
}

int csipm_mtm(csipm_op_t h, const void* x, void* y) {
  if (!h || !x || !y) return fail(CSIPM_ERR_ARG, "null argument");
  if (OP_DOUBLE(h))
    return dispatch_mtm<double, SP_BETA>(h, (const double*)x, epi_store<double>((double*)y), nullptr);
  return dispatch_mtm<float, SP_BETA>(h, (const float*)x, epi_store<float>((float*)y), nullptr);
}

int csipm_kkt_apply(csipm_op_t h, const void* x, void* y, const void* L1, const void* L2,
                    double alpha, double beta) {
  if (!h || !x || !y || !L1 || !L2) return fail(CSIPM_ERR_ARG, "null argument");
  if (OP_DOUBLE(h))
    return kkt_apply_t<double, SP_BETA>(h, (const double*)x, (double*)y, (const double*)L1,
                                        (const double*)L2, alpha, beta, nullptr, nullptr);
  return kkt_apply_t<float, SP_BETA>(h, (const float*)x, (float*)y, (const float*)L1,
                                     (const float*)L2, alpha, beta, nullptr, nullptr);
}

int csipm_precond_build(csipm_op_t h, const void* s1, const void* s2, void* L1, void* L2,
                        void* P11, void* P12, void* P22) {
  if (!h || !s1 || !s2 || !L1 || !L2 || !P11 || !P12 || !P22)
    return fail(CSIPM_ERR_ARG, "null argument");
  const int grid = (int)((h->N + kNT - 1) / kNT);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_precond_build<double>), dim3(grid), dim3(kNT), 0, h->stream,
                       (const double*)s1, (const double*)s2, (double*)L1, (double*)L2,
                       (double*)P11, (double*)P12, (double*)P22, h->N);
  else
    hipLaunchKernelGGL((k_precond_build<float>), dim3(grid), dim3(kNT), 0, h->stream,
                       (const float*)s1, (const float*)s2, (float*)L1, (float*)L2, (float*)P11,
                       (float*)P12, (float*)P22, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_precond_apply(csipm_op_t h, const void* x, void* y, const void* P11, const void* P12,
                        const void* P22, double alpha, double beta) {
  if (!h || !x || !y || !P11 || !P12 || !P22) return fail(CSIPM_ERR_ARG, "null argument");
  const int grid = (int)((h->N + kNT - 1) / kNT);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_precond_apply<double>), dim3(grid), dim3(kNT), 0, h->stream,
                       (const double*)x, (double*)y, (const double*)P11, (const double*)P12,
                       (const double*)P22, alpha, beta, h->N);
  else
    hipLaunchKernelGGL((k_precond_apply<float>), dim3(grid), dim3(kNT), 0, h->stream,
                       (const float*)x, (float*)y, (const float*)P11, (const float*)P12,
                       (const float*)P22, (float)alpha, (float)beta, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_cg_solve(csipm_op_t h, const void* b, void* x, const void* L1, const void* L2,
                   const void* P11, const void* P12, const void* P22, double atol, double rtol,
                   int64_t itmax, int64_t* iters, double* resid, int* status) {
  if (!h || !b || !x || !L1 || !L2 || !P11 || !P12 || !P22 || !iters || !resid || !status)
    return fail(CSIPM_ERR_ARG, "null argument");
  if (OP_DOUBLE(h))
    return cg_solve_solo<double>(h, (const double*)b, (double*)x, (const double*)L1,
                                 (const double*)L2, (const double*)P11, (const double*)P12,
                                 (const double*)P22, atol, rtol, itmax, iters, resid, status);
  return cg_solve_solo<float>(h, (const float*)b, (float*)x, (const float*)L1, (const float*)L2,
                              (const float*)P11, (const float*)P12, (const float*)P22, atol,
                              rtol, itmax, iters, resid, status);
}

int csipm_solve(csipm_op_t h, const double* z0, double lambda, double tol, int64_t max_iter,
                double cg_atol, double cg_rtol, int64_t cg_itmax, double* beta_out,
                double* info) {
  if (!h || !z0 || !beta_out || !info) return fail(CSIPM_ERR_ARG, "null argument");
  if (!(lambda >= 0) || !(tol > 0) || max_iter < 0)
    return fail(CSIPM_ERR_ARG, "bad lambda / tol / max_iter");
  if (!(cg_atol >= 0) || !(cg_rtol >= 0) || !(cg_atol > 0 || cg_rtol > 0))
    return fail(CSIPM_ERR_ARG, "bad cg_atol / cg_rtol");
  if (h->prec == CSIPM_F64)
    return solve_solo<double>(h, z0, lambda, tol, max_iter, cg_atol, cg_rtol, cg_itmax, beta_out,
                              info);
  if (h->prec == CSIPM_MIXED)
    return solve_solo<float>(h, z0, lambda, tol, max_iter, cg_atol, cg_rtol, cg_itmax, beta_out,
                             info);
  return fail(CSIPM_ERR_UNSUPPORTED, "csipm_solve needs a CSIPM_F64 or CSIPM_MIXED handle");
}

int csipm_admm(csipm_op_t h, const double* z0, double lambda, double rho, int64_t maxt,
               double tol, double* beta_out, double* info) {
  if (!h || !z0 || !beta_out || !info) return fail(CSIPM_ERR_ARG, "null argument");
  if (!(lambda >= 0) || !(rho > 0) || !(tol > 0) || maxt < 1)
    return fail(CSIPM_ERR_ARG, "bad lambda / rho / tol / maxt");
  if (h->prec == CSIPM_F64)
    return admm_solo<double>(h, z0, lambda, rho, maxt, tol, beta_out, info);
  if (h->prec == CSIPM_MIXED)
    return admm_solo<float>(h, z0, lambda, rho, maxt, tol, beta_out, info);
  return fail(CSIPM_ERR_UNSUPPORTED, "csipm_admm needs a CSIPM_F64 or CSIPM_MIXED handle");
}

#define DISPATCH_T(h, F64EXPR, F32EXPR) (OP_DOUBLE(h) ? (F64EXPR) : (F32EXPR))

extern "C++" {
// N-vector scratch of the NLP entry points, made on first use and kept
static int nlp_scratch(csipm_op* h) {
  if (h->scratch) return CSIPM_OK;
  CHK(no_capture(h->stream, "an NLP-model entry point"));
  HIPL(dmalloc(&h->scratch, h->esize() * (size_t)h->N));
  h->scratch_bytes = h->esize() * (size_t)h->N;
  return CSIPM_OK;
}

template <typename T>
static int nlp_obj_t(csipm_op* h, const T* x, const T* mtz0, double lam, double* f) {
  CHK(nlp_scratch(h));
  T* y = (T*)h->scratch;
  int rc;
  if (h->dim == 1) rc = apply_m_perp_beta<T, 1>(h, x, y);
  else if (h->dim == 2) rc = apply_m_perp_beta<T, 2>(h, x, y);
  else rc = apply_m_perp_beta<T, 3>(h, x, y);
  if (rc) return rc;
  const int gv = grid_for(h->N, 4);
  hipLaunchKernelGGL((k_obj_parts<T>), dim3(gv), dim3(kNT), 0, h->stream, y, x, mtz0, h->N,
                     h->part);
  std::vector<double> hp(3 * (size_t)gv);
  HIPCHK(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double a = 0, b = 0, c = 0;
  for (int i = 0; i < gv; ++i) {
    a += hp[i];
    b += hp[gv + i];
    c += hp[2 * gv + i];
  }
  *f = 0.5 * a - b + lam * c;
  return CSIPM_OK;
}

template <typename T>
static int nlp_grad_t(csipm_op* h, const T* x, const T* mtz0, double lam, T* g) {
  CHK((dispatch_mtm<T, SP_BETA>(h, x, epi_store<T>(g), nullptr)));
  hipLaunchKernelGGL((k_grad_tail<T>), dim3(grid_for(h->N, 4)), dim3(kNT), 0, h->stream, g, mtz0,
                     (T)lam, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T>
static int kkt_full_t(csipm_op* h, const T* x, T* y, double alpha, double beta) {
  CHK(nlp_scratch(h));
  T* Mb = (T*)h->scratch;
  CHK((dispatch_mtm<T, SP_BETA>(h, x, epi_store<T>(Mb), nullptr)));
  hipLaunchKernelGGL((k_kkt_full<T>), dim3(grid_for(h->N, 4)), dim3(kNT), 0, h->stream, x, y, Mb,
                     (T)alpha, (T)beta, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}
}  // extern "C++"

int csipm_nlp_obj(csipm_op_t h, const void* x, const void* Mtz0, double lambda, double* f) {
  if (!h || !x || !Mtz0 || !f) return fail(CSIPM_ERR_ARG, "null argument");
  CHK(no_capture(h->stream, "csipm_nlp_obj"));
  return DISPATCH_T(h, nlp_obj_t<double>(h, (const double*)x, (const double*)Mtz0, lambda, f),
                    nlp_obj_t<float>(h, (const float*)x, (const float*)Mtz0, lambda, f));
}

int csipm_nlp_grad(csipm_op_t h, const void* x, const void* Mtz0, double lambda, void* g) {
  if (!h || !x || !Mtz0 || !g) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, nlp_grad_t<double>(h, (const double*)x, (const double*)Mtz0, lambda, (double*)g),
                    nlp_grad_t<float>(h, (const float*)x, (const float*)Mtz0, lambda, (float*)g));
}

int csipm_nlp_cons(csipm_op_t h, const void* x, void* c) {
  if (!h || !x || !c) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_cons<double>), dim3(g), dim3(kNT), 0, h->stream, (const double*)x, (double*)c, h->N);
  else
    hipLaunchKernelGGL((k_cons<float>), dim3(g), dim3(kNT), 0, h->stream, (const float*)x, (float*)c, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_kkt_mul_full(csipm_op_t h, const void* x, void* y, double alpha, double beta) {
  if (!h || !x || !y) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, kkt_full_t<double>(h, (const double*)x, (double*)y, alpha, beta),
                    kkt_full_t<float>(h, (const float*)x, (float*)y, alpha, beta));
}

int csipm_jtprod(csipm_op_t h, const void* y, void* out) {
  if (!h || !y || !out) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_jtprod<double>), dim3(g), dim3(kNT), 0, h->stream, (const double*)y, (double*)out, h->N);
  else
    hipLaunchKernelGGL((k_jtprod<float>), dim3(g), dim3(kNT), 0, h->stream, (const float*)y, (float*)out, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_condense_rhs(csipm_op_t h, const void* w, const void* s1, const void* s2, void* b) {
  if (!h || !w || !s1 || !s2 || !b) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_condense<double>), dim3(g), dim3(kNT), 0, h->stream, (const double*)w,
                       (const double*)s1, (const double*)s2, (double*)b, h->N);
  else
    hipLaunchKernelGGL((k_condense<float>), dim3(g), dim3(kNT), 0, h->stream, (const float*)w,
                       (const float*)s1, (const float*)s2, (float*)b, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_unpack_solution(csipm_op_t h, void* w, const void* x, const void* s1, const void* s2) {
  if (!h || !w || !x || !s1 || !s2) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (OP_DOUBLE(h))
    hipLaunchKernelGGL((k_unpack<double>), dim3(g), dim3(kNT), 0, h->stream, (double*)w,
                       (const double*)x, (const double*)s1, (const double*)s2, h->N);
  else
    hipLaunchKernelGGL((k_unpack<float>), dim3(g), dim3(kNT), 0, h->stream, (float*)w,
                       (const float*)x, (const float*)s1, (const float*)s2, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_kkt_apply_timed(csipm_op_t h, const void* x, void* y, const void* L1, const void* L2,
                          int64_t nrep, int layout, double* pass_ms, double* pass_bytes,
                          int* npass) {
  if (!h || !x || !y || !L1 || !L2 || !pass_ms || !npass) return fail(CSIPM_ERR_ARG, "null argument");
  if (nrep < 1) return fail(CSIPM_ERR_ARG, "nrep < 1");
  if (layout < 0 || layout > 2) return fail(CSIPM_ERR_ARG, "layout must be 0, 1 or 2");
  CHK(no_capture(h->stream, "csipm_kkt_apply_timed"));
  const char* names[8];
  double bytes[8];
  const int np = kkt_pass_info(h, layout, names, bytes);
  hipEvent_t ev[9];
  for (int i = 0; i <= np; ++i) HIPCHK(hipEventCreate(&ev[i]));
  std::vector<double> acc(np, 0.0);
  for (int64_t r = 0; r < nrep; ++r) {
    HIPCHK(hipEventRecord(ev[0], h->stream));
    int rc;
    if (layout == 2)
      rc = OP_DOUBLE(h)
               ? red_apply_t<double, SP_U>(h, (const double*)x, (double*)y, (const double*)L1,
                                           1.0, 0.0, h->part, nullptr, ev + 1)
               : red_apply_t<float, SP_U>(h, (const float*)x, (float*)y, (const float*)L1, 1.0,
                                          0.0, h->part, nullptr, ev + 1);
    else if (layout == 1)
      rc = OP_DOUBLE(h)
               ? kkt_apply_t<double, SP_U>(h, (const double*)x, (double*)y, (const double*)L1,
                                           (const double*)L2, 1.0, 0.0, h->part, nullptr, ev + 1)
               : kkt_apply_t<float, SP_U>(h, (const float*)x, (float*)y, (const float*)L1,
                                          (const float*)L2, 1.0, 0.0, h->part, nullptr, ev + 1);
    else
      rc = OP_DOUBLE(h)
               ? kkt_apply_t<double, SP_BETA>(h, (const double*)x, (double*)y, (const double*)L1,
                                              (const double*)L2, 1.0, 0.0, nullptr, nullptr, ev + 1)
               : kkt_apply_t<float, SP_BETA>(h, (const float*)x, (float*)y, (const float*)L1,
                                             (const float*)L2, 1.0, 0.0, nullptr, nullptr, ev + 1);
    if (rc) return rc;
    HIPCHK(hipEventSynchronize(ev[np]));
    for (int p = 0; p < np; ++p) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev[p], ev[p + 1]));
      acc[p] += ms;
    }
  }
  for (int i = 0; i <= np; ++i) (void)hipEventDestroy(ev[i]);
  for (int p = 0; p < np; ++p) {
    pass_ms[p] = acc[p] / (double)nrep;
    if (pass_bytes) pass_bytes[p] = bytes[p];
  }
  *npass = np;
  return CSIPM_OK;
}

const char* csipm_kkt_pass_name(csipm_op_t h, int layout, int pass) {
  const char* names[8];
  double bytes[8];
  if (!h || layout < 0 || layout > 2) return "";
  const int np = kkt_pass_info(h, layout, names, bytes);
  return (pass >= 0 && pass < np) ? names[pass] : "";
}

// ---------------------------------------------------------------------------
// batch engine
// ---------------------------------------------------------------------------
struct csipm_batch {
  Engine* E = nullptr;
  int prec = CSIPM_F64;
  std::vector<std::string> names;   // of the last csipm_batch_apply_timed
};

#define BATCH_TC(b, EXPR_D, EXPR_F) ((b)->prec == CSIPM_F64 ? (EXPR_D) : (EXPR_F))

int csipm_batch_create(csipm_batch_t* out, int precision, int64_t slots, int max_instances,
                       void* stream) {
  if (!out) return fail(CSIPM_ERR_ARG, "null argument");
  *out = nullptr;
  if (precision != CSIPM_F64 && precision != CSIPM_MIXED)
    return fail(CSIPM_ERR_ARG, "batch precision must be CSIPM_F64 or CSIPM_MIXED");
  if (slots < 1 || max_instances < 1) return fail(CSIPM_ERR_ARG, "empty batch");
  csipm_batch* b = new (std::nothrow) csipm_batch();
  if (!b) return fail(CSIPM_ERR_ALLOC, "out of host memory");
  b->prec = precision;
  const int rc = engine_create(&b->E, slots, max_instances, precision == CSIPM_F64 ? 8 : 4,
                               precision == CSIPM_MIXED, (hipStream_t)stream);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return CSIPM_OK;
}

int csipm_batch_destroy(csipm_batch_t b) {
  if (!b) return CSIPM_OK;
  delete b->E;
  delete b;
  return CSIPM_OK;
}

int csipm_batch_slots(csipm_op_t h, int64_t* slots) {
  if (!h || !slots) return fail(CSIPM_ERR_ARG, "null argument");
  *slots = slots_of(h);
  return CSIPM_OK;
}

int csipm_batch_bytes(int precision, int64_t slots, int64_t* bytes) {
  if (!bytes) return fail(CSIPM_ERR_ARG, "null argument");
  *bytes = (int64_t)engine_bytes_per_slot(precision == CSIPM_F64 ? 8 : 4) * slots;
  return CSIPM_OK;
}

int csipm_batch_admit(csipm_batch_t b, csipm_op_t h, const double* z0, double lambda,
                      double* beta_out, int64_t cg_itmax, int* id) {
  if (!b || !h || !z0 || !beta_out || !id) return fail(CSIPM_ERR_ARG, "null argument");
  if (h->prec != b->prec) return fail(CSIPM_ERR_ARG, "handle precision differs from the batch");
  if (!(lambda >= 0)) return fail(CSIPM_ERR_ARG, "bad lambda");
  CHK(engine_enter(b->E));
  return engine_admit(b->E, h, z0, lambda, beta_out, cg_itmax, id);
}

int csipm_batch_admit_many(csipm_batch_t b, int n, const csipm_op_t* ops,
                           const double* const* z0, const double* lambda,
                           double* const* beta_out, int64_t cg_itmax, int* ids, int* rcs) {
  if (!b || n < 0 || (n > 0 && (!ops || !z0 || !lambda || !beta_out || !ids || !rcs)))
    return fail(CSIPM_ERR_ARG, "null argument");
  CHK(engine_enter(b->E));
  std::unordered_set<csipm_op_t> seen;
  for (int k = 0; k < n; ++k) {
    ids[k] = -1;
    int rc = CSIPM_OK;
    if (!ops[k] || !z0[k] || !beta_out[k]) rc = fail(CSIPM_ERR_ARG, "null argument %d", k);
    else if (ops[k]->prec != b->prec)
      rc = fail(CSIPM_ERR_ARG, "handle %d: precision differs from the batch", k);
    else if (!(lambda[k] >= 0)) rc = fail(CSIPM_ERR_ARG, "bad lambda %d", k);
    else if (!seen.insert(ops[k]).second) rc = fail(CSIPM_ERR_ARG, "handle %d appears twice", k);
    else rc = engine_admit(b->E, ops[k], z0[k], lambda[k], beta_out[k], cg_itmax, &ids[k]);
    rcs[k] = rc;
    if (rc == CSIPM_ERR_HIP) return rc;
  }
  return CSIPM_OK;
}

int csipm_batch_harvest_many(csipm_batch_t b, int n, const int* ids, double* info) {
  if (!b || n < 0 || (n > 0 && (!ids || !info))) return fail(CSIPM_ERR_ARG, "null argument");
  Engine* E = b->E;
  std::vector<char> seen((size_t)E->maxinst, 0);
  for (int k = 0; k < n; ++k) {
    const int id = ids[k];
    if (id < 0 || id >= E->maxinst || !E->inst[(size_t)id].h || E->inst[(size_t)id].fresh)
      return fail(CSIPM_ERR_ARG, "no admitted instance %d", id);
    if (seen[(size_t)id]++) return fail(CSIPM_ERR_ARG, "instance %d appears twice", id);
  }
  CHK(engine_enter(E));
  for (int k = 0; k < n; ++k) CHK(engine_harvest(E, ids[k], info + 8 * (size_t)k, false));
  return engine_leave(E);
}

int csipm_batch_step(csipm_batch_t b, double tol, int64_t max_iter, double cg_atol,
                     double cg_rtol, int* running) {
  if (!b || !running) return fail(CSIPM_ERR_ARG, "null argument");
  if (!(tol > 0) || max_iter < 0) return fail(CSIPM_ERR_ARG, "bad tol / max_iter");
  if (!(cg_atol >= 0) || !(cg_rtol >= 0) || !(cg_atol > 0 || cg_rtol > 0))
    return fail(CSIPM_ERR_ARG, "bad cg_atol / cg_rtol");
  const int mi = (int)std::min<int64_t>(max_iter, 1 << 30);
  CHK(engine_enter(b->E));
  CHK(BATCH_TC(b, engine_step<double>(b->E, tol, mi, cg_atol, cg_rtol, running),
               engine_step<float>(b->E, tol, mi, cg_atol, cg_rtol, running)));
  return engine_leave(b->E);
}

int csipm_batch_poll(csipm_batch_t b, int* ids, int cap, int* n) {
  if (!b || !n || (cap > 0 && !ids)) return fail(CSIPM_ERR_ARG, "null argument");
  Engine* E = b->E;
  HIPCHK(hipMemcpyAsync(E->flags_h, E->flags, sizeof(BFlags) * (size_t)E->maxinst,
                        hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  int k = 0;
  for (int id : E->active)
    if (!E->inst[(size_t)id].fresh && E->flags_h[id].done) {
      if (k < cap) ids[k] = id;
      ++k;
    }
  *n = k;
  return CSIPM_OK;
}

int csipm_batch_harvest(csipm_batch_t b, int id, double* info) {
  if (!b) return fail(CSIPM_ERR_ARG, "null argument");
  Engine* E = b->E;
  if (id < 0 || id >= E->maxinst || !E->inst[(size_t)id].h || E->inst[(size_t)id].fresh)
    return fail(CSIPM_ERR_ARG, "no admitted instance %d", id);
  CHK(engine_enter(E));
  CHK(engine_harvest(E, id, info, false));
  return engine_leave(E);
}

int csipm_batch_solve(csipm_batch_t b, int n, const csipm_op_t* ops, const double* const* z0,
                      const double* lambda, double tol, int64_t max_iter, double cg_atol,
                      double cg_rtol, int64_t cg_itmax, double* const* beta_out, double* info) {
  if (!b || n < 0 || (n > 0 && (!ops || !z0 || !lambda || !beta_out || !info)))
    return fail(CSIPM_ERR_ARG, "null argument");
  std::vector<int> ids((size_t)n, -1);
  for (int i = 0; i < n; ++i) {
    const int rc = csipm_batch_admit(b, ops[i], z0[i], lambda[i], beta_out[i], cg_itmax, &ids[i]);
    if (rc) {
      for (int j = 0; j < i; ++j) engine_drop(b->E, ids[(size_t)j]);
      return rc;
    }
  }
  for (;;) {
    int running = 0;
    const int rc = csipm_batch_step(b, tol, max_iter, cg_atol, cg_rtol, &running);
    if (rc) {
      for (int i = 0; i < n; ++i) engine_drop(b->E, ids[(size_t)i]);
      return rc;
    }
    if (running == 0) break;
  }
  for (int i = 0; i < n; ++i)
    CHK(engine_harvest(b->E, ids[(size_t)i], info + 8 * (size_t)i, false));
  HIPCHK(hipStreamSynchronize(b->E->stream));
  return engine_leave(b->E);
}

int csipm_batch_apply_timed(csipm_batch_t b, int64_t nrep, double* pass_ms,
                            double* pass_bytes, int cap, int* nlaunch) {
  if (!b || !pass_ms || !pass_bytes || !nlaunch || nrep < 1)
    return fail(CSIPM_ERR_ARG, "bad argument");
  b->names.assign((size_t)std::max(cap, 0), std::string());
  CHK(engine_enter(b->E));
  return BATCH_TC(b,
                  engine_apply_timed<double>(b->E, nrep, pass_ms, pass_bytes, b->names.data(),
                                             cap, nlaunch),
                  engine_apply_timed<float>(b->E, nrep, pass_ms, pass_bytes, b->names.data(),
                                            cap, nlaunch));
}

const char* csipm_batch_launch_name(csipm_batch_t b, int k) {
  if (!b || k < 0 || k >= (int)b->names.size()) return "";
  return b->names[(size_t)k].c_str();
}

int csipm_batch_stats(csipm_batch_t b, double* out) {
  if (!b || !out) return fail(CSIPM_ERR_ARG, "null argument");
  for (int i = 0; i < Engine::ST_N; ++i) out[i] = b->E->stats[i];
  return CSIPM_OK;
}


int csipm_gondzio_obj(csipm_op_t h, const void* x, double lambda, double* f) {
  if (!h || !x || !f) return fail(CSIPM_ERR_ARG, "null argument");
  CHK(no_capture(h->stream, "csipm_gondzio_obj"));
  return DISPATCH_T(h, gz_obj_t<double>(h, (const double*)x, lambda, f),
                    gz_obj_t<float>(h, (const float*)x, lambda, f));
}

int csipm_gondzio_grad(csipm_op_t h, const void* x, double lambda, void* g) {
  if (!h || !x || !g) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_grad_t<double>(h, (const double*)x, lambda, (double*)g),
                    gz_grad_t<float>(h, (const float*)x, lambda, (float*)g));
}

int csipm_gondzio_cons(csipm_op_t h, const void* x, const void* z0, void* c) {
  if (!h || !x || !z0 || !c) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_cons_t<double>(h, (const double*)x, (const double*)z0, (double*)c),
                    gz_cons_t<float>(h, (const float*)x, (const float*)z0, (float*)c));
}

int csipm_gondzio_jtprod(csipm_op_t h, const void* v, void* out) {
  if (!h || !v || !out) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_jtprod_t<double>(h, (const double*)v, (double*)out),
                    gz_jtprod_t<float>(h, (const float*)v, (float*)out));
}

int csipm_gondzio_kkt_apply(csipm_op_t h, const void* x, void* y, const void* Wp, const void* Wq,
                            double alpha, double beta) {
  if (!h || !x || !y || !Wp || !Wq) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h,
                    gz_kkt_apply_t<double>(h, (const double*)x, (double*)y, (const double*)Wp,
                                           (const double*)Wq, alpha, beta),
                    gz_kkt_apply_t<float>(h, (const float*)x, (float*)y, (const float*)Wp,
                                          (const float*)Wq, alpha, beta));
}

int csipm_gondzio_kkt_mul_full(csipm_op_t h, const void* x, void* y, double alpha, double beta) {
  if (!h || !x || !y) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_kkt_full_t<double>(h, (const double*)x, (double*)y, alpha, beta),
                    gz_kkt_full_t<float>(h, (const float*)x, (float*)y, alpha, beta));
}

int csipm_gondzio_precond_build(csipm_op_t h, const void* Wp, const void* Wq, void* P11,
                                void* P12, void* P22) {
  if (!h || !Wp || !Wq || !P11 || !P12 || !P22) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h,
                    gz_precond_build_t<double>(h, (const double*)Wp, (const double*)Wq,
                                               (double*)P11, (double*)P12, (double*)P22),
                    gz_precond_build_t<float>(h, (const float*)Wp, (const float*)Wq,
                                              (float*)P11, (float*)P12, (float*)P22));
}

int csipm_gondzio_condense_rhs(csipm_op_t h, const void* w, void* rhs) {
  if (!h || !w || !rhs) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_condense_t<double>(h, (const double*)w, (double*)rhs),
                    gz_condense_t<float>(h, (const float*)w, (float*)rhs));
}

int csipm_gondzio_unpack_solution(csipm_op_t h, void* w, const void* dx) {
  if (!h || !w || !dx) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, gz_unpack_t<double>(h, (double*)w, (const double*)dx),
                    gz_unpack_t<float>(h, (float*)w, (const float*)dx));
}

int csipm_gondzio_cg_solve(csipm_op_t h, const void* b, void* x, const void* Wp, const void* Wq,
                           const void* P11, const void* P12, const void* P22, double atol,
                           double rtol, int64_t itmax, int64_t* iters, double* resid,
                           int* status) {
  if (!h || !b || !x || !Wp || !Wq || !P11 || !P12 || !P22 || !iters || !resid || !status)
    return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(
      h,
      gz_cg_t<double>(h, (const double*)b, (double*)x, (const double*)Wp, (const double*)Wq,
                      (const double*)P11, (const double*)P12, (const double*)P22, atol, rtol,
                      itmax, iters, resid, status),
      gz_cg_t<float>(h, (const float*)b, (float*)x, (const float*)Wp, (const float*)Wq,
                     (const float*)P11, (const float*)P12, (const float*)P22, atol, rtol, itmax,
                     iters, resid, status));
}

int csipm_gondzio_solve(csipm_op_t h, const double* z0, double lambda, double tol,
                        int64_t max_iter, double cg_atol, double cg_rtol, int64_t cg_itmax,
                        double* x_out, double* y_out, double* info) {
  if (!h || !z0 || !x_out || !y_out || !info) return fail(CSIPM_ERR_ARG, "null argument");
  if (!(lambda >= 0) || !(tol > 0) || max_iter < 0)
    return fail(CSIPM_ERR_ARG, "bad lambda / tol / max_iter");
  if (!(cg_atol >= 0) || !(cg_rtol >= 0) || !(cg_atol > 0 || cg_rtol > 0))
    return fail(CSIPM_ERR_ARG, "bad cg_atol / cg_rtol");
  if (h->prec != CSIPM_F64 && h->prec != CSIPM_MIXED)
    return fail(CSIPM_ERR_UNSUPPORTED, "csipm_gondzio_solve needs a CSIPM_F64 or CSIPM_MIXED handle");
  const int64_t n = h->N;
  // the same central path as the FFT model under p = (c + beta)/2,
  // q = (c - beta)/2 (slacks c + beta = 2p, c - beta = 2q)
  double* beta = x_out + 2 * n;   // r's place holds beta until r is formed
  double* c = y_out;
  const int rc = h->prec == CSIPM_F64
                     ? solve_solo<double>(h, z0, lambda, tol, max_iter, cg_atol, cg_rtol,
                                          cg_itmax, beta, info, c)
                     : solve_solo<float>(h, z0, lambda, tol, max_iter, cg_atol, cg_rtol,
                                         cg_itmax, beta, info, c);
  if (rc) return rc;
  const double secs = info[6];
  CHK(each(h, n, [=] __device__(int64_t i) {
    x_out[i] = 0.5 * (c[i] + beta[i]);
    x_out[n + i] = 0.5 * (c[i] - beta[i]);
  }));
  // r = z0 - M (p - q), y = r (stationarity in r of MadNLP's Lagrangian)
  double *d, *md;
  CHK(gz_scratch<double>(h, &d, &md));
  CHK(each(h, n, [=] __device__(int64_t i) { d[i] = x_out[i] - x_out[n + i]; }));
  CHK(op_m<double>(h, d, md));
  CHK(each(h, n, [=] __device__(int64_t i) {
    const double r = z0[i] - md[i];
    x_out[2 * n + i] = r;
    y_out[i] = r;
  }));
  double f = 0;
  CHK(gz_obj_t<double>(h, x_out, lambda, &f));
  info[0] = f;
  info[6] = secs;
  return CSIPM_OK;
}

}  // extern "C"
