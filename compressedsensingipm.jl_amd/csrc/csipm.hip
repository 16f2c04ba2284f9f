// libcsipm: C ABI over the gfx950 kernels of the matrix-free IPM hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "csipm.h"
#include "csipm_kernels.hpp"

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

#define CHK(expr)                    \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != CSIPM_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
namespace {

constexpr int kNT = 256;      // threads per workgroup of the FFT passes
constexpr int kNTBig = 1024;  // dimension-1 passes with long lines (1D)
constexpr int kE = 16;        // complex elements staged per thread and stage
constexpr size_t kLdsSmall = 64 * 1024;
constexpr size_t kLdsMax = 160 * 1024;

struct LineCfg {
  int lg = 0, ntile1 = 0, grid = 0;
  size_t lds = 0;
};
struct Dim1Cfg {
  int nt = 0, lg = 0, ps = 0, npairs = 0, grid = 0;
  size_t lds = 0;
};

// per precision: twiddles and launch shapes (the plans are shared)
struct PrecState {
  void* tw[3] = {nullptr, nullptr, nullptr};
  LineCfg lines[3];
  Dim1Cfg d1;
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
  FftPlan plan[3]{};
  PrecState p64, p32;
  uint32_t* miss = nullptr;
  int64_t nmiss = 0;
  uint32_t* valid = nullptr;     // bit per u slot
  void* buf = nullptr;           // half-spectrum scratch (Nc complex doubles)
  double* part = nullptr;        // per-block partial sums
  int npart = 0;
  CgState* cg = nullptr;         // device
  CgState* cg_host = nullptr;    // pinned mirror
  void* cgw = nullptr;           // CG vectors r, z, p, q (2*max(N, U) each, op precision)
  void* conv = nullptr;          // reference-layout CG: u copies of b, x, L1, L2, P11, P12, P22
  double* ipm = nullptr;         // interior-point workspace
  void* ipm_lo = nullptr;        // its CG-precision part
  double* scal = nullptr;
  double* scal_host = nullptr;
  bool has64() const { return prec == CSIPM_F64 || prec == CSIPM_MIXED; }
  bool has32() const { return prec == CSIPM_F32 || prec == CSIPM_MIXED; }
  size_t esize() const { return prec == CSIPM_F64 ? 8 : 4; }   // op / CG precision
  int64_t vlen() const { return std::max<int64_t>(N, U); }
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

// radix sequence for length L: 16, 8, 4, 2 for the power of two, then odd
// radices 17, 13, 11, 7, 15 (=3*5), 9, 5, 3
static bool factorize(int L, FftPlan& p) {
  p.L = L;
  p.nstages = 0;
  int m = L;
  auto push = [&](int r) {
    if (p.nstages >= kMaxStages) return false;
    p.radix[p.nstages++] = r;
    return true;
  };
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

template <typename T>
static int upload_twiddles(int L, void** out) {
  std::vector<cpx<T>> h(L);
  const long double two_pi = 2.0L * acosl(-1.0L);
  for (int t = 0; t < L; ++t) {
    const long double a = two_pi * (long double)t / (long double)L;
    h[t].x = (T)cosl(a);
    h[t].y = (T)(-sinl(a));
  }
  HIPCHK(hipMalloc(out, sizeof(cpx<T>) * (size_t)L));
  HIPCHK(hipMemcpy(*out, h.data(), sizeof(cpx<T>) * (size_t)L, hipMemcpyHostToDevice));
  return CSIPM_OK;
}

// line FFTs along dim 2/3: 2^lg lines of length L per workgroup, as many as
// the registers (lines*L <= NT*E) and 64 KB of LDS allow, fewer when that
// would leave the chip with under ~1024 workgroups
static LineCfg line_cfg(int H1, int L, int64_t others, size_t cs) {
  LineCfg c{};
  int lg = 0;
  while ((2 << lg) * L <= kNT * kE && (size_t)(2 << lg) * L * cs <= kLdsSmall && (1 << lg) < H1)
    ++lg;
  while (lg > 0 && (int64_t)((H1 + (1 << lg) - 1) >> lg) * others < 1024) --lg;
  c.lg = lg;
  c.ntile1 = (H1 + (1 << lg) - 1) >> lg;
  c.grid = (int)(c.ntile1 * others);
  c.lds = (size_t)(1 << lg) * L * cs;
  return c;
}

static Dim1Cfg dim1_cfg(int N1, int nlines, size_t cs) {
  Dim1Cfg c{};
  c.npairs = (nlines + 1) / 2;
  c.nt = (N1 <= kNT * kE) ? kNT : kNTBig;
  int lg = 0;
  auto lds_of = [&](int l) { return (size_t)N1 * ((1 << l) + (l > 0 ? 1 : 0)) * cs; };
  while ((2 << lg) * N1 <= c.nt * kE && lds_of(lg + 1) <= kLdsSmall && (1 << lg) < c.npairs) ++lg;
  while (lg > 0 && ((c.npairs + (1 << lg) - 1) >> lg) < 1024) --lg;
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

template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes > kLdsSmall)
    HIPCHK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
  return CSIPM_OK;
}

template <typename T>
PassArgs<T> base_args(csipm_op* h, int axis) {
  PassArgs<T> a{};
  a.g = h->g;
  a.plan = h->plan[axis];
  a.tw = (const cpx<T>*)h->ps<T>().tw[axis];
  a.H1 = h->H1;
  a.buf = (cpx<T>*)h->buf;
  a.h1m = magic(h->H1);
  a.n1m = magic(h->n[0]);
  a.miss = h->miss;
  return a;
}

template <typename T, int D, int SRC, int DST, bool FWD>
int launch_lines(csipm_op* h, int axis, const void* src, const Epi<T>& epi, const int* done) {
  const LineCfg& c = h->ps<T>().lines[axis];
  PassArgs<T> a = base_args<T>(h, axis);
  a.axis = axis;
  a.lg = c.lg;
  a.ps = 1 << c.lg;
  a.ntile1 = c.ntile1;
  a.beta_in = (const T*)src;
  a.u_in = (const cpx<T>*)src;
  a.epi = epi;
  a.done = done;
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

// number of workgroups of the pass that ends in beta / u space
template <typename T>
int final_grid(csipm_op* h) {
  if (h->dim == 1) return h->ps<T>().d1.grid;
  return h->ps<T>().lines[h->dim - 1].grid;
}

// ---------------------------------------------------------------------------
// preconditioned CG (Krylov.jl cg recurrence), device-resident scalars.
// nb = 2: the condensed system K = [[M^T M + L1, L2], [L2, L1]] on vectors of
// 2n entries with the block preconditioner [[P11, P12], [P12, P22]];
// nb = 1: the reduced system M^T M + D (D passed as L1) on n entries with the
// diagonal preconditioner P11.  reliable > 0: every `reliable` iterations the
// residual is recomputed as b - K x and the recurrence restarts from it.
// ---------------------------------------------------------------------------
template <typename T, int SP>
int cg_apply(csipm_op* h, int nb, const T* x, T* y, const T* L1, const T* L2, double* dot_part,
             const int* done) {
  if (nb == 2) return kkt_apply_t<T, SP>(h, x, y, L1, L2, 1.0, 0.0, dot_part, done);
  return red_apply_t<T, SP>(h, x, y, L1, 1.0, 0.0, dot_part, done);
}

template <typename T, int SP>
int cg_run(csipm_op* h, int nb, const T* b, T* x, const T* L1, const T* L2, const T* P11,
           const T* P12, const T* P22, double atol, double rtol, int64_t itmax, int reliable,
           int64_t* iters, double* resid, int* status) {
  const int64_t n = SP == SP_U ? h->U : h->N;
  const int64_t vl = 2 * h->vlen();
  T* r = (T*)h->cgw;
  T* z = r + vl;
  T* p = z + vl;
  T* q = p + vl;
  CgState init{};
  init.atol = atol;
  init.rtol = rtol;
  init.itmax = itmax > 0 ? itmax : 4 * h->N;
  HIPCHK(hipMemcpyAsync(h->cg, &init, sizeof(CgState), hipMemcpyHostToDevice, h->stream));
  const int gv = grid_for(n, 4);
  hipLaunchKernelGGL((k_cg_init<T>), dim3(gv), dim3(kNT), 0, h->stream, b, x, r, z, p, P11, P12,
                     P22, n, nb, h->part);
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gv);
  HIPCHK(hipGetLastError());
  const int* done = &h->cg->done;
  const int gk = final_grid<T>(h);
  const int check_every = 4;
  for (int64_t it = 0;; ++it) {
    if (it % check_every == 0) {
      HIPCHK(hipMemcpyAsync(h->cg_host, h->cg, sizeof(CgState), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      if (h->cg_host->done) break;
    }
    CHK((cg_apply<T, SP>(h, nb, p, q, L1, L2, h->part, done)));
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gk);
    hipLaunchKernelGGL((k_cg_update<T>), dim3(gv), dim3(kNT), 0, h->stream, x, r, z, p, q, P11,
                       P12, P22, n, nb, h->cg, h->part);
    hipLaunchKernelGGL(k_cg_gamma, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gv);
    if (reliable > 0 && (it + 1) % reliable == 0) {
      CHK((cg_apply<T, SP>(h, nb, x, q, L1, L2, nullptr, done)));
      hipLaunchKernelGGL((k_cg_residual<T>), dim3(gv), dim3(kNT), 0, h->stream, b, q, r, z, P11,
                         P12, P22, n, nb, h->cg, h->part);
      hipLaunchKernelGGL(k_cg_regamma, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gv);
    }
    hipLaunchKernelGGL((k_cg_pupdate<T>), dim3(gv), dim3(kNT), 0, h->stream, p, z, nb * n, h->cg);
    HIPCHK(hipGetLastError());
  }
  *iters = h->cg_host->iters;
  *resid = h->cg_host->rnorm;
  *status = h->cg_host->status;
  return CSIPM_OK;
}

// beta-layout <-> u conversions of one N-vector
// sign = 1 for a vector, 0 for a diagonal matrix (whose entries keep their sign)
template <typename T>
int to_u(csipm_op* h, const T* beta, T* u, int sign = 1) {
  const int grid = (int)((h->Nc + kNT - 1) / kNT);
  cpx<T>* uu = (cpx<T>*)u;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_beta_to_u<1, T>), dim3(grid), dim3(kNT), 0, h->stream, beta, uu, h->g, h->H1, h->Nc, sign);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_beta_to_u<2, T>), dim3(grid), dim3(kNT), 0, h->stream, beta, uu, h->g, h->H1, h->Nc, sign);
  else
    hipLaunchKernelGGL((k_beta_to_u<3, T>), dim3(grid), dim3(kNT), 0, h->stream, beta, uu, h->g, h->H1, h->Nc, sign);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T>
int from_u(csipm_op* h, const T* u, T* beta) {
  const int grid = (int)((h->Nc + kNT - 1) / kNT);
  const cpx<T>* uu = (const cpx<T>*)u;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_u_to_beta<1, T>), dim3(grid), dim3(kNT), 0, h->stream, uu, beta, h->g, h->H1, h->Nc);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_u_to_beta<2, T>), dim3(grid), dim3(kNT), 0, h->stream, uu, beta, h->g, h->H1, h->Nc);
  else
    hipLaunchKernelGGL((k_u_to_beta<3, T>), dim3(grid), dim3(kNT), 0, h->stream, uu, beta, h->g, h->H1, h->Nc);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// the reference-layout CG (csipm_cg_solve): converted to the spectrum order,
// solved there, converted back
template <typename T>
int cg_solve_ref(csipm_op* h, const T* b, T* x, const T* L1, const T* L2, const T* P11,
                 const T* P12, const T* P22, double atol, double rtol, int64_t itmax,
                 int64_t* iters, double* resid, int* status) {
  const int64_t N = h->N, U = h->U;
  if (!h->conv) HIPCHK(hipMalloc(&h->conv, sizeof(T) * (size_t)U * 9));
  T* ub = (T*)h->conv;
  T* ux = ub + 2 * U;
  T* uL1 = ux + 2 * U;
  T* uL2 = uL1 + U;
  T* uP11 = uL2 + U;
  T* uP12 = uP11 + U;
  T* uP22 = uP12 + U;
  CHK(to_u<T>(h, b, ub));
  CHK(to_u<T>(h, b + N, ub + U));
  CHK(to_u<T>(h, L1, uL1, 0));
  CHK(to_u<T>(h, L2, uL2, 0));
  CHK(to_u<T>(h, P11, uP11, 0));
  CHK(to_u<T>(h, P12, uP12, 0));
  CHK(to_u<T>(h, P22, uP22, 0));
  CHK((cg_run<T, SP_U>(h, 2, ub, ux, uL1, uL2, uP11, uP12, uP22, atol, rtol, itmax, 0, iters,
                       resid, status)));
  CHK(from_u<T>(h, ux, x));
  CHK(from_u<T>(h, ux + U, x + N));
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// interior-point driver: Mehrotra predictor-corrector on the condensed FFT-KKT
// system, all vectors in the spectrum order (u), control in FP64.  TC is the
// precision of the K apply inside CG (double, or float with one FP64
// refinement step of each Newton direction).
// ---------------------------------------------------------------------------
enum { W_B = 0, W_BETA, W_C, W_Z1, W_Z2, W_QB, W_S1, W_S2, W_DT1, W_DT2, W_DZ1, W_DZ2,
       W_L1D, W_L2D, W_DGD, W_NVEC, W_RHSD = W_NVEC, W_DD = W_NVEC + 2, W_TMP = W_NVEC + 4,
       W_TOTAL = W_NVEC + 6 };
enum { C_L1 = 0, C_L2, C_P11, C_P12, C_P22, C_NVEC, C_RHS = C_NVEC, C_D = C_NVEC + 2,
       C_TOTAL = C_NVEC + 4 };

int ipm_finalize(csipm_op* h, int nparts, std::initializer_list<int> ops, int mode, int first,
                 double tau) {
  RedSpec spec{};
  spec.nq = 0;
  for (int op : ops) spec.op[spec.nq++] = op;
  hipLaunchKernelGGL(k_ipm_finalize, dim3(1), dim3(kNT), 0, h->stream, h->part, nparts, spec,
                     h->scal, mode, first, 0.5 / (double)h->N, tau);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename TC>
int ipm_solve_t(csipm_op* h, const double* z0, double lam, double tol, int64_t max_iter,
                double cg_rtol, int64_t cg_itmax, double* beta_out, double* info) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t U = h->U;
  const bool mixed = sizeof(TC) == 4;
  if (!h->ipm) {
    HIPCHK(hipMalloc(&h->ipm, sizeof(double) * (size_t)U * W_TOTAL));
    HIPCHK(hipMalloc(&h->ipm_lo, sizeof(TC) * (size_t)U * C_TOTAL));
    HIPCHK(hipMalloc(&h->scal, sizeof(double) * S_NSLOTS));
    HIPCHK(hipHostMalloc(&h->scal_host, sizeof(double) * S_NSLOTS, hipHostMallocDefault));
  }
  double* W[W_TOTAL];
  for (int i = 0; i < W_TOTAL; ++i) W[i] = h->ipm + (size_t)i * U;
  TC* C[C_TOTAL];
  for (int i = 0; i < C_TOTAL; ++i) C[i] = (TC*)h->ipm_lo + (size_t)i * U;
  double *b = W[W_B], *beta = W[W_BETA], *c = W[W_C], *z1 = W[W_Z1], *z2 = W[W_Z2], *Qb = W[W_QB];
  double *s1 = W[W_S1], *s2 = W[W_S2], *L1d = W[W_L1D], *L2d = W[W_L2D];
  double *dt1 = W[W_DT1], *dt2 = W[W_DT2], *dz1 = W[W_DZ1], *dz2 = W[W_DZ2];
  double *rhsd = W[W_RHSD], *dd = W[W_DD], *tmp = W[W_TMP], *Dgd = W[W_DGD];
  TC *L1 = C[C_L1], *L2 = C[C_L2], *P11 = C[C_P11], *P12 = C[C_P12], *P22 = C[C_P22];
  TC *rhs = C[C_RHS], *d = C[C_D];
  const uint32_t* valid = h->valid;
  const int gv = grid_for(U, 4);
  const double tau = 0.995;
  const double rtol = mixed ? std::max(cg_rtol, 1e-5) : cg_rtol;
  const int reliable = mixed ? 25 : 0;
  // b = M^T z0, in the spectrum order
  CHK((dispatch_m_perpt_z<double, SP_U>(h, z0, b)));
  hipLaunchKernelGGL(k_absmax, dim3(gv), dim3(kNT), 0, h->stream, b, U, h->part);
  CHK(ipm_finalize(h, gv, {1}, 0, S_BNORM, tau));
  hipLaunchKernelGGL(k_ipm_init, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, U, h->scal,
                     lam, valid);
  HIPCHK(hipGetLastError());
  int64_t cg_total = 0;
  int status = 1;
  double f = 0, dual = 0, mu = 0;
  int64_t it = 0;
  // One Newton solve.  FP64: CG on the condensed system (the reference's
  // K), direction in d.  Mixed: the c-block is eliminated in FP64
  // (dc = (h_c - L2 db) / L1 leaves (M^T M + D) db = h_b - (L2/L1) h_c with
  // D = L1 - L2^2/L1 = 4 S1 S2 / (S1 + S2) > 0), the reduced system is solved
  // by FP32 CG with the diagonal preconditioner 1/(1 + D), then one FP64
  // refinement step; direction in dd.  In FP32 the condensed K itself is
  // unusable near the solution: L1 and |L2| agree to more than 24 bits there.
  auto newton = [&]() -> int {
    int64_t its;
    double res;
    int st;
    if (!mixed) {
      CHK((cg_run<TC, SP_U>(h, 2, rhs, d, L1, L2, P11, P12, P22, 0.0, rtol, cg_itmax, reliable,
                            &its, &res, &st)));
      cg_total += its;
      return CSIPM_OK;
    }
    CHK((cg_run<TC, SP_U>(h, 1, rhs, d, L1, nullptr, P11, nullptr, nullptr, 0.0, rtol, cg_itmax,
                          reliable, &its, &res, &st)));
    cg_total += its;
    hipLaunchKernelGGL((k_widen<TC>), dim3(gv), dim3(kNT), 0, h->stream, dd, d, U);
    HIPCHK(hipMemcpyAsync(tmp, rhsd, sizeof(double) * U, hipMemcpyDeviceToDevice, h->stream));
    CHK((red_apply_t<double, SP_U>(h, dd, tmp, Dgd, -1.0, 1.0, nullptr, nullptr)));
    hipLaunchKernelGGL((k_narrow<TC>), dim3(gv), dim3(kNT), 0, h->stream, tmp, rhs, U);
    CHK((cg_run<TC, SP_U>(h, 1, rhs, d, L1, nullptr, P11, nullptr, nullptr, 0.0, rtol, cg_itmax,
                          reliable, &its, &res, &st)));
    cg_total += its;
    hipLaunchKernelGGL((k_widen_add<TC>), dim3(gv), dim3(kNT), 0, h->stream, dd, d, U);
    hipLaunchKernelGGL(k_red_back, dim3(gv), dim3(kNT), 0, h->stream, dd, rhsd, L1d, L2d, U,
                       valid);
    HIPCHK(hipGetLastError());
    return CSIPM_OK;
  };
  for (it = 0; it <= max_iter; ++it) {
    CHK((dispatch_mtm<double, SP_U>(h, beta, epi_store<double>(Qb), nullptr)));
    hipLaunchKernelGGL(k_ipm_resid, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, Qb, b, lam,
                       U, h->part, valid);
    CHK(ipm_finalize(h, gv, {1, 1, 0, 0, 0, 0}, 0, S_RB, tau));
    HIPCHK(hipMemcpyAsync(h->scal_host, h->scal, sizeof(double) * S_NSLOTS,
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double* sc = h->scal_host;
    dual = std::max(sc[S_RB], sc[S_RC]);
    mu = sc[S_MU];
    f = 0.5 * sc[S_BQB] - sc[S_BB] + lam * sc[S_SUMC];
    if (!std::isfinite(f) || !std::isfinite(mu) || !std::isfinite(dual)) {
      status = 2;
      break;
    }
    if (dual <= tol * (1.0 + sc[S_BNORM]) && mu <= tol * (1.0 + std::fabs(f) / (double)h->N)) {
      status = 0;
      break;
    }
    if (it == max_iter) break;
    if (mixed)
      hipLaunchKernelGGL((k_ipm_prep_red<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2,
                         Qb, b, lam, U, s1, s2, L1, P11, rhs, L1d, L2d, Dgd, rhsd, valid);
    else
      hipLaunchKernelGGL((k_ipm_prep<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, Qb,
                         b, lam, U, s1, s2, L1, L2, P11, P12, P22, rhs, valid);
    HIPCHK(hipGetLastError());
    CHK(newton());
    if (mixed)
      hipLaunchKernelGGL((k_ipm_dir_aff<double>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1,
                         z2, s1, s2, dd, U, dt1, dt2, dz1, dz2, h->part, valid);
    else
      hipLaunchKernelGGL((k_ipm_dir_aff<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2,
                         s1, s2, d, U, dt1, dt2, dz1, dz2, h->part, valid);
    CHK(ipm_finalize(h, gv, {2, 2}, 1, 0, tau));
    hipLaunchKernelGGL(k_ipm_muaff, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, dt1, dt2,
                       dz1, dz2, h->scal, U, h->part, valid);
    CHK(ipm_finalize(h, gv, {0}, 2, 0, tau));
    hipLaunchKernelGGL((k_ipm_rhs_corr<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2,
                       Qb, b, dt1, dt2, dz1, dz2, h->scal, lam, U, rhs, mixed ? rhsd : nullptr,
                       L1d, L2d, valid);
    HIPCHK(hipGetLastError());
    CHK(newton());
    if (mixed) {
      hipLaunchKernelGGL((k_ipm_dir<double>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2,
                         s1, s2, dd, h->scal, U, dt1, dt2, dz1, dz2, h->part, valid);
      CHK(ipm_finalize(h, gv, {2, 2}, 3, 0, tau));
      hipLaunchKernelGGL((k_ipm_update<double>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1,
                         z2, dd, dz1, dz2, h->scal, U, valid);
    } else {
      hipLaunchKernelGGL((k_ipm_dir<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, s1,
                         s2, d, h->scal, U, dt1, dt2, dz1, dz2, h->part, valid);
      CHK(ipm_finalize(h, gv, {2, 2}, 3, 0, tau));
      hipLaunchKernelGGL((k_ipm_update<TC>), dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2,
                         d, dz1, dz2, h->scal, U, valid);
    }
    HIPCHK(hipGetLastError());
  }
  CHK(from_u<double>(h, beta, beta_out));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  info[0] = f;
  info[1] = (double)it;
  info[2] = (double)cg_total;
  info[3] = (double)status;
  info[4] = dual;
  info[5] = mu;
  info[6] = secs;
  info[7] = 0.0;
  return CSIPM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* csipm_version(void) { return CSIPM_VERSION; }

const char* csipm_last_error(void) { return g_err.c_str(); }

int csipm_op_destroy(csipm_op_t h) {
  if (!h) return CSIPM_OK;
  for (int i = 0; i < 3; ++i) {
    if (h->p64.tw[i]) (void)hipFree(h->p64.tw[i]);
    if (h->p32.tw[i]) (void)hipFree(h->p32.tw[i]);
  }
  void* bufs[] = {h->miss, h->valid, h->buf, h->part, h->cg, h->cgw, h->conv,
                  h->ipm, h->ipm_lo, h->scal};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (h->cg_host) (void)hipHostFree(h->cg_host);
  if (h->scal_host) (void)hipHostFree(h->scal_host);
  delete h;
  return CSIPM_OK;
}

static int set_shapes(csipm_op* h, PrecState& ps, size_t cs) {
  const int dim = h->dim;
  ps.d1 = dim1_cfg(h->n[0], (int)(h->N / h->n[0]), cs);
  if (ps.d1.nt * kE < h->n[0] || ps.d1.lds > kLdsMax)
    return fail(CSIPM_ERR_UNSUPPORTED, "dimension-1 length %d exceeds one workgroup", h->n[0]);
  if (dim >= 2) ps.lines[1] = line_cfg(h->H1, h->n[1], h->n[2], cs);
  if (dim == 3) ps.lines[2] = line_cfg(h->H1, h->n[2], h->n[1], cs);
  for (int ax = 1; ax < dim; ++ax)
    if (h->n[ax] > kNT * kE || ps.lines[ax].lds > kLdsMax)
      return fail(CSIPM_ERR_UNSUPPORTED, "dimension-%d length %d exceeds one workgroup", ax + 1,
                  h->n[ax]);
  return CSIPM_OK;
}

static int create_impl(csipm_op* h, int dim, const int64_t sizes[3], int rdft, int precision,
                       const int64_t* missing, int64_t n_missing, void* stream) {
  if (dim < 1 || dim > 3) return fail(CSIPM_ERR_ARG, "dim must be 1, 2 or 3 (got %d)", dim);
  if (!sizes) return fail(CSIPM_ERR_ARG, "sizes is null");
  if (rdft != 1) return fail(CSIPM_ERR_UNSUPPORTED, "only the rdft=true path is implemented");
  if (precision != CSIPM_F64 && precision != CSIPM_F32 && precision != CSIPM_MIXED)
    return fail(CSIPM_ERR_ARG, "bad precision %d", precision);
  h->dim = dim;
  h->prec = precision;
  h->stream = (hipStream_t)stream;
  for (int i = 0; i < dim; ++i) {
    if (sizes[i] < 2 || sizes[i] > 16384)
      return fail(CSIPM_ERR_SHAPE, "size %d = %lld out of range [2, 16384]", i + 1,
                  (long long)sizes[i]);
    h->n[i] = (int)sizes[i];
  }
  h->g = make_map_geom(h->dim, h->n);
  h->N = h->g.N;
  if (h->N > ((int64_t)1 << 31)) return fail(CSIPM_ERR_SHAPE, "N too large");
  h->H1 = h->n[0] / 2 + 1;
  h->Nc = (int64_t)h->H1 * h->n[1] * h->n[2];
  h->U = 2 * h->Nc;
  for (int i = 0; i < dim; ++i) {
    if (!factorize(h->n[i], h->plan[i]))
      return fail(CSIPM_ERR_UNSUPPORTED,
                  "size %d has a prime factor above 17 (Bluestein not yet implemented)", h->n[i]);
    if (h->has64()) CHK(upload_twiddles<double>(h->n[i], &h->p64.tw[i]));
    if (h->has32()) CHK(upload_twiddles<float>(h->n[i], &h->p32.tw[i]));
  }
  if (h->has64()) CHK(set_shapes(h, h->p64, sizeof(cpx<double>)));
  if (h->has32()) CHK(set_shapes(h, h->p32, sizeof(cpx<float>)));
  // missing-sample bitmask
  if (n_missing < 0) return fail(CSIPM_ERR_ARG, "n_missing < 0");
  if (n_missing > 0 && !missing) return fail(CSIPM_ERR_ARG, "missing is null");
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
    HIPCHK(hipMalloc(&h->miss, sizeof(uint32_t) * (size_t)words));
    HIPCHK(hipMemcpy(h->miss, bits.data(), sizeof(uint32_t) * (size_t)words,
                     hipMemcpyHostToDevice));
  }
  // u-slot validity
  const int64_t vwords = (h->U + 31) / 32;
  HIPCHK(hipMalloc(&h->valid, sizeof(uint32_t) * (size_t)vwords));
  {
    const int grid = (int)((vwords + kNT - 1) / kNT);
    if (dim == 1)
      hipLaunchKernelGGL((k_u_valid<1>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    else if (dim == 2)
      hipLaunchKernelGGL((k_u_valid<2>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    else
      hipLaunchKernelGGL((k_u_valid<3>), dim3(grid), dim3(kNT), 0, h->stream, h->valid, h->g, h->H1, h->Nc);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMalloc(&h->buf, (size_t)h->Nc * sizeof(cpx<double>)));
  int np = 4096;
  for (PrecState* ps : {&h->p64, &h->p32}) {
    np = std::max(np, ps->d1.grid);
    for (int ax = 1; ax < dim; ++ax) np = std::max(np, ps->lines[ax].grid);
  }
  h->npart = 8 * np;   // up to 8 reduced quantities per pass
  HIPCHK(hipMalloc(&h->part, sizeof(double) * (size_t)h->npart));
  HIPCHK(hipMalloc(&h->cg, sizeof(CgState)));
  HIPCHK(hipHostMalloc(&h->cg_host, sizeof(CgState), hipHostMallocDefault));
  HIPCHK(hipMalloc(&h->cgw, h->esize() * (size_t)(8 * h->vlen())));
  HIPCHK(hipStreamSynchronize(h->stream));
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
    csipm_op_destroy(h);
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
    return cg_solve_ref<double>(h, (const double*)b, (double*)x, (const double*)L1,
                                (const double*)L2, (const double*)P11, (const double*)P12,
                                (const double*)P22, atol, rtol, itmax, iters, resid, status);
  return cg_solve_ref<float>(h, (const float*)b, (float*)x, (const float*)L1, (const float*)L2,
                             (const float*)P11, (const float*)P12, (const float*)P22, atol, rtol,
                             itmax, iters, resid, status);
}

int csipm_solve(csipm_op_t h, const double* z0, double lambda, double tol, int64_t max_iter,
                double cg_rtol, int64_t cg_itmax, double* beta_out, double* info) {
  if (!h || !z0 || !beta_out || !info) return fail(CSIPM_ERR_ARG, "null argument");
  if (!(lambda >= 0) || !(tol > 0) || max_iter < 0)
    return fail(CSIPM_ERR_ARG, "bad lambda / tol / max_iter");
  if (h->prec == CSIPM_F64)
    return ipm_solve_t<double>(h, z0, lambda, tol, max_iter, cg_rtol, cg_itmax, beta_out, info);
  if (h->prec == CSIPM_MIXED)
    return ipm_solve_t<float>(h, z0, lambda, tol, max_iter, cg_rtol, cg_itmax, beta_out, info);
  return fail(CSIPM_ERR_UNSUPPORTED, "csipm_solve needs a CSIPM_F64 or CSIPM_MIXED handle");
}

#define DISPATCH_T(h, F64EXPR, F32EXPR) (OP_DOUBLE(h) ? (F64EXPR) : (F32EXPR))

extern "C++" {
template <typename T>
static int nlp_obj_t(csipm_op* h, const T* x, const T* mtz0, double lam, double* f) {
  T* y = (T*)h->cgw;   // N-vector scratch
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
  T* Mb = (T*)h->cgw;
  CHK((dispatch_mtm<T, SP_BETA>(h, x, epi_store<T>(Mb), nullptr)));
  hipLaunchKernelGGL((k_kkt_full<T>), dim3(grid_for(h->N, 4)), dim3(kNT), 0, h->stream, x, y, Mb,
                     (T)alpha, (T)beta, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}
}  // extern "C++"

int csipm_nlp_obj(csipm_op_t h, const void* x, const void* Mtz0, double lambda, double* f) {
  if (!h || !x || !Mtz0 || !f) return fail(CSIPM_ERR_ARG, "null argument");
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

}  // extern "C"
