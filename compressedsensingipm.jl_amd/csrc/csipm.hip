// libcsipm: C ABI over the gfx950 kernels of the matrix-free IPM hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <initializer_list>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "csipm.h"
#include "csipm_kernels.hpp"

using namespace csipm;

#define CSIPM_VERSION "0.1.0"

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

#define CHK(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != CSIPM_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
namespace {

constexpr int kNT = 256;     // threads per workgroup of line kernels
constexpr int kNTBig = 1024; // dimension-1 kernels with long lines
constexpr int kE = 16;       // complex elements staged per thread and stage

struct LineCfg {
  int nl, ntile1, grid;
  size_t lds;
};
struct Dim1Cfg {
  int nt, np, npairs, grid;
  size_t lds;
};

}  // namespace

struct csipm_op {
  int dim = 0;
  int n[3] = {1, 1, 1};
  int prec = CSIPM_F64;
  hipStream_t stream = nullptr;
  MapGeom g{};
  int H1 = 0;
  int64_t N = 0, Nc = 0;
  FftPlan plan[3]{};
  void* tw[3] = {nullptr, nullptr, nullptr};  // twiddles in op precision
  uint32_t* miss = nullptr;
  int64_t nmiss = 0;
  void* buf = nullptr;     // half spectrum scratch (Nc complex)
  double* part = nullptr;  // per-block partial sums
  int npart = 0;
  CgState* cg = nullptr;   // device
  CgState* cg_host = nullptr;  // pinned mirror
  void* cg_r = nullptr;
  void* cg_z = nullptr;
  void* cg_p = nullptr;
  void* cg_q = nullptr;
  LineCfg lines[3]{};   // per line axis (index = dim 0-based: 1 or 2)
  Dim1Cfg d1{};
  // interior-point workspace (FP64), allocated by the first csipm_solve
  double* ipm = nullptr;
  double* scal = nullptr;       // device scalars (S_* slots)
  double* scal_host = nullptr;  // pinned mirror
  size_t esize() const { return prec == CSIPM_F64 ? 8 : 4; }
};

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
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
  const int odd[] = {17, 13, 11, 7, 5};
  for (int r : odd)
    while (m % r == 0) { if (!push(r)) return false; m /= r; }
  while (m % 9 == 0) { if (!push(9)) return false; m /= 9; }
  while (m % 3 == 0) { if (!push(3)) return false; m /= 3; }
  return m == 1;
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

static void make_geom(csipm_op* h) { h->g = make_map_geom(h->dim, h->n); }

// workgroups of line FFTs: nl lines of length L per group, enough groups to
// fill the chip when the problem allows it
static LineCfg line_cfg(int H1, int L, int64_t others, size_t es) {
  LineCfg c{};
  int nl_max = (kNT * kE) / L;
  if (nl_max < 1) nl_max = 1;
  if (nl_max > H1) nl_max = H1;
  int64_t want = (int64_t)H1 * others / 1024;   // aim for >= ~1024 groups
  int nl = (int)std::max<int64_t>(1, std::min<int64_t>(nl_max, want));
  c.nl = nl;
  c.ntile1 = (H1 + nl - 1) / nl;
  c.grid = (int)(c.ntile1 * others);
  c.lds = (size_t)nl * L * 2 * es;
  return c;
}

static Dim1Cfg dim1_cfg(int N1, int nlines, size_t es) {
  Dim1Cfg c{};
  c.npairs = (nlines + 1) / 2;
  c.nt = (N1 <= kNT * kE) ? kNT : kNTBig;
  int np_max = (c.nt * kE) / N1;
  if (np_max < 1) np_max = 1;
  int want = c.npairs / 1024;
  int np = std::max(1, std::min(np_max, want));
  c.np = np;
  c.grid = (c.npairs + np - 1) / np;
  c.lds = (size_t)np * N1 * 2 * es;
  return c;
}

// ---------------------------------------------------------------------------
// pass launchers
// ---------------------------------------------------------------------------
namespace {

enum class Out { BUF, BETA, REAL };

template <typename T>
Epi<T> epi_store(T* out) {
  Epi<T> e{};
  e.kind = 0;
  e.out = out;
  return e;
}

template <typename T, int D, int SRC, int DST, bool FWD>
int launch_lines(csipm_op* h, int axis, const T* beta_in, const Epi<T>& epi, const int* done) {
  const LineCfg& c = h->lines[axis];
  LineArgs<T> a{};
  a.g = h->g;
  a.plan = h->plan[axis];
  a.tw = (const cpx<T>*)h->tw[axis];
  a.axis = axis;
  a.H1 = h->H1;
  a.nl = c.nl;
  a.ntile1 = c.ntile1;
  a.buf = (cpx<T>*)h->buf;
  a.beta_in = beta_in;
  a.epi = epi;
  a.done = done;
  if (c.lds > 160 * 1024) return fail(CSIPM_ERR_UNSUPPORTED, "line of %d too long", a.plan.L);
  hipLaunchKernelGGL((k_lines<T, kNT, kE, D, SRC, DST, FWD>), dim3(c.grid), dim3(kNT), c.lds,
                     h->stream, a);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T, int D, int SRC, int DST, int MODE>
int launch_dim1(csipm_op* h, const T* beta_in, const T* real_in, T* real_out,
                const Epi<T>& epi, T scale, const int* done) {
  const Dim1Cfg& c = h->d1;
  Dim1Args<T> a{};
  a.g = h->g;
  a.plan = h->plan[0];
  a.tw = (const cpx<T>*)h->tw[0];
  a.N1 = h->n[0];
  a.H1 = h->H1;
  a.nlines = (int)(h->N / h->n[0]);
  a.npairs = c.npairs;
  a.np = c.np;
  a.buf = (cpx<T>*)h->buf;
  a.beta_in = beta_in;
  a.real_in = real_in;
  a.real_out = real_out;
  a.epi = epi;
  a.miss = h->miss;
  a.scale = scale;
  a.done = done;
  if (c.lds > 160 * 1024) return fail(CSIPM_ERR_UNSUPPORTED, "dimension 1 of %d too long", a.N1);
  if (c.nt == kNT)
    hipLaunchKernelGGL((k_dim1<T, kNT, kE, D, SRC, DST, MODE>), dim3(c.grid), dim3(kNT), c.lds,
                       h->stream, a);
  else
    hipLaunchKernelGGL((k_dim1<T, kNTBig, kE, D, SRC, DST, MODE>), dim3(c.grid), dim3(kNTBig),
                       c.lds, h->stream, a);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// y = M^T M x (epi kind 0) or the KKT epilogue; the whole fused operator.
// ev (optional): an event is recorded after every pass (per-pass timing).
template <typename T, int D>
int apply_mtm(csipm_op* h, const T* x, const Epi<T>& epi, const int* done, hipEvent_t* ev) {
  const T s = (T)(1.0 / (double)h->N);
  int k = 0;
  auto mark = [&]() -> int {
    if (ev) HIPCHK(hipEventRecord(ev[k++], h->stream));
    return CSIPM_OK;
  };
  if (D == 1) {
    CHK((launch_dim1<T, 1, SRC_BETA, DST_BETA, MODE_MTM>(h, x, nullptr, nullptr, epi, s, done)));
    return mark();
  }
  if (D == 2) {
    CHK((launch_lines<T, 2, SRC_BETA, DST_BUF, false>(h, 1, x, epi, done)));
    CHK(mark());
    CHK((launch_dim1<T, 2, SRC_BUF, DST_BUF, MODE_MTM>(h, nullptr, nullptr, nullptr, epi, s, done)));
    CHK(mark());
    CHK((launch_lines<T, 2, SRC_BUF, DST_BETA, true>(h, 1, nullptr, epi, done)));
    return mark();
  }
  CHK((launch_lines<T, 3, SRC_BETA, DST_BUF, false>(h, 2, x, epi, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, false>(h, 1, nullptr, epi, done)));
  CHK(mark());
  CHK((launch_dim1<T, 3, SRC_BUF, DST_BUF, MODE_MTM>(h, nullptr, nullptr, nullptr, epi, s, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, true>(h, 1, nullptr, epi, done)));
  CHK(mark());
  CHK((launch_lines<T, 3, SRC_BUF, DST_BETA, true>(h, 2, nullptr, epi, done)));
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

template <typename T, int D>
int apply_m_perpt_z(csipm_op* h, const T* z, T* beta) {
  const T s = (T)(1.0 / std::sqrt((double)h->N));
  Epi<T> e = epi_store<T>(beta);
  if (D == 1)
    return launch_dim1<T, 1, SRC_REAL, DST_BETA, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr);
  if (D == 2) {
    CHK((launch_dim1<T, 2, SRC_REAL, DST_BUF, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr)));
    return launch_lines<T, 2, SRC_BUF, DST_BETA, true>(h, 1, nullptr, e, nullptr);
  }
  CHK((launch_dim1<T, 3, SRC_REAL, DST_BUF, MODE_R2C>(h, nullptr, z, nullptr, e, s, nullptr)));
  CHK((launch_lines<T, 3, SRC_BUF, DST_BUF, true>(h, 1, nullptr, e, nullptr)));
  return launch_lines<T, 3, SRC_BUF, DST_BETA, true>(h, 2, nullptr, e, nullptr);
}

template <typename T>
int dispatch_mtm(csipm_op* h, const T* x, const Epi<T>& epi, const int* done,
                 hipEvent_t* ev = nullptr) {
  if (h->dim == 1) return apply_mtm<T, 1>(h, x, epi, done, ev);
  if (h->dim == 2) return apply_mtm<T, 2>(h, x, epi, done, ev);
  return apply_mtm<T, 3>(h, x, epi, done, ev);
}

int grid_for(int64_t n, int per_thread = 1) {
  int64_t g = (n + (int64_t)kNT * per_thread - 1) / ((int64_t)kNT * per_thread);
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

template <typename T>
int kkt_apply_t(csipm_op* h, const T* x, T* y, const T* L1, const T* L2, double alpha,
                double beta, double* dot_part, const int* done, hipEvent_t* ev = nullptr) {
  Epi<T> e{};
  e.kind = 1;
  e.xb = x;
  e.xz = x + h->N;
  e.L1 = L1;
  e.L2 = L2;
  e.yb = y;
  e.yz = y + h->N;
  e.alpha = (T)alpha;
  e.beta = (T)beta;
  e.dot_part = dot_part;
  return dispatch_mtm<T>(h, x, e, done, ev);
}

// names and algorithmic HBM bytes of the passes of one K apply
int kkt_pass_info(csipm_op* h, const char** names, double* bytes) {
  const double s = (double)h->esize(), N = (double)h->N, Nc = (double)h->Nc;
  const double mask = h->nmiss > 0 ? std::ceil(N / 8.0) : 0.0;
  if (h->dim == 1) {
    names[0] = "dim1_mtm_kkt";
    bytes[0] = 6 * N * s + mask;   // x_beta, x_z, L1, L2 in; y_beta, y_z out
    return 1;
  }
  const double spec = 2 * Nc * s;   // one pass over the half spectrum
  if (h->dim == 2) {
    names[0] = "lines2_inv_from_beta"; bytes[0] = N * s + spec;
    names[1] = "dim1_c2r_mask_r2c";    bytes[1] = 2 * spec + mask;
    names[2] = "lines2_fwd_kkt";       bytes[2] = spec + 6 * N * s;
    return 3;
  }
  names[0] = "lines3_inv_from_beta"; bytes[0] = N * s + spec;
  names[1] = "lines2_inv";           bytes[1] = 2 * spec;
  names[2] = "dim1_c2r_mask_r2c";    bytes[2] = 2 * spec + mask;
  names[3] = "lines2_fwd";           bytes[3] = 2 * spec;
  names[4] = "lines3_fwd_kkt";       bytes[4] = spec + 6 * N * s;
  return 5;
}

// number of workgroups of the pass that ends in beta space
int final_grid(csipm_op* h) {
  if (h->dim == 1) return h->d1.grid;
  return h->lines[h->dim - 1].grid;
}

template <typename T>
int cg_solve_t(csipm_op* h, const T* b, T* x, const T* L1, const T* L2, const T* P11,
               const T* P12, const T* P22, double atol, double rtol, int64_t itmax,
               int64_t* iters, double* resid, int* status) {
  const int64_t n = h->N;
  T* r = (T*)h->cg_r;
  T* z = (T*)h->cg_z;
  T* p = (T*)h->cg_p;
  T* q = (T*)h->cg_q;
  CgState init{};
  init.atol = atol;
  init.rtol = rtol;
  init.itmax = itmax > 0 ? itmax : 4 * n;
  HIPCHK(hipMemcpyAsync(h->cg, &init, sizeof(CgState), hipMemcpyHostToDevice, h->stream));
  const int gv = grid_for(n, 4);
  hipLaunchKernelGGL((k_cg_init<T>), dim3(gv), dim3(kNT), 0, h->stream, b, x, r, z, p, P11, P12,
                     P22, n, h->part);
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gv);
  HIPCHK(hipGetLastError());
  const int* done = &h->cg->done;
  const int gk = final_grid(h);
  const int check_every = 4;
  for (int64_t it = 0;; ++it) {
    if (it % check_every == 0) {
      HIPCHK(hipMemcpyAsync(h->cg_host, h->cg, sizeof(CgState), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      if (h->cg_host->done) break;
    }
    CHK(kkt_apply_t<T>(h, p, q, L1, L2, 1.0, 0.0, h->part, done));
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gk);
    hipLaunchKernelGGL((k_cg_update<T>), dim3(gv), dim3(kNT), 0, h->stream, x, r, z, p, q, P11,
                       P12, P22, n, h->cg, h->part);
    hipLaunchKernelGGL(k_cg_gamma, dim3(1), dim3(kNT), 0, h->stream, h->cg, h->part, gv);
    hipLaunchKernelGGL((k_cg_pupdate<T>), dim3(gv), dim3(kNT), 0, h->stream, p, z, 2 * n, h->cg);
    HIPCHK(hipGetLastError());
  }
  *iters = h->cg_host->iters;
  *resid = h->cg_host->rnorm;
  *status = h->cg_host->status;
  return CSIPM_OK;
}


// ---------------------------------------------------------------------------
// interior-point driver (FP64): the Mehrotra predictor-corrector on the
// condensed FFT-KKT system, two preconditioned CG solves per iteration
// ---------------------------------------------------------------------------
enum { W_B = 0, W_BETA, W_C, W_Z1, W_Z2, W_QB, W_S1, W_S2, W_L1, W_L2, W_P11, W_P12, W_P22,
       W_DT1, W_DT2, W_DZ1, W_DZ2, W_NVEC, W_RHS = W_NVEC, W_D = W_NVEC + 2, W_TOTAL = W_NVEC + 4 };

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

int ipm_solve_f64(csipm_op* h, const double* z0, double lam, double tol, int64_t max_iter,
                  double cg_rtol, int64_t cg_itmax, double* beta_out, double* info) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = h->N;
  if (!h->ipm) {
    HIPCHK(hipMalloc(&h->ipm, sizeof(double) * (size_t)n * W_TOTAL));
    HIPCHK(hipMalloc(&h->scal, sizeof(double) * S_NSLOTS));
    HIPCHK(hipHostMalloc(&h->scal_host, sizeof(double) * S_NSLOTS, hipHostMallocDefault));
  }
  double* W[W_TOTAL];
  for (int i = 0; i < W_TOTAL; ++i) W[i] = h->ipm + (size_t)i * n;
  double *b = W[W_B], *beta = W[W_BETA], *c = W[W_C], *z1 = W[W_Z1], *z2 = W[W_Z2], *Qb = W[W_QB];
  double *s1 = W[W_S1], *s2 = W[W_S2], *L1 = W[W_L1], *L2 = W[W_L2];
  double *P11 = W[W_P11], *P12 = W[W_P12], *P22 = W[W_P22];
  double *dt1 = W[W_DT1], *dt2 = W[W_DT2], *dz1 = W[W_DZ1], *dz2 = W[W_DZ2];
  double *rhs = W[W_RHS], *d = W[W_D];
  const int gv = grid_for(n, 4);
  const double tau = 0.995;
  // b = M^T z0
  if (h->dim == 1) CHK((apply_m_perpt_z<double, 1>(h, z0, b)));
  else if (h->dim == 2) CHK((apply_m_perpt_z<double, 2>(h, z0, b)));
  else CHK((apply_m_perpt_z<double, 3>(h, z0, b)));
  hipLaunchKernelGGL(k_absmax, dim3(gv), dim3(kNT), 0, h->stream, b, n, h->part);
  CHK(ipm_finalize(h, gv, {1}, 0, S_BNORM, tau));
  hipLaunchKernelGGL(k_ipm_init, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, n, h->scal,
                     lam);
  HIPCHK(hipGetLastError());
  int64_t cg_total = 0;
  int status = 1;
  double f = 0, dual = 0, mu = 0;
  int64_t it = 0;
  for (it = 0; it <= max_iter; ++it) {
    CHK(dispatch_mtm<double>(h, beta, epi_store<double>(Qb), nullptr));
    hipLaunchKernelGGL(k_ipm_resid, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, Qb, b,
                       lam, n, h->part);
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
    if (dual <= tol * (1.0 + sc[S_BNORM]) && mu <= tol * (1.0 + std::fabs(f) / (double)n)) {
      status = 0;
      break;
    }
    if (it == max_iter) break;
    hipLaunchKernelGGL(k_ipm_prep, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, Qb, b, lam,
                       n, s1, s2, L1, L2, P11, P12, P22, rhs);
    HIPCHK(hipGetLastError());
    int64_t its;
    double res;
    int st;
    CHK(cg_solve_t<double>(h, rhs, d, L1, L2, P11, P12, P22, 0.0, cg_rtol, cg_itmax, &its, &res,
                           &st));
    cg_total += its;
    hipLaunchKernelGGL(k_ipm_dir_aff, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, s1, s2,
                       d, n, dt1, dt2, dz1, dz2, h->part);
    CHK(ipm_finalize(h, gv, {2, 2}, 1, 0, tau));
    hipLaunchKernelGGL(k_ipm_muaff, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, dt1, dt2,
                       dz1, dz2, h->scal, n, h->part);
    CHK(ipm_finalize(h, gv, {0}, 2, 0, tau));
    hipLaunchKernelGGL(k_ipm_rhs_corr, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, Qb, b,
                       dt1, dt2, dz1, dz2, h->scal, lam, n, rhs);
    HIPCHK(hipGetLastError());
    CHK(cg_solve_t<double>(h, rhs, d, L1, L2, P11, P12, P22, 0.0, cg_rtol, cg_itmax, &its, &res,
                           &st));
    cg_total += its;
    hipLaunchKernelGGL(k_ipm_dir, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, s1, s2, d,
                       h->scal, n, dt1, dt2, dz1, dz2, h->part);
    CHK(ipm_finalize(h, gv, {2, 2}, 3, 0, tau));
    hipLaunchKernelGGL(k_ipm_update, dim3(gv), dim3(kNT), 0, h->stream, beta, c, z1, z2, d, dz1,
                       dz2, h->scal, n);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(beta_out, beta, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice,
                        h->stream));
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
  for (int i = 0; i < 3; ++i)
    if (h->tw[i]) (void)hipFree(h->tw[i]);
  if (h->miss) (void)hipFree(h->miss);
  if (h->buf) (void)hipFree(h->buf);
  if (h->part) (void)hipFree(h->part);
  if (h->cg) (void)hipFree(h->cg);
  if (h->cg_host) (void)hipHostFree(h->cg_host);
  if (h->cg_r) (void)hipFree(h->cg_r);
  if (h->cg_z) (void)hipFree(h->cg_z);
  if (h->cg_p) (void)hipFree(h->cg_p);
  if (h->cg_q) (void)hipFree(h->cg_q);
  if (h->ipm) (void)hipFree(h->ipm);
  if (h->scal) (void)hipFree(h->scal);
  if (h->scal_host) (void)hipHostFree(h->scal_host);
  delete h;
  return CSIPM_OK;
}

static int create_impl(csipm_op* h, int dim, const int64_t sizes[3], int rdft, int precision,
                       const int64_t* missing, int64_t n_missing, void* stream) {
  if (dim < 1 || dim > 3) return fail(CSIPM_ERR_ARG, "dim must be 1, 2 or 3 (got %d)", dim);
  if (!sizes) return fail(CSIPM_ERR_ARG, "sizes is null");
  if (rdft != 1) return fail(CSIPM_ERR_UNSUPPORTED, "only the rdft=true path is implemented");
  if (precision != CSIPM_F64 && precision != CSIPM_F32)
    return fail(CSIPM_ERR_ARG, "bad precision %d", precision);
  h->dim = dim;
  h->prec = precision;
  h->stream = (hipStream_t)stream;
  for (int i = 0; i < dim; ++i) {
    if (sizes[i] < 2 || sizes[i] > (1 << 20))
      return fail(CSIPM_ERR_SHAPE, "size %d = %lld out of range [2, 2^20]", i + 1,
                  (long long)sizes[i]);
    h->n[i] = (int)sizes[i];
  }
  make_geom(h);
  h->N = h->g.N;
  if (h->N > ((int64_t)1 << 31)) return fail(CSIPM_ERR_SHAPE, "N too large");
  h->H1 = h->n[0] / 2 + 1;
  h->Nc = (int64_t)h->H1 * h->n[1] * h->n[2];
  const size_t es = h->esize();
  for (int i = 0; i < dim; ++i) {
    if (!factorize(h->n[i], h->plan[i]))
      return fail(CSIPM_ERR_UNSUPPORTED,
                  "size %d has a prime factor above 17 (Bluestein not yet implemented)", h->n[i]);
    const int rc = precision == CSIPM_F64 ? upload_twiddles<double>(h->n[i], &h->tw[i])
                                          : upload_twiddles<float>(h->n[i], &h->tw[i]);
    if (rc) return rc;
  }
  // launch shapes
  h->d1 = dim1_cfg(h->n[0], (int)(h->N / h->n[0]), es);
  if (h->d1.nt * kE < h->n[0] || h->d1.lds > 160 * 1024)
    return fail(CSIPM_ERR_UNSUPPORTED, "dimension-1 length %d exceeds one workgroup", h->n[0]);
  if (dim >= 2) h->lines[1] = line_cfg(h->H1, h->n[1], h->n[2], es);
  if (dim == 3) h->lines[2] = line_cfg(h->H1, h->n[2], h->n[1], es);
  for (int ax = 1; ax < dim; ++ax)
    if (h->n[ax] > kNT * kE || h->lines[ax].lds > 160 * 1024)
      return fail(CSIPM_ERR_UNSUPPORTED, "dimension-%d length %d exceeds one workgroup", ax + 1,
                  h->n[ax]);
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
  HIPCHK(hipMalloc(&h->buf, (size_t)h->Nc * 2 * es));
  int np = 4096;
  np = std::max(np, h->d1.grid);
  for (int ax = 1; ax < dim; ++ax) np = std::max(np, h->lines[ax].grid);
  h->npart = 8 * np;   // up to 8 reduced quantities per pass
  HIPCHK(hipMalloc(&h->part, sizeof(double) * (size_t)h->npart));
  HIPCHK(hipMalloc(&h->cg, sizeof(CgState)));
  HIPCHK(hipHostMalloc(&h->cg_host, sizeof(CgState), hipHostMallocDefault));
  const size_t vb = (size_t)(2 * h->N) * es;
  HIPCHK(hipMalloc(&h->cg_r, vb));
  HIPCHK(hipMalloc(&h->cg_z, vb));
  HIPCHK(hipMalloc(&h->cg_p, vb));
  HIPCHK(hipMalloc(&h->cg_q, vb));
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
  if (h->prec == CSIPM_F64) {
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
  if (h->prec == CSIPM_F64) {
    if (h->dim == 1) return apply_m_perpt_z<double, 1>(h, (const double*)z, (double*)beta);
    if (h->dim == 2) return apply_m_perpt_z<double, 2>(h, (const double*)z, (double*)beta);
    return apply_m_perpt_z<double, 3>(h, (const double*)z, (double*)beta);
  }
  if (h->dim == 1) return apply_m_perpt_z<float, 1>(h, (const float*)z, (float*)beta);
  if (h->dim == 2) return apply_m_perpt_z<float, 2>(h, (const float*)z, (float*)beta);
  return apply_m_perpt_z<float, 3>(h, (const float*)z, (float*)beta);
}

int csipm_mtm(csipm_op_t h, const void* x, void* y) {
  if (!h || !x || !y) return fail(CSIPM_ERR_ARG, "null argument");
  if (h->prec == CSIPM_F64) return dispatch_mtm<double>(h, (const double*)x, epi_store<double>((double*)y), nullptr);
  return dispatch_mtm<float>(h, (const float*)x, epi_store<float>((float*)y), nullptr);
}

int csipm_kkt_apply(csipm_op_t h, const void* x, void* y, const void* L1, const void* L2,
                    double alpha, double beta) {
  if (!h || !x || !y || !L1 || !L2) return fail(CSIPM_ERR_ARG, "null argument");
  if (h->prec == CSIPM_F64)
    return kkt_apply_t<double>(h, (const double*)x, (double*)y, (const double*)L1,
                               (const double*)L2, alpha, beta, nullptr, nullptr);
  return kkt_apply_t<float>(h, (const float*)x, (float*)y, (const float*)L1, (const float*)L2,
                            alpha, beta, nullptr, nullptr);
}

int csipm_precond_build(csipm_op_t h, const void* s1, const void* s2, void* L1, void* L2,
                        void* P11, void* P12, void* P22) {
  if (!h || !s1 || !s2 || !L1 || !L2 || !P11 || !P12 || !P22)
    return fail(CSIPM_ERR_ARG, "null argument");
  const int grid = (int)((h->N + kNT - 1) / kNT);
  if (h->prec == CSIPM_F64)
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
  if (h->prec == CSIPM_F64)
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
  if (h->prec == CSIPM_F64)
    return cg_solve_t<double>(h, (const double*)b, (double*)x, (const double*)L1,
                              (const double*)L2, (const double*)P11, (const double*)P12,
                              (const double*)P22, atol, rtol, itmax, iters, resid, status);
  return cg_solve_t<float>(h, (const float*)b, (float*)x, (const float*)L1, (const float*)L2,
                           (const float*)P11, (const float*)P12, (const float*)P22, atol, rtol,
                           itmax, iters, resid, status);
}

int csipm_solve(csipm_op_t h, const double* z0, double lambda, double tol, int64_t max_iter,
                double cg_rtol, int64_t cg_itmax, double* beta_out, double* info) {
  if (!h || !z0 || !beta_out || !info) return fail(CSIPM_ERR_ARG, "null argument");
  if (h->prec != CSIPM_F64)
    return fail(CSIPM_ERR_UNSUPPORTED, "csipm_solve needs an FP64 handle");
  if (!(lambda >= 0) || !(tol > 0) || max_iter < 0)
    return fail(CSIPM_ERR_ARG, "bad lambda / tol / max_iter");
  return ipm_solve_f64(h, z0, lambda, tol, max_iter, cg_rtol, cg_itmax, beta_out, info);
}


#define DISPATCH_T(h, F64EXPR, F32EXPR) ((h)->prec == CSIPM_F64 ? (F64EXPR) : (F32EXPR))

extern "C++" {
template <typename T>
static int nlp_obj_t(csipm_op* h, const T* x, const T* mtz0, double lam, double* f) {
  T* y = (T*)h->cg_q;   // N-vector scratch
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
}  // extern "C++"

int csipm_nlp_obj(csipm_op_t h, const void* x, const void* Mtz0, double lambda, double* f) {
  if (!h || !x || !Mtz0 || !f) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, nlp_obj_t<double>(h, (const double*)x, (const double*)Mtz0, lambda, f),
                    nlp_obj_t<float>(h, (const float*)x, (const float*)Mtz0, lambda, f));
}

extern "C++" {
template <typename T>
static int nlp_grad_t(csipm_op* h, const T* x, const T* mtz0, double lam, T* g) {
  CHK(dispatch_mtm<T>(h, x, epi_store<T>(g), nullptr));
  hipLaunchKernelGGL((k_grad_tail<T>), dim3(grid_for(h->N, 4)), dim3(kNT), 0, h->stream, g, mtz0,
                     (T)lam, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}
}  // extern "C++"

int csipm_nlp_grad(csipm_op_t h, const void* x, const void* Mtz0, double lambda, void* g) {
  if (!h || !x || !Mtz0 || !g) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, nlp_grad_t<double>(h, (const double*)x, (const double*)Mtz0, lambda, (double*)g),
                    nlp_grad_t<float>(h, (const float*)x, (const float*)Mtz0, lambda, (float*)g));
}

int csipm_nlp_cons(csipm_op_t h, const void* x, void* c) {
  if (!h || !x || !c) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (h->prec == CSIPM_F64)
    hipLaunchKernelGGL((k_cons<double>), dim3(g), dim3(kNT), 0, h->stream, (const double*)x, (double*)c, h->N);
  else
    hipLaunchKernelGGL((k_cons<float>), dim3(g), dim3(kNT), 0, h->stream, (const float*)x, (float*)c, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

extern "C++" {
template <typename T>
static int kkt_full_t(csipm_op* h, const T* x, T* y, double alpha, double beta) {
  T* Mb = (T*)h->cg_q;
  CHK(dispatch_mtm<T>(h, x, epi_store<T>(Mb), nullptr));
  hipLaunchKernelGGL((k_kkt_full<T>), dim3(grid_for(h->N, 4)), dim3(kNT), 0, h->stream, x, y, Mb,
                     (T)alpha, (T)beta, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}
}  // extern "C++"

int csipm_kkt_mul_full(csipm_op_t h, const void* x, void* y, double alpha, double beta) {
  if (!h || !x || !y) return fail(CSIPM_ERR_ARG, "null argument");
  return DISPATCH_T(h, kkt_full_t<double>(h, (const double*)x, (double*)y, alpha, beta),
                    kkt_full_t<float>(h, (const float*)x, (float*)y, alpha, beta));
}

int csipm_jtprod(csipm_op_t h, const void* y, void* out) {
  if (!h || !y || !out) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (h->prec == CSIPM_F64)
    hipLaunchKernelGGL((k_jtprod<double>), dim3(g), dim3(kNT), 0, h->stream, (const double*)y, (double*)out, h->N);
  else
    hipLaunchKernelGGL((k_jtprod<float>), dim3(g), dim3(kNT), 0, h->stream, (const float*)y, (float*)out, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

int csipm_condense_rhs(csipm_op_t h, const void* w, const void* s1, const void* s2, void* b) {
  if (!h || !w || !s1 || !s2 || !b) return fail(CSIPM_ERR_ARG, "null argument");
  const int g = grid_for(h->N, 4);
  if (h->prec == CSIPM_F64)
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
  if (h->prec == CSIPM_F64)
    hipLaunchKernelGGL((k_unpack<double>), dim3(g), dim3(kNT), 0, h->stream, (double*)w,
                       (const double*)x, (const double*)s1, (const double*)s2, h->N);
  else
    hipLaunchKernelGGL((k_unpack<float>), dim3(g), dim3(kNT), 0, h->stream, (float*)w,
                       (const float*)x, (const float*)s1, (const float*)s2, h->N);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}


int csipm_kkt_apply_timed(csipm_op_t h, const void* x, void* y, const void* L1, const void* L2,
                          int64_t nrep, double* pass_ms, double* pass_bytes, int* npass) {
  if (!h || !x || !y || !L1 || !L2 || !pass_ms || !npass) return fail(CSIPM_ERR_ARG, "null argument");
  if (nrep < 1) return fail(CSIPM_ERR_ARG, "nrep < 1");
  const char* names[8];
  double bytes[8];
  const int np = kkt_pass_info(h, names, bytes);
  hipEvent_t ev[9];
  for (int i = 0; i <= np; ++i) HIPCHK(hipEventCreate(&ev[i]));
  std::vector<double> acc(np, 0.0);
  for (int64_t r = 0; r < nrep; ++r) {
    HIPCHK(hipEventRecord(ev[0], h->stream));
    int rc = h->prec == CSIPM_F64
                 ? kkt_apply_t<double>(h, (const double*)x, (double*)y, (const double*)L1,
                                       (const double*)L2, 1.0, 0.0, nullptr, nullptr, ev + 1)
                 : kkt_apply_t<float>(h, (const float*)x, (float*)y, (const float*)L1,
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

const char* csipm_kkt_pass_name(csipm_op_t h, int pass) {
  const char* names[8];
  double bytes[8];
  if (!h) return "";
  const int np = kkt_pass_info(h, names, bytes);
  return (pass >= 0 && pass < np) ? names[pass] : "";
}

}  // extern "C"
