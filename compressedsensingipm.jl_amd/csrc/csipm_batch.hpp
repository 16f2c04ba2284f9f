// Interior-point and CG vector kernels over a batch of instances (included
// once, by csipm.hip, after csipm_kernels.hpp).
//
// Layout.  Every instance owns a slot range [off, off + n) of the
// concatenated spectrum-ordered vectors (n = U = 2 Nc of that instance, off a
// multiple of 256, so the validity bitmask of the batch is the concatenation
// of the instances' masks).  Two-block vectors ([x_beta; x_c], the CG vectors
// and KKT right-hand sides) hold the instance at [2 off, 2 off + 2n).
// A launch covers all instances: block b works for instance seg.inst[b] as
// its (b - seg.first[inst])-th block; partial sums of instance i go to
// part[info[i].pbase + q * nblocks_i + block].  An instance whose interior-
// point run has ended (flags[i].done) is skipped by every kernel.
#pragma once

#include "csipm_kernels.hpp"

namespace csipm {

struct BInst {
  int64_t off;    // first slot (1-block vectors)
  int64_t n;      // slots, U
  int64_t N;      // beta entries
  double lam;
  int64_t pbase;  // first partial-sum slot
  int gcg;        // partial sums written by the last pass of the CG apply
  int itmax;      // CG iteration cap
};

struct BFlags {
  int done;       // interior-point run ended
  int status;     // 0 solved, 1 iteration cap, 2 non-finite
  int iters;      // interior-point steps taken
  int cgits;      // CG iterations over all solves
  double f, dual, mu;
};

// zeroed interior-point flags of newly admitted instances (ids[0, n))
__global__ __launch_bounds__(256) void bk_reset_flags(const int* __restrict__ ids, int n,
                                                      BFlags* flags) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) flags[ids[i]] = BFlags{};
}

#define BSEG_BEGIN                                                  \
  const int inst = uniform(seg.inst[blockIdx.x]);                   \
  const int f0_ = uniform(seg.first[inst]);                         \
  const int lb = (int)blockIdx.x - f0_;                             \
  const int nbk = uniform(seg.first[inst + 1]) - f0_;               \
  const BInst& I = info[inst];                                      \
  if (flags && flags[inst].done) return;
#define BFOR(i)                                                                     \
  for (int64_t i = I.off + (int64_t)lb * blockDim.x + threadIdx.x; i < I.off + I.n; \
       i += (int64_t)nbk * blockDim.x)
#define BVALID(v, i) (!(v) || slot_valid((v), (i)))

template <int NQ>
__device__ __forceinline__ void bstore_partials(double* part, const double (&acc)[NQ],
                                                const int (&ops)[NQ], int lb, int nbk) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double r;
    if (ops[q] == 0) r = block_reduce(acc[q], OpSum());
    else if (ops[q] == 1) r = block_reduce(acc[q], OpMax());
    else r = block_reduce(acc[q], OpMin());
    if (threadIdx.x == 0) part[q * nbk + lb] = r;
  }
}

// ===========================================================================
// preconditioned CG (Krylov.jl cg recurrence), one state per instance.
// nb = 2: condensed system, block preconditioner [[P11, P12], [P12, P22]];
// nb = 1: reduced system, diagonal preconditioner P11.
// ===========================================================================
struct CgState {
  double gamma;
  double pq;
  double alpha;
  double beta;
  double rnorm;
  double eps;
  int64_t iters;
  int64_t itmax;
  int done;
  int status;     // 0 converged, 1 iteration cap, 2 zero curvature
};

__device__ __forceinline__ double sum_partials(const double* part, int nparts) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  return block_reduce(s, OpSum());
}

// x = 0, r = b, z = P r, p = z; partial r'z
template <typename T>
__global__ __launch_bounds__(256) void bk_cg_init(BSeg seg, const BInst* __restrict__ info,
                                                  const BFlags* flags, const T* __restrict__ b,
                                                  T* x, T* r, T* z, T* p,
                                                  const T* __restrict__ P11,
                                                  const T* __restrict__ P12,
                                                  const T* __restrict__ P22, int nb,
                                                  double* part) {
  BSEG_BEGIN
  double acc = 0.0;
  BFOR(i) {
    const int64_t j = i + I.off;
    const T rb = b[j];
    if (nb == 1) {
      const T zb = P11[i] * rb;
      x[j] = 0;
      r[j] = rb;
      z[j] = zb;
      p[j] = zb;
      acc += (double)rb * zb;
      continue;
    }
    const T rz = b[j + I.n];
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    x[j] = 0;
    x[j + I.n] = 0;
    r[j] = rb;
    r[j + I.n] = rz;
    z[j] = zb;
    z[j + I.n] = zz;
    p[j] = zb;
    p[j + I.n] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part + I.pbase, acc, lb);
}

// warm start of the reduced system (nb = 1): x and its residual r = b - K x
// are given; z = P r, p = z; partials r'z and, nbk slots further, b'Pb (the
// stopping test stays relative to the right-hand side)
template <typename T>
__global__ __launch_bounds__(256) void bk_cg_init_warm(BSeg seg, const BInst* __restrict__ info,
                                                       const BFlags* flags,
                                                       const T* __restrict__ b,
                                                       const T* __restrict__ r, T* z, T* p,
                                                       const T* __restrict__ P11, double* part) {
  BSEG_BEGIN
  double acc[2] = {0.0, 0.0};
  BFOR(i) {
    const int64_t j = i + I.off;
    const T rb = r[j], bb = b[j], pd = P11[i];
    const T zb = pd * rb;
    z[j] = zb;
    p[j] = zb;
    acc[0] += (double)rb * zb;
    acc[1] += (double)bb * (pd * bb);
  }
  const int ops[2] = {0, 0};
  bstore_partials<2>(part + I.pbase, acc, ops, lb, nbk);
}

// one block per instance: gamma = r'z, eps = atol + rtol sqrt(gamma) (warm:
// rtol sqrt(b'Pb), from the second partial set)
__global__ void bk_cg_start(BSeg vseg, const int* ids, const BInst* __restrict__ info,
                            const BFlags* flags, CgState* cg, const double* part, double atol,
                            double rtol, int warm) {
  const int inst = ids[blockIdx.x];
  const BInst& I = info[inst];
  const int np = vseg.first[inst + 1] - vseg.first[inst];
  const double g = sum_partials(part + I.pbase, np);
  const double gb = warm ? sum_partials(part + I.pbase + np, np) : g;
  if (threadIdx.x == 0) {
    CgState& st = cg[inst];
    st.gamma = g;
    st.rnorm = sqrt(g > 0 ? g : 0.0);
    st.eps = atol + rtol * sqrt(gb > 0 ? gb : 0.0);
    st.iters = 0;
    st.itmax = I.itmax;
    st.status = 1;
    st.done = 0;
    st.beta = 0.0;
    if (flags && flags[inst].done) {
      st.done = 1;
    } else if (st.rnorm <= st.eps) {
      st.done = 1;
      st.status = 0;
    }
  }
}

// after q = K p (its last pass left gcg partials of p'q): alpha = gamma / p'q
// (every workgroup sums the apply's p'q partials itself,
// in the same order, so all agree; the first one records it and a zero
// curvature), then x += alpha p; r -= alpha q; z = P r; partial r'z at
// pbase + gcg, beside the p'q partials other workgroups are still reading
template <typename T>
__global__ __launch_bounds__(256) void bk_cg_update(BSeg seg, const BInst* __restrict__ info,
                                                    CgState* cg, T* x, T* r, T* z,
                                                    const T* __restrict__ p,
                                                    const T* __restrict__ q,
                                                    const T* __restrict__ P11,
                                                    const T* __restrict__ P12,
                                                    const T* __restrict__ P22, int nb,
                                                    double* part) {
  const BFlags* flags = nullptr;
  BSEG_BEGIN
  CgState& st = cg[inst];
  if (st.done) return;
  __shared__ double s_pq;
  const double pq0 = sum_partials(part + I.pbase, I.gcg);
  if (threadIdx.x == 0) s_pq = pq0;
  __syncthreads();
  const double pq = s_pq;
  if (!(pq > 0.0)) {
    if (lb == 0 && threadIdx.x == 0) {
      st.pq = pq;
      st.done = 1;
      st.status = 2;
    }
    return;
  }
  const double ad = st.gamma / pq;
  if (lb == 0 && threadIdx.x == 0) {
    st.pq = pq;
    st.alpha = ad;
  }
  const T alpha = (T)ad;
  double acc = 0.0;
  BFOR(i) {
    const int64_t j = i + I.off;
    if (nb == 1) {
      x[j] += alpha * p[j];
      const T rb = r[j] - alpha * q[j];
      const T zb = P11[i] * rb;
      r[j] = rb;
      z[j] = zb;
      acc += (double)rb * zb;
      continue;
    }
    const int64_t k = j + I.n;
    x[j] += alpha * p[j];
    x[k] += alpha * p[k];
    const T rb = r[j] - alpha * q[j];
    const T rz = r[k] - alpha * q[k];
    r[j] = rb;
    r[k] = rz;
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    z[j] = zb;
    z[k] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part + I.pbase + I.gcg, acc, lb);
}

// gamma <- r'z, beta = gamma_new / gamma_old, stopping tests
__global__ void bk_cg_gamma(BSeg vseg, const int* ids, const BInst* __restrict__ info,
                            CgState* cg, const double* part) {
  const int inst = ids[blockIdx.x];
  CgState& st = cg[inst];
  if (st.done) return;
  const BInst& I = info[inst];
  const double g =
      sum_partials(part + I.pbase + I.gcg, vseg.first[inst + 1] - vseg.first[inst]);
  if (threadIdx.x == 0) {
    st.iters += 1;
    st.rnorm = sqrt(g > 0 ? g : 0.0);
    st.beta = g / st.gamma;
    st.gamma = g;
    if (st.rnorm <= st.eps) {
      st.done = 1;
      st.status = 0;
    } else if (st.iters >= st.itmax) {
      st.done = 1;
      st.status = 1;
    }
  }
}

// reliable update: r = b - q with q = K x just recomputed, z = P r, partial r'z
template <typename T>
__global__ __launch_bounds__(256) void bk_cg_residual(BSeg seg, const BInst* __restrict__ info,
                                                      const CgState* cg,
                                                      const T* __restrict__ b,
                                                      const T* __restrict__ q, T* r, T* z,
                                                      const T* __restrict__ P11,
                                                      const T* __restrict__ P12,
                                                      const T* __restrict__ P22, int nb,
                                                      double* part) {
  const BFlags* flags = nullptr;
  BSEG_BEGIN
  if (cg[inst].done) return;
  double acc = 0.0;
  BFOR(i) {
    const int64_t j = i + I.off;
    if (nb == 1) {
      const T rb = b[j] - q[j];
      const T zb = P11[i] * rb;
      r[j] = rb;
      z[j] = zb;
      acc += (double)rb * zb;
      continue;
    }
    const int64_t k = j + I.n;
    const T rb = b[j] - q[j], rz = b[k] - q[k];
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    r[j] = rb;
    r[k] = rz;
    z[j] = zb;
    z[k] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part + I.pbase, acc, lb);
}

// after a reliable update: gamma from the true residual, restart p = z
__global__ void bk_cg_regamma(BSeg vseg, const int* ids, const BInst* __restrict__ info,
                              CgState* cg, const double* part) {
  const int inst = ids[blockIdx.x];
  CgState& st = cg[inst];
  if (st.done) return;
  const BInst& I = info[inst];
  const double g = sum_partials(part + I.pbase, vseg.first[inst + 1] - vseg.first[inst]);
  if (threadIdx.x == 0) {
    st.rnorm = sqrt(g > 0 ? g : 0.0);
    st.gamma = g;
    st.beta = 0.0;
    if (st.rnorm <= st.eps) {
      st.done = 1;
      st.status = 0;
    }
  }
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(256) void bk_cg_pupdate(BSeg seg, const BInst* __restrict__ info,
                                                     const CgState* cg, T* p,
                                                     const T* __restrict__ z, int nb) {
  const BFlags* flags = nullptr;
  BSEG_BEGIN
  if (cg[inst].done) return;
  const T beta = (T)cg[inst].beta;
  BFOR(i) {
    const int64_t j = i + I.off;
    p[j] = z[j] + beta * p[j];
    if (nb == 2) p[j + I.n] = z[j + I.n] + beta * p[j + I.n];
  }
}

// CG iterations of the solve just finished, into the instance's total
__global__ void bk_cg_account(int n, const int* ids, const CgState* cg, BFlags* flags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = ids[k];
  if (!flags[i].done) flags[i].cgits += (int)cg[i].iters;
}

// ===========================================================================
// Interior-point driver (FP64 control).  State: beta, c, duals z1, z2 of the
// slacks t1 = c + beta >= 0, t2 = c - beta >= 0 of FFTNLPModel's constraints
// -beta - c <= 0, beta - c <= 0 (src/fft_model.jl:92-101).  Slots outside
// the instance's beta (padding of the spectrum order) stay inert: beta = 0,
// c = 1, z = 0, zero right-hand sides and directions.
// ===========================================================================
struct RedSpec {
  int nq;
  int op[8];  // 0 sum, 1 max, 2 min
};

__device__ __forceinline__ double red_ident(int op) {
  return op == 0 ? 0.0 : (op == 1 ? -1e300 : 1e300);
}

// ipm scalar slots (per instance)
enum {
  S_RB = 0, S_RC, S_TZ, S_BQB, S_BB, S_SUMC,  // residual pass
  S_MU, S_AP_AFF, S_AD_AFF, S_SIGMA, S_AP, S_AD, S_BNORM, S_NSLOTS = 16
};

// One block per instance (ids[block]).  mode 0: reductions into scal[first + q]; with
// first == S_RB also mu, the objective, the dual residual and the stopping
// test (tol, max_iter); 1: affine step lengths; 2: sigma = min(1,
// (mu_aff/mu)^3); 3: corrector step lengths with tau; 4: affine step lengths
// (reductions 0, 1) and sigma from the bilinear form of mu_aff in them
// (sums 2-5: t'z, t'dz, dt'z, dt'dz).
__global__ void bk_ipm_finalize(BSeg vseg, const int* ids, const BInst* __restrict__ info,
                                BFlags* flags, const double* part, RedSpec spec,
                                double* scal_all, int mode, int first, double tau, double tol,
                                int max_iter) {
  const int inst = ids[blockIdx.x];
  if (flags[inst].done) return;
  const BInst& I = info[inst];
  const int np = vseg.first[inst + 1] - vseg.first[inst];
  const double* pp = part + I.pbase;
  double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double inv2n = 0.5 / (double)I.N;
  __shared__ double s_sum[4];
  for (int q = 0; q < spec.nq; ++q) {
    const int op = spec.op[q];
    double acc = red_ident(op);
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
      const double v = pp[q * np + i];
      acc = op == 0 ? acc + v : (op == 1 ? fmax(acc, v) : fmin(acc, v));
    }
    double r;
    if (op == 0) r = block_reduce(acc, OpSum());
    else if (op == 1) r = block_reduce(acc, OpMax());
    else r = block_reduce(acc, OpMin());
    if (threadIdx.x == 0) {
      if (mode == 0) scal[first + q] = r;
      else if (mode == 1 || (mode == 4 && q < 2)) scal[S_AP_AFF + q] = fmin(1.0, r);
      else if (mode == 4) s_sum[q - 2] = r;
      else if (mode == 2) {
        const double ratio = r * inv2n / scal[S_MU];
        scal[S_SIGMA] = fmin(1.0, ratio * ratio * ratio);
      } else {
        scal[S_AP + q] = fmin(1.0, tau * r);
      }
    }
    __syncthreads();
  }
  if (mode == 4 && threadIdx.x == 0) {
    const double ap = scal[S_AP_AFF], ad = scal[S_AD_AFF];
    const double muaff = (s_sum[0] + ad * s_sum[1] + ap * s_sum[2] + ap * ad * s_sum[3]) * inv2n;
    const double ratio = muaff / scal[S_MU];
    scal[S_SIGMA] = fmin(1.0, ratio * ratio * ratio);
  }
  if (mode == 0 && first == S_RB && threadIdx.x == 0) {
    BFlags& fl = flags[inst];
    const double mu = scal[S_TZ] * inv2n;
    scal[S_MU] = mu;
    const double f = 0.5 * scal[S_BQB] - scal[S_BB] + I.lam * scal[S_SUMC];
    const double dual = fmax(scal[S_RB], scal[S_RC]);
    fl.f = f;
    fl.dual = dual;
    fl.mu = mu;
    if (!isfinite(f) || !isfinite(mu) || !isfinite(dual)) {
      fl.status = 2;
      fl.done = 1;
    } else if (dual <= tol * (1.0 + scal[S_BNORM]) && mu <= tol * (1.0 + fabs(f) / (double)I.N)) {
      fl.status = 0;
      fl.done = 1;
    } else if (fl.iters >= max_iter) {
      fl.status = 1;
      fl.done = 1;
    } else {
      fl.iters += 1;
    }
  }
}

__global__ __launch_bounds__(256) void bk_absmax(BSeg seg, const BInst* __restrict__ info,
                                                 const BFlags* flags,
                                                 const double* __restrict__ x, double* part) {
  BSEG_BEGIN
  double a[1] = {0.0};
  BFOR(i) a[0] = fmax(a[0], fabs(x[i]));
  const int ops[1] = {1};
  bstore_partials<1>(part + I.pbase, a, ops, lb, nbk);
}

__global__ __launch_bounds__(256) void bk_ipm_init(BSeg seg, const BInst* __restrict__ info,
                                                   BFlags* flags, double* beta, double* c,
                                                   double* z1, double* z2, double* Qb,
                                                   const double* scal_all,
                                                   const uint32_t* valid) {
  BSEG_BEGIN
  const double cval = 1.0 + scal_all[(int64_t)inst * S_NSLOTS + S_BNORM];
  const double zval = I.lam > 0 ? 0.5 * I.lam : 1.0;
  BFOR(i) {
    const bool ok = BVALID(valid, i);
    beta[i] = 0.0;
    Qb[i] = 0.0;   // M^T M beta
    c[i] = ok ? cval : 1.0;
    z1[i] = ok ? zval : 0.0;
    z2[i] = ok ? zval : 0.0;
  }
}

// residuals and objective pieces: max|r_b|, max|r_c|, sum t'z, beta'Qbeta, beta'b, sum c
__global__ __launch_bounds__(256) void bk_ipm_resid(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ Qb, const double* __restrict__ b, double* part,
    const uint32_t* valid) {
  BSEG_BEGIN
  const double lam = I.lam;
  double a[6] = {0, 0, 0, 0, 0, 0};
  BFOR(i) {
    if (!BVALID(valid, i)) continue;
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i], q = Qb[i], bb = b[i];
    const double t1 = ce + be, t2 = ce - be;
    const double rb = q - bb - y1 + y2;
    const double rc = lam - y1 - y2;
    a[0] = fmax(a[0], fabs(rb));
    a[1] = fmax(a[1], fabs(rc));
    a[2] += t1 * y1 + t2 * y2;
    a[3] += be * q;
    a[4] += be * bb;
    a[5] += ce;
  }
  const int ops[6] = {1, 1, 0, 0, 0, 0};
  bstore_partials<6>(part + I.pbase, a, ops, lb, nbk);
}

// FP64 path: slack diagonals, preconditioner of the condensed system (in the
// CG precision TC) and the predictor right-hand side
// [-r_b + g1 - g2; -r_c + g1 + g2] with g_i = -z_i (two-block rhs)
template <typename TC>
__global__ __launch_bounds__(256) void bk_ipm_prep(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ Qb, const double* __restrict__ b, double* s1o, double* s2o,
    TC* L1, TC* L2, TC* P11, TC* P12, TC* P22, TC* rhs, const uint32_t* valid) {
  BSEG_BEGIN
  const double lam = I.lam;
  BFOR(i) {
    double s1 = 0, s2 = 0, rb = 0, rc = 0, y1 = 0, y2 = 0;
    double l1 = 0, l2 = 0, p11 = 1, p12 = 0, p22 = 1;
    const bool ok = BVALID(valid, i);
    if (ok) {
      const double be = beta[i], ce = c[i];
      y1 = z1[i];
      y2 = z2[i];
      const double t1 = ce + be, t2 = ce - be;
      s1 = y1 / t1;
      s2 = y2 / t2;
      precond_entries<double>(s1, s2, l1, l2, p11, p12, p22);
      rb = Qb[i] - b[i] - y1 + y2;
      rc = lam - y1 - y2;
    }
    s1o[i] = s1;
    s2o[i] = s2;
    L1[i] = (TC)l1;
    L2[i] = (TC)l2;
    P11[i] = (TC)p11;
    P12[i] = (TC)p12;
    P22[i] = (TC)p22;
    const int64_t j = i + I.off;
    rhs[j] = (TC)(ok ? -rb - y1 + y2 : 0.0);
    rhs[j + I.n] = (TC)(ok ? -rc - y1 - y2 : 0.0);
  }
}

// Mixed path: the residual partials of bk_ipm_resid and, from the same
// loads, the reduced system (the slack diagonals S1, S2; with L1 = S1 + S2,
// L2 = S1 - S2, the reduced diagonal D = 4 S1 S2 / L1 in FP64 and in TC with
// the preconditioner 1/(1 + D); rhsd = [h_b - (L2/L1) h_c; h_c] in FP64 and
// its first block in TC).  L1 and L2 are not stored: their users form them
// from S1 and S2.  An instance whose run ends at this step computes the
// system too, unused.
template <typename TC>
__global__ __launch_bounds__(256) void bk_ipm_resid_prep_red(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ Qb, const double* __restrict__ b, double* part, double* s1o,
    double* s2o, TC* D, TC* P, TC* rhs, double* Dd, double* rhsd, const uint32_t* valid) {
  BSEG_BEGIN
  const double lam = I.lam;
  double a[6] = {0, 0, 0, 0, 0, 0};
  BFOR(i) {
    double s1 = 0, s2 = 0, dg = 0, hb = 0, hc = 0;
    if (BVALID(valid, i)) {
      const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i], q = Qb[i], bb = b[i];
      const double t1 = ce + be, t2 = ce - be;
      const double rb = q - bb - y1 + y2;
      const double rc = lam - y1 - y2;
      a[0] = fmax(a[0], fabs(rb));
      a[1] = fmax(a[1], fabs(rc));
      a[2] += t1 * y1 + t2 * y2;
      a[3] += be * q;
      a[4] += be * bb;
      a[5] += ce;
      s1 = y1 / t1;
      s2 = y2 / t2;
      const double l1 = s1 + s2, l2 = s1 - s2;
      dg = 4.0 * s1 * s2 / l1;
      hb = -rb - y1 + y2;
      hc = -rc - y1 - y2;
      hb -= (l2 / l1) * hc;
    }
    s1o[i] = s1;
    s2o[i] = s2;
    Dd[i] = dg;
    D[i] = (TC)dg;
    P[i] = (TC)(1.0 / (1.0 + dg));
    const int64_t j = i + I.off;
    rhsd[j] = hb;
    rhsd[j + I.n] = hc;
    rhs[j] = (TC)hb;
  }
  const int ops[6] = {1, 1, 0, 0, 0, 0};
  bstore_partials<6>(part + I.pbase, a, ops, lb, nbk);
}

// y = x (first nb blocks), with the precision change TI -> TO
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void bk_convert(BSeg seg, const BInst* __restrict__ info,
                                                  const BFlags* flags,
                                                  const TI* __restrict__ x, TO* y, int nb) {
  BSEG_BEGIN
  BFOR(i) {
    const int64_t j = i + I.off;
    y[j] = (TO)x[j];
    if (nb == 2) y[j + I.n] = (TO)x[j + I.n];
  }
}

// the refinement's correction x (TC, reduced system) into the FP64 direction
// y, and M^T M x = b - r - D x (b, r: that CG's right-hand side and final
// recursive residual, D the reduced diagonal) into the second block of mtm,
// whose first block the FP64 apply of the uncorrected y left M^T M y in
template <typename TC>
__global__ __launch_bounds__(256) void bk_refine_add(BSeg seg, const BInst* __restrict__ info,
                                                     const BFlags* flags, double* y,
                                                     const TC* __restrict__ x, double* mtm,
                                                     const TC* __restrict__ b,
                                                     const TC* __restrict__ r,
                                                     const TC* __restrict__ D) {
  BSEG_BEGIN
  BFOR(i) {
    const int64_t j = i + I.off;
    const double xv = (double)x[j];
    y[j] += xv;
    mtm[j + I.n] += (double)b[j] - (double)r[j] - (double)D[i] * xv;
  }
}

// direction of the step: dt_i, dz_i and the largest steps to the boundary.
// corr = 0: affine (g_i = -z_i); corr = 1: corrector with
// g_i = (sigma mu - t_i z_i - dt_i^aff dz_i^aff) / t_i (dt, dz hold the affine
// direction on entry and are overwritten).  d: two-block direction in TD.
template <typename TD>
__global__ __launch_bounds__(256) void bk_ipm_dir(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ s1, const double* __restrict__ s2,
    const TD* __restrict__ d, const double* scal_all, int corr, double* dt1o, double* dt2o,
    double* dz1o, double* dz2o, double* part, const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double smu = corr ? scal[S_SIGMA] * scal[S_MU] : 0.0;
  double a[2] = {1e300, 1e300};
  BFOR(i) {
    if (!BVALID(valid, i)) {
      dt1o[i] = dt2o[i] = dz1o[i] = dz2o[i] = 0.0;
      continue;
    }
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    double g1 = -y1, g2 = -y2;
    if (corr) {
      g1 = (smu - t1 * y1 - dt1o[i] * dz1o[i]) / t1;
      g2 = (smu - t2 * y2 - dt2o[i] * dz2o[i]) / t2;
    }
    const int64_t j = i + I.off;
    const double db = (double)d[j], dc = (double)d[j + I.n];
    const double dt1 = dc + db, dt2 = dc - db;
    const double dz1 = g1 - s1[i] * dt1;
    const double dz2 = g2 - s2[i] * dt2;
    dt1o[i] = dt1;
    dt2o[i] = dt2;
    dz1o[i] = dz1;
    dz2o[i] = dz2;
    if (dt1 < 0) a[0] = fmin(a[0], -t1 / dt1);
    if (dt2 < 0) a[0] = fmin(a[0], -t2 / dt2);
    if (dz1 < 0) a[1] = fmin(a[1], -y1 / dz1);
    if (dz2 < 0) a[1] = fmin(a[1], -y2 / dz2);
  }
  const int ops[2] = {2, 2};
  bstore_partials<2>(part + I.pbase, a, ops, lb, nbk);
}

// Mixed path direction: the c-block of the reduced system's solution by the
// back-substitution dc = (h_c - L2 db) / L1 (db = x, h_c the second block of
// rhsd, L1 = S1 + S2, L2 = S1 - S2), then dt_i, dz_i and the largest steps to
// the boundary as bk_ipm_dir.  corr = 0: x is the predictor's CG solution
// (TX = TC) and the affine dt, dz are written; corr = 1: x is the refined
// corrector direction (TX = double) and dz and dc (into the second block of
// dco, for the update, which takes db and dc) are written, dt is not.  The
// affine pass also leaves the sums t'z, t'dz, dt'z, dt'dz (partials 2-5), of
// which the complementarity after the affine step is the bilinear form
// mu_aff 2N = t'z + a_d t'dz + a_p dt'z + a_p a_d dt'dz (finalize mode 4).
template <typename TX>
__global__ __launch_bounds__(256) void bk_ipm_dir_red(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ s1, const double* __restrict__ s2, const TX* x,
    const double* __restrict__ rhsd, const double* scal_all, int corr, double* dt1o,
    double* dt2o, double* dz1o, double* dz2o, double* dco, double* part,
    const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double smu = corr ? scal[S_SIGMA] * scal[S_MU] : 0.0;
  double a[6] = {1e300, 1e300, 0.0, 0.0, 0.0, 0.0};
  BFOR(i) {
    const int64_t j = i + I.off;
    if (!BVALID(valid, i)) {
      dz1o[i] = dz2o[i] = 0.0;
      if (corr) dco[j + I.n] = 0.0;
      else dt1o[i] = dt2o[i] = 0.0;
      continue;
    }
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double sa = s1[i], sb = s2[i];
    const double t1 = ce + be, t2 = ce - be;
    double g1 = -y1, g2 = -y2;
    if (corr) {
      g1 = (smu - t1 * y1 - dt1o[i] * dz1o[i]) / t1;
      g2 = (smu - t2 * y2 - dt2o[i] * dz2o[i]) / t2;
    }
    const double db = (double)x[j];
    const double dc = (rhsd[j + I.n] - (sa - sb) * db) / (sa + sb);
    const double dt1 = dc + db, dt2 = dc - db;
    const double dz1 = g1 - sa * dt1;
    const double dz2 = g2 - sb * dt2;
    if (corr) {
      dco[j + I.n] = dc;
    } else {
      dt1o[i] = dt1;
      dt2o[i] = dt2;
      a[2] += t1 * y1 + t2 * y2;
      a[3] += t1 * dz1 + t2 * dz2;
      a[4] += dt1 * y1 + dt2 * y2;
      a[5] += dt1 * dz1 + dt2 * dz2;
    }
    dz1o[i] = dz1;
    dz2o[i] = dz2;
    if (dt1 < 0) a[0] = fmin(a[0], -t1 / dt1);
    if (dt2 < 0) a[0] = fmin(a[0], -t2 / dt2);
    if (dz1 < 0) a[1] = fmin(a[1], -y1 / dz1);
    if (dz2 < 0) a[1] = fmin(a[1], -y2 / dz2);
  }
  if (corr) {
    const double b2[2] = {a[0], a[1]};
    const int ops[2] = {2, 2};
    bstore_partials<2>(part + I.pbase, b2, ops, lb, nbk);
  } else {
    const int ops[6] = {2, 2, 0, 0, 0, 0};
    bstore_partials<6>(part + I.pbase, a, ops, lb, nbk);
  }
}

// complementarity after the affine step
__global__ __launch_bounds__(256) void bk_ipm_muaff(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal_all,
    double* part, const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double ap = scal[S_AP_AFF], ad = scal[S_AD_AFF];
  double a[1] = {0.0};
  BFOR(i) {
    if (!BVALID(valid, i)) continue;
    const double be = beta[i], ce = c[i];
    const double t1 = ce + be, t2 = ce - be;
    a[0] += (t1 + ap * dt1[i]) * (z1[i] + ad * dz1[i]) + (t2 + ap * dt2[i]) * (z2[i] + ad * dz2[i]);
  }
  const int ops[1] = {0};
  bstore_partials<1>(part + I.pbase, a, ops, lb, nbk);
}

// corrector right-hand side of the condensed system (FP64 path), two-block,
// in TC
template <typename TC>
__global__ __launch_bounds__(256) void bk_ipm_rhs_corr(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ z1, const double* __restrict__ z2,
    const double* __restrict__ Qb, const double* __restrict__ b,
    const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal_all,
    TC* rhs, const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double smu = scal[S_SIGMA] * scal[S_MU];
  const double lam = I.lam;
  BFOR(i) {
    double hb = 0, hc = 0;
    if (BVALID(valid, i)) {
      const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
      const double t1 = ce + be, t2 = ce - be;
      const double g1 = (smu - t1 * y1 - dt1[i] * dz1[i]) / t1;
      const double g2 = (smu - t2 * y2 - dt2[i] * dz2[i]) / t2;
      const double rb = Qb[i] - b[i] - y1 + y2;
      const double rc = lam - y1 - y2;
      hb = -rb + g1 - g2;
      hc = -rc + g1 + g2;
    }
    const int64_t j = i + I.off;
    rhs[j] = (TC)hb;
    rhs[j + I.n] = (TC)hc;
  }
}

// Mixed path corrector right-hand side, from the predictor's: the
// corrector's g_i = (sigma mu - t_i z_i - dt_i dz_i) / t_i (affine dt, dz)
// exceed the predictor's g_i = -z_i by delta_i = (sigma mu - dt_i dz_i) / t_i,
// so h_b grows by delta_1 - delta_2 and h_c by delta_1 + delta_2: rhsd (the
// predictor's reduced right-hand side on entry) gets
// [dh_b - (L2/L1) dh_c; dh_c] added, rhs is its new first block in TC.
// r (when given) holds the predictor CG's final residual h_b - K db and gets
// dh_b added: the residual of the predictor's db for the corrector's system,
// the corrector CG's warm start.  Invalid slots keep their zeros.
template <typename TC>
__global__ __launch_bounds__(256) void bk_ipm_rhs_corr_red(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags,
    const double* __restrict__ beta, const double* __restrict__ c,
    const double* __restrict__ s1, const double* __restrict__ s2,
    const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal_all,
    TC* rhs, double* rhsd, TC* r, const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double smu = scal[S_SIGMA] * scal[S_MU];
  BFOR(i) {
    if (!BVALID(valid, i)) continue;
    const double be = beta[i], ce = c[i], sa = s1[i], sb = s2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double d1 = (smu - dt1[i] * dz1[i]) / t1;
    const double d2 = (smu - dt2[i] * dz2[i]) / t2;
    const double dhc = d1 + d2;
    const double dhb = (d1 - d2) - ((sa - sb) / (sa + sb)) * dhc;
    const int64_t j = i + I.off;
    const double hb = rhsd[j] + dhb;
    rhsd[j] = hb;
    rhsd[j + I.n] += dhc;
    rhs[j] = (TC)hb;
    if (r) r[j] = (TC)((double)r[j] + dhb);
  }
}

// step: beta, c by the primal length, z by the dual length
template <typename TD>
__global__ __launch_bounds__(256) void bk_ipm_update(
    BSeg seg, const BInst* __restrict__ info, const BFlags* flags, double* beta, double* c,
    double* z1, double* z2, const TD* __restrict__ d, const double* __restrict__ dz1,
    const double* __restrict__ dz2, double* Qb, const double* __restrict__ mtm,
    const double* scal_all, const uint32_t* valid) {
  BSEG_BEGIN
  const double* scal = scal_all + (int64_t)inst * S_NSLOTS;
  const double ap = scal[S_AP], ad = scal[S_AD];
  BFOR(i) {
    if (!BVALID(valid, i)) continue;
    const int64_t j = i + I.off;
    beta[i] += ap * (double)d[j];
    if (Qb) Qb[i] += ap * mtm[j + I.n];   // Q beta carried with beta (M^T M d in mtm)
    c[i] += ap * (double)d[j + I.n];
    z1[i] += ad * dz1[i];
    z2[i] += ad * dz2[i];
  }
}

// ===========================================================================
// ADMM (src/admm.jl:8-61) over a batch, in the spectrum order: b0 = M^T z0,
// x0 (previous x), z, y in FP64; each iteration solves
// (M^T M + rho I) x = b0 + rho z - y by CG (the reduced-system apply with
// D = rho, no preconditioner), then z = S_{lambda/rho}(x + y/rho),
// y += rho (x - z), err = max(||x - x0||, ||y - y0||, ||z - z0||).  The soft
// threshold is odd, so it commutes with the signed permutation into u.
// ===========================================================================
__global__ __launch_bounds__(256) void k_fill(double* x, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

template <typename TC>
__global__ __launch_bounds__(256) void bk_admm_setup(BSeg seg, const BInst* __restrict__ info,
                                                     const BFlags* flags, const double* x0,
                                                     double* z, double* y, TC* D, TC* P,
                                                     double rho, const uint32_t* valid) {
  BSEG_BEGIN
  BFOR(i) {
    const bool ok = BVALID(valid, i);
    z[i] = x0[i];
    y[i] = 0.0;
    D[i] = (TC)(ok ? rho : 0.0);
    P[i] = (TC)1;
  }
}

// b = b0 + rho z - y (first block of the CG right-hand side)
template <typename TC>
__global__ __launch_bounds__(256) void bk_admm_rhs(BSeg seg, const BInst* __restrict__ info,
                                                   const BFlags* flags,
                                                   const double* __restrict__ b0,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ y, double rho,
                                                   TC* rhs, const uint32_t* valid) {
  BSEG_BEGIN
  BFOR(i) rhs[i + I.off] = (TC)(BVALID(valid, i) ? b0[i] + rho * z[i] - y[i] : 0.0);
}

__device__ __forceinline__ double soft_threshold(double x, double t) {
  return x > t ? x - t : (x < -t ? x + t : 0.0);
}

// x from CG; z, y updates in place; partials of ||x - x0||^2, ||y - y0||^2, ||z - z0||^2
template <typename TC>
__global__ __launch_bounds__(256) void bk_admm_update(BSeg seg, const BInst* __restrict__ info,
                                                      const BFlags* flags,
                                                      const TC* __restrict__ xs, double* x0,
                                                      double* z, double* y, double rho,
                                                      double* part, const uint32_t* valid) {
  BSEG_BEGIN
  const double thr = I.lam / rho;
  double a[3] = {0, 0, 0};
  BFOR(i) {
    if (!BVALID(valid, i)) continue;
    const double x1 = (double)xs[i + I.off];
    const double yo = y[i], zo = z[i];
    const double z1 = soft_threshold(x1 + yo / rho, thr);
    const double y1 = yo + rho * (x1 - z1);
    const double dx = x1 - x0[i], dy = y1 - yo, dz = z1 - zo;
    a[0] += dx * dx;
    a[1] += dy * dy;
    a[2] += dz * dz;
    x0[i] = x1;
    z[i] = z1;
    y[i] = y1;
  }
  const int ops[3] = {0, 0, 0};
  bstore_partials<3>(part + I.pbase, a, ops, lb, nbk);
}

// one block per instance: err, the iteration count and the stopping test
// (admm.jl:32: while t < maxt and err > tol)
__global__ void bk_admm_finalize(BSeg vseg, const int* ids, const BInst* __restrict__ info,
                                 BFlags* flags, const double* part, double tol, int maxt) {
  const int inst = ids[blockIdx.x];
  if (flags[inst].done) return;
  const BInst& I = info[inst];
  const int np = vseg.first[inst + 1] - vseg.first[inst];
  double e[3];
  for (int q = 0; q < 3; ++q) {
    e[q] = sum_partials(part + I.pbase + q * np, np);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    BFlags& fl = flags[inst];
    const double err = fmax(sqrt(e[0]), fmax(sqrt(e[1]), sqrt(e[2])));
    fl.iters += 1;
    fl.dual = err;
    if (!isfinite(err)) {
      fl.status = 2;
      fl.done = 1;
    } else if (err <= tol) {
      fl.status = 0;
      fl.done = 1;
    } else if (fl.iters >= maxt) {
      fl.status = 1;
      fl.done = 1;
    }
  }
}

#undef BSEG_BEGIN
#undef BFOR
#undef BVALID

}  // namespace csipm
