// __global__ kernels of libcsipm (included once, by csipm.hip).
#pragma once

#include "csipm_device.hpp"

namespace csipm {

// ===========================================================================
// Standalone packing maps (FP64, bit-exact for even sizes)
// ===========================================================================

// v[k1 + ld1*(k2 + N2*k3)] for every element of the (ld1, N2, N3) spectrum.
template <int D>
__global__ __launch_bounds__(256) void k_beta_to_dft(const double* __restrict__ beta,
                                                     cpx<double>* __restrict__ v, MapGeom g,
                                                     int ld1, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int k1 = (int)(t % ld1);
  const int64_t r = t / ld1;
  const int k2 = D >= 2 ? (int)(r % g.n[1]) : 0;
  const int k3 = D >= 3 ? (int)(r / g.n[1]) : 0;
  v[t] = beta_to_dft_elem<D, double>(beta, k1, k2, k3, g);
}

// Gather form of DFT_to_beta, one beta index at a time
// (src/mapping_cpu_kernels.jl:11-21, 38-87, 109-351).  Decodes the block of
// index i, the element it represents and reads it from the half spectrum
// (3D second-half dim-1 elements: conjugate partner, Im negated).
__host__ __device__ __forceinline__ int self_index(int s, const MapGeom& g, int dim) {
  return s == 0 ? 0 : g.M[dim];
}

template <int D>
__host__ __device__ __forceinline__ double dft_to_beta_elem(const cpx<double>* __restrict__ v,
                                                            int64_t i, const MapGeom& g,
                                                            int ld1) {
#pragma clang fp contract(off)
  int e[3] = {0, 0, 0};   // represented element (full-spectrum coordinates)
  int comp = 0;           // 0 Re, 1 Im
  bool vertex = false;
  const int P1 = g.P[0], P2 = g.P[1], P3 = g.P[2];
  const int S2 = g.S[1], S3 = g.S[2];
  // R offset r -> index r+1 ; H offset h -> index M+1+h
  if (D == 1) {
    if (i < g.off[1]) {
      vertex = true;
      e[0] = self_index((int)i, g, 0);
    } else {
      const int64_t t = i - g.off[1];
      comp = (int)(t / P1);
      e[0] = (int)(t % P1) + 1;
    }
  } else if (D == 2) {
    if (i < g.off[1]) {
      vertex = true;
      e[0] = self_index((int)i / S2, g, 0);
      e[1] = self_index((int)i % S2, g, 1);
    } else if (i < g.off[2]) {
      const int64_t t = i - g.off[1];
      const int q = (int)(t / P2);
      comp = q & 1;
      e[0] = self_index(q >> 1, g, 0);
      e[1] = (int)(t % P2) + 1;
    } else if (i < g.off[3]) {
      const int64_t t = i - g.off[2];
      const int q = (int)(t / P1);
      comp = q & 1;
      e[1] = self_index(q >> 1, g, 1);
      e[0] = (int)(t % P1) + 1;
    } else {
      const int64_t t = i - g.off[3];
      const int q = (int)(t / g.P12);   // 0: Re(R,R) 1: Im(R,R) 2: Re(R,H) 3: Im(R,H)
      const int64_t u = t % g.P12;
      comp = q & 1;
      e[0] = (int)(u % P1) + 1;
      const int c2 = (int)(u / P1);
      e[1] = (q >= 2) ? g.M[1] + 1 + c2 : c2 + 1;
    }
  } else {
    if (i < g.off[1]) {
      vertex = true;
      const int t = (int)i;
      e[2] = self_index(t % S3, g, 2);
      e[1] = self_index((t / S3) % S2, g, 1);
      e[0] = self_index(t / (S3 * S2), g, 0);
    } else if (i < g.off[2]) {           // edges along dim 3, groups (s1, s2)
      const int64_t t = i - g.off[1];
      const int q = (int)(t / P3);
      comp = q & 1;
      const int grp = q >> 1;
      e[0] = self_index(grp / S2, g, 0);
      e[1] = self_index(grp % S2, g, 1);
      e[2] = (int)(t % P3) + 1;
    } else if (i < g.off[3]) {           // edges along dim 2, groups (s1, s3)
      const int64_t t = i - g.off[2];
      const int q = (int)(t / P2);
      comp = q & 1;
      const int grp = q >> 1;
      e[0] = self_index(grp / S3, g, 0);
      e[2] = self_index(grp % S3, g, 2);
      e[1] = (int)(t % P2) + 1;
    } else if (i < g.off[4]) {           // edges along dim 1, groups (s2, s3)
      const int64_t t = i - g.off[3];
      const int q = (int)(t / P1);
      comp = q & 1;
      const int grp = q >> 1;
      e[1] = self_index(grp / S3, g, 1);
      e[2] = self_index(grp % S3, g, 2);
      e[0] = (int)(t % P1) + 1;
    } else if (i < g.off[5]) {           // faces, self-conjugate dim 1
      const int64_t t = i - g.off[4];
      const int q = (int)(t / g.P23);
      const int64_t u = t % g.P23;
      comp = q & 1;
      e[0] = self_index(q >> 2, g, 0);
      const int c2 = (int)(u % P2);
      e[2] = (int)(u / P2) + 1;
      e[1] = ((q & 2) ? g.M[1] + 1 + c2 : c2 + 1);
    } else if (i < g.off[6]) {           // faces, self-conjugate dim 2
      const int64_t t = i - g.off[5];
      const int q = (int)(t / g.P13);
      const int64_t u = t % g.P13;
      comp = q & 1;
      e[1] = self_index(q >> 2, g, 1);
      const int c1 = (int)(u % P1);
      e[2] = (int)(u / P1) + 1;
      e[0] = ((q & 2) ? g.M[0] + 1 + c1 : c1 + 1);
    } else if (i < g.off[7]) {           // faces, self-conjugate dim 3
      const int64_t t = i - g.off[6];
      const int q = (int)(t / g.P12);
      const int64_t u = t % g.P12;
      comp = q & 1;
      e[2] = self_index(q >> 2, g, 2);
      const int c1 = (int)(u % P1);
      e[1] = (int)(u / P1) + 1;
      e[0] = ((q & 2) ? g.M[0] + 1 + c1 : c1 + 1);
    } else {                             // interior
      const int64_t t = i - g.off[7];
      const int q = (int)(t / g.P123);   // 2*blk + comp, blk: (R,R) (H,R) (R,H) (H,H)
      const int64_t u = t % g.P123;
      comp = q & 1;
      const int blk = q >> 1;
      const int c1 = (int)(u % P1);
      const int64_t u2 = u / P1;
      const int c2 = (int)(u2 % P2);
      e[2] = (int)(u2 / P2) + 1;
      e[0] = (blk & 1) ? g.M[0] + 1 + c1 : c1 + 1;
      e[1] = (blk & 2) ? g.M[1] + 1 + c2 : c2 + 1;
    }
  }
  double sign = 1.0;
  if (e[0] > g.M[0]) {  // second half of dim 1: read the conjugate partner
    e[0] = g.n[0] - e[0];
    e[1] = e[1] == 0 ? 0 : g.n[1] - e[1];
    e[2] = e[2] == 0 ? 0 : g.n[2] - e[2];
    sign = -1.0;
  }
  const cpx<double> val = v[e[0] + (int64_t)ld1 * (e[1] + (int64_t)g.n[1] * e[2])];
  const double sq2 = 1.4142135623730951;
  if (vertex) return val.x;
  if (comp == 0) return sq2 * val.x;
  return (sign * sq2) * val.y;
}

template <int D>
__global__ __launch_bounds__(256) void k_dft_to_beta(const cpx<double>* __restrict__ v,
                                                     double* __restrict__ beta, MapGeom g,
                                                     int ld1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  beta[i] = dft_to_beta_elem<D>(v, i, g, ld1);
}

// ===========================================================================
// Destinations of the forward passes that end in beta space
// ===========================================================================
template <typename T>
struct Epi {
  int kind;             // 0: out[i] = value          1: condensed KKT epilogue
  T* out;
  const T* xb;
  const T* xz;
  const T* L1;
  const T* L2;
  T* yb;
  T* yz;
  T alpha, beta;
  double* dot_part;     // per-block partial sums of x'y (CG p'q), or null
};

// writes value m of beta index i; returns its contribution to x'y
template <typename T>
__device__ __forceinline__ double epi_apply(const Epi<T>& e, int64_t i, T m) {
  if (e.kind == 0) {
    e.out[i] = m;
    return 0.0;
  }
  const T xb = e.xb[i], xz = e.xz[i], l1 = e.L1[i], l2 = e.L2[i];
  T qb = e.alpha * (m + l1 * xb + l2 * xz);
  T qz = e.alpha * (l2 * xb + l1 * xz);
  if (e.beta != (T)0) {
    qb += e.beta * e.yb[i];
    qz += e.beta * e.yz[i];
  }
  e.yb[i] = qb;
  e.yz[i] = qz;
  return (double)xb * (double)qb + (double)xz * (double)qz;
}

// value v of half-spectrum element (k1,k2,k3) -> the beta entries reading it
template <int D, typename T>
__device__ __forceinline__ double scatter_elem(const Epi<T>& e, int k1, int k2, int k3,
                                               cpx<T> v, const MapGeom& g) {
  int64_t re, im;
  bool vertex, neg;
  if (!scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) return 0.0;
  const T sq2 = (T)1.4142135623730951;
  if (vertex) return epi_apply(e, re, v.x);
  double acc = epi_apply(e, re, sq2 * v.x);
  acc += epi_apply(e, im, (neg ? -sq2 : sq2) * v.y);
  return acc;
}

// per-block partial of a dot product (deterministic; finalised by a later kernel)
__device__ __forceinline__ void store_partial(double* part, double acc) {
  const double s = block_reduce(acc, OpSum());
  if (threadIdx.x == 0 && part) part[blockIdx.x] = s;
}

// ===========================================================================
// Line FFTs along dimension 2 or 3 of the (H1, N2, N3) half spectrum.
// A workgroup takes nl lines with consecutive k1 (coalesced strided access).
// ===========================================================================
enum { SRC_BUF = 0, SRC_BETA = 1, SRC_REAL = 2 };
enum { DST_BUF = 0, DST_BETA = 1, DST_REAL = 2 };

template <typename T>
struct LineArgs {
  MapGeom g;
  FftPlan plan;
  const cpx<T>* tw;
  int axis;     // 1: lines along dim 2; 2: lines along dim 3
  int H1;
  int nl;
  int ntile1;
  cpx<T>* buf;
  const T* beta_in;
  Epi<T> epi;
  const int* done;
};

template <typename T, int NT, int E, int D, int SRC, int DST, bool FWD>
__global__ __launch_bounds__(NT) void k_lines(LineArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const MapGeom& g = a.g;
  const int L = a.plan.L;
  const int tile1 = blockIdx.x % a.ntile1;
  const int o = blockIdx.x / a.ntile1;
  const int k10 = tile1 * a.nl;
  const int nl = min(a.nl, a.H1 - k10);
  const int n2 = g.n[1];
  const int64_t s2 = a.H1, s3 = (int64_t)a.H1 * n2;
  const int64_t sl = a.axis == 1 ? s2 : s3;                  // stride along the line
  const int64_t base = k10 + (a.axis == 1 ? (int64_t)o * s3 : (int64_t)o * s2);
  const int total = nl * L;
  for (int idx = threadIdx.x; idx < total; idx += NT) {
    const int line = idx % nl;
    const int el = idx / nl;
    cpx<T> val;
    if (SRC == SRC_BUF) {
      val = a.buf[base + line + el * sl];
    } else {
      const int k2 = a.axis == 1 ? el : o;
      const int k3 = a.axis == 1 ? o : el;
      val = beta_to_dft_elem<D, T>(a.beta_in, k10 + line, k2, k3, g);
    }
    lds[line * L + el] = val;
  }
  __syncthreads();
  lds_fft<T, E, FWD>(lds, nl, L, a.plan, a.tw);
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < total; idx += NT) {
    const int line = idx % nl;
    const int el = idx / nl;
    const cpx<T> val = lds[line * L + el];
    if (DST == DST_BUF) {
      a.buf[base + line + el * sl] = val;
    } else {
      const int k2 = a.axis == 1 ? el : o;
      const int k3 = a.axis == 1 ? o : el;
      acc += scatter_elem<D, T>(a.epi, k10 + line, k2, k3, val, g);
    }
  }
  if (DST == DST_BETA && a.epi.dot_part) store_partial(a.epi.dot_part, acc);
}

// ===========================================================================
// Dimension-1 pass: two real lines per complex FFT of length N1.
// MODE 0 (MtM): half spectra -> C2R -> mask*scale -> R2C -> half spectra
// MODE 1 (C2R): half spectra -> C2R -> mask*scale -> real samples
// MODE 2 (R2C): real samples*scale -> R2C -> half spectra
// ===========================================================================
enum { MODE_MTM = 0, MODE_C2R = 1, MODE_R2C = 2 };

template <typename T>
struct Dim1Args {
  MapGeom g;
  FftPlan plan;
  const cpx<T>* tw;
  int N1, H1;
  int nlines;     // real lines = N2*N3
  int npairs;
  int np;         // pairs per workgroup
  cpx<T>* buf;
  const T* beta_in;
  const T* real_in;
  T* real_out;
  Epi<T> epi;
  const uint32_t* miss;   // bit set = missing sample (nullable)
  T scale;
  const int* done;
};

template <typename T, int D, int SRC>
__device__ __forceinline__ cpx<T> d1_load(const Dim1Args<T>& a, int line, int k) {
  if (SRC == SRC_BUF) return a.buf[(int64_t)line * a.H1 + k];
  return beta_to_dft_elem<D, T>(a.beta_in, k, 0, 0, a.g);   // 1D: line 0
}

__device__ __forceinline__ bool miss_bit(const uint32_t* m, int64_t s) {
  return m && ((m[s >> 5] >> (s & 31)) & 1u);
}

template <typename T, int NT, int E, int D, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void k_dim1(Dim1Args<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const int N1 = a.N1, H1 = a.H1;
  const int p0 = blockIdx.x * a.np;
  const int npl = min(a.np, a.npairs - p0);
  const bool even = (N1 & 1) == 0;
  if (MODE != MODE_R2C) {
    const int tot = npl * H1;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      const int p = idx / H1;
      const int k = idx - p * H1;
      const int la = 2 * (p0 + p), lb = la + 1;
      cpx<T> A = d1_load<T, D, SRC>(a, la, k);
      cpx<T> B = lb < a.nlines ? d1_load<T, D, SRC>(a, lb, k) : cpx<T>{0, 0};
      if (k == 0 || (even && k == H1 - 1)) {  // DC / Nyquist of a real line are real
        A.y = 0;
        B.y = 0;
      }
      cpx<T>* z = lds + p * N1;
      z[k] = {A.x - B.y, A.y + B.x};
      if (k > 0 && N1 - k != k) z[N1 - k] = {A.x + B.y, B.x - A.y};
    }
    __syncthreads();
    lds_fft<T, E, false>(lds, npl, N1, a.plan, a.tw);
    const int totr = npl * N1;
    for (int idx = threadIdx.x; idx < totr; idx += NT) {
      const int p = idx / N1;
      const int n = idx - p * N1;
      const int la = 2 * (p0 + p), lb = la + 1;
      cpx<T> z = lds[idx];
      const T sa = miss_bit(a.miss, (int64_t)la * N1 + n) ? (T)0 : a.scale;
      const T sb = (lb < a.nlines && !miss_bit(a.miss, (int64_t)lb * N1 + n)) ? a.scale : (T)0;
      z.x *= sa;
      z.y *= sb;
      if (MODE == MODE_C2R) {
        a.real_out[(int64_t)la * N1 + n] = z.x;
        if (lb < a.nlines) a.real_out[(int64_t)lb * N1 + n] = z.y;
      } else {
        lds[idx] = z;
      }
    }
    if (MODE == MODE_C2R) return;
    __syncthreads();
  } else {
    const int totr = npl * N1;
    for (int idx = threadIdx.x; idx < totr; idx += NT) {
      const int p = idx / N1;
      const int n = idx - p * N1;
      const int la = 2 * (p0 + p), lb = la + 1;
      const T xa = a.real_in[(int64_t)la * N1 + n] * a.scale;
      const T xb = lb < a.nlines ? a.real_in[(int64_t)lb * N1 + n] * a.scale : (T)0;
      lds[idx] = {xa, xb};
    }
    __syncthreads();
  }
  lds_fft<T, E, true>(lds, npl, N1, a.plan, a.tw);
  double acc = 0.0;
  const int tot = npl * H1;
  const T half = (T)0.5;
  for (int idx = threadIdx.x; idx < tot; idx += NT) {
    const int p = idx / H1;
    const int k = idx - p * H1;
    const int la = 2 * (p0 + p), lb = la + 1;
    const cpx<T>* z = lds + p * N1;
    const cpx<T> X = z[k];
    const cpx<T> Y = z[k == 0 ? 0 : N1 - k];
    const cpx<T> Va = {half * (X.x + Y.x), half * (X.y - Y.y)};
    const cpx<T> Vb = {half * (X.y + Y.y), half * (Y.x - X.x)};
    if (DST == DST_BUF) {
      a.buf[(int64_t)la * H1 + k] = Va;
      if (lb < a.nlines) a.buf[(int64_t)lb * H1 + k] = Vb;
    } else {
      acc += scatter_elem<D, T>(a.epi, k, 0, 0, Va, a.g);   // 1D: line 0 only
    }
  }
  if (DST == DST_BETA && a.epi.dot_part) store_partial(a.epi.dot_part, acc);
}

// ===========================================================================
// Vector kernels: preconditioner and CG
// ===========================================================================

// fft_kkt.jl:297-305.  The reference forms Minv = 1/(1+L1),
// S = 1/(L1 - L2^2 Minv), P12 = -L2 S Minv, P11 = Minv + Minv L2 S L2 Minv.  With
// det = (1+L1) L1 - L2^2 = L1 + 4 s1 s2 these are exactly P11 = L1/det,
// P12 = -L2/det, P22 = (1+L1)/det, evaluated here without the cancellation of
// L1 - L2^2 Minv when s1 and s2 differ by many orders of magnitude.
template <typename T>
__device__ __forceinline__ void precond_entries(T a, T b, T& l1, T& l2, T& p11, T& p12, T& p22) {
  l1 = a + b;
  l2 = a - b;
  const T inv = (T)1 / (l1 + (T)4 * a * b);
  p11 = l1 * inv;
  p12 = -l2 * inv;
  p22 = ((T)1 + l1) * inv;
}

template <typename T>
__global__ __launch_bounds__(256) void k_precond_build(const T* __restrict__ s1,
                                                       const T* __restrict__ s2, T* L1, T* L2,
                                                       T* P11, T* P12, T* P22, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T l1, l2, p11, p12, p22;
  precond_entries<T>(s1[i], s2[i], l1, l2, p11, p12, p22);
  L1[i] = l1;
  L2[i] = l2;
  P11[i] = p11;
  P12[i] = p12;
  P22[i] = p22;
}

// fft_kkt.jl:81-82
template <typename T>
__global__ __launch_bounds__(256) void k_precond_apply(const T* __restrict__ x, T* y,
                                                       const T* __restrict__ P11,
                                                       const T* __restrict__ P12,
                                                       const T* __restrict__ P22, T alpha,
                                                       T beta, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T xb = x[i], xz = x[n + i];
  T yb = alpha * (P11[i] * xb + P12[i] * xz);
  T yz = alpha * (P12[i] * xb + P22[i] * xz);
  if (beta != (T)0) {
    yb += beta * y[i];
    yz += beta * y[n + i];
  }
  y[i] = yb;
  y[n + i] = yz;
}

// CG control block, device resident.  status: 0 solved, 1 itmax, 2 zero curvature
struct CgState {
  double gamma;
  double pq;
  double alpha;
  double beta;
  double rnorm;
  double eps;
  double atol, rtol;
  int64_t iters;
  int64_t itmax;
  int done;
  int status;
};

// finalize a dot product from per-block partials (one block)
__device__ __forceinline__ double sum_partials(const double* part, int nparts) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  return block_reduce(s, OpSum());
}

// x = 0, r = b, z = P r, p = z; partial r'z
template <typename T>
__global__ __launch_bounds__(256) void k_cg_init(const T* __restrict__ b, T* x, T* r, T* z,
                                                 T* p, const T* __restrict__ P11,
                                                 const T* __restrict__ P12,
                                                 const T* __restrict__ P22, int64_t n,
                                                 double* part) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T rb = b[i], rz = b[n + i];
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    x[i] = 0;
    x[n + i] = 0;
    r[i] = rb;
    r[n + i] = rz;
    z[i] = zb;
    z[n + i] = zz;
    p[i] = zb;
    p[n + i] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part, acc);
}

__global__ void k_cg_start(CgState* st, const double* part, int nparts) {
  const double g = sum_partials(part, nparts);
  if (threadIdx.x == 0) {
    st->gamma = g;
    st->rnorm = sqrt(g > 0 ? g : 0.0);
    st->eps = st->atol + st->rtol * st->rnorm;
    st->iters = 0;
    st->status = 1;
    st->done = 0;
    if (st->rnorm <= st->eps) {
      st->done = 1;
      st->status = 0;
    }
  }
}

// after q = K p: alpha = gamma / p'q (zero curvature stops)
__global__ void k_cg_alpha(CgState* st, const double* part, int nparts) {
  if (st->done) return;
  const double pq = sum_partials(part, nparts);
  if (threadIdx.x == 0) {
    st->pq = pq;
    if (!(pq > 0.0)) {
      st->done = 1;
      st->status = 2;
    } else {
      st->alpha = st->gamma / pq;
    }
  }
}

// x += alpha p; r -= alpha q; z = P r; partial r'z
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(T* x, T* r, T* z, const T* __restrict__ p,
                                                   const T* __restrict__ q,
                                                   const T* __restrict__ P11,
                                                   const T* __restrict__ P12,
                                                   const T* __restrict__ P22, int64_t n,
                                                   const CgState* st, double* part) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    x[n + i] += alpha * p[n + i];
    const T rb = r[i] - alpha * q[i];
    const T rz = r[n + i] - alpha * q[n + i];
    r[i] = rb;
    r[n + i] = rz;
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    z[i] = zb;
    z[n + i] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part, acc);
}

__global__ void k_cg_gamma(CgState* st, const double* part, int nparts) {
  if (st->done) return;
  const double g = sum_partials(part, nparts);
  if (threadIdx.x == 0) {
    st->iters += 1;
    st->rnorm = sqrt(g > 0 ? g : 0.0);
    st->beta = g / st->gamma;
    st->gamma = g;
    if (st->rnorm <= st->eps) {
      st->done = 1;
      st->status = 0;
    } else if (st->iters >= st->itmax) {
      st->done = 1;
      st->status = 1;
    }
  }
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_cg_pupdate(T* p, const T* __restrict__ z, int64_t n2,
                                                    const CgState* st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}


// ===========================================================================
// Interior-point driver kernels (FP64 control).  State: beta, c, duals z1, z2
// of the slacks t1 = c + beta >= 0, t2 = c - beta >= 0 of FFTNLPModel's
// constraints -beta - c <= 0, beta - c <= 0 (src/fft_model.jl:92-101).
// ===========================================================================
struct RedSpec {
  int nq;
  int op[8];  // 0 sum, 1 max, 2 min
};

__device__ __forceinline__ double red_ident(int op) {
  return op == 0 ? 0.0 : (op == 1 ? -1e300 : 1e300);
}

// per-block partials of up to 8 quantities: part[q * gridDim.x + block]
template <int NQ>
__device__ __forceinline__ void store_partials(double* part, const double (&acc)[NQ],
                                               const int (&ops)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double r;
    if (ops[q] == 0) r = block_reduce(acc[q], OpSum());
    else if (ops[q] == 1) r = block_reduce(acc[q], OpMax());
    else r = block_reduce(acc[q], OpMin());
    if (threadIdx.x == 0) part[q * gridDim.x + blockIdx.x] = r;
  }
}

// ipm scalar slots
enum {
  S_RB = 0, S_RC, S_TZ, S_BQB, S_BB, S_SUMC,  // residual pass
  S_MU, S_AP_AFF, S_AD_AFF, S_SIGMA, S_AP, S_AD, S_BNORM, S_NSLOTS = 16
};

// mode 0: out[q] = reductions (+ mu from S_TZ); 1: affine step lengths;
// 2: sigma = min(1, (mu_aff/mu)^3); 3: corrector step lengths with tau
__global__ void k_ipm_finalize(const double* part, int nparts, RedSpec spec, double* scal,
                               int mode, int first, double inv2n, double tau) {
  for (int q = 0; q < spec.nq; ++q) {
    const int op = spec.op[q];
    double acc = red_ident(op);
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
      const double v = part[q * nparts + i];
      acc = op == 0 ? acc + v : (op == 1 ? fmax(acc, v) : fmin(acc, v));
    }
    double r;
    if (op == 0) r = block_reduce(acc, OpSum());
    else if (op == 1) r = block_reduce(acc, OpMax());
    else r = block_reduce(acc, OpMin());
    if (threadIdx.x == 0) {
      if (mode == 0) scal[first + q] = r;
      else if (mode == 1) scal[S_AP_AFF + q] = fmin(1.0, r);
      else if (mode == 2) {
        const double mu_aff = r * inv2n;
        const double ratio = mu_aff / scal[S_MU];
        scal[S_SIGMA] = fmin(1.0, ratio * ratio * ratio);
      } else {
        scal[S_AP + q] = fmin(1.0, tau * r);
      }
    }
    __syncthreads();
  }
  if (mode == 0 && threadIdx.x == 0 && first == S_RB) scal[S_MU] = scal[S_TZ] * inv2n;
}

__global__ __launch_bounds__(256) void k_ipm_init(double* beta, double* c, double* z1, double* z2,
                                                  int64_t n, const double* scal, double lam) {
  const double cval = 1.0 + scal[S_BNORM];
  const double zval = lam > 0 ? 0.5 * lam : 1.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    beta[i] = 0.0;
    c[i] = cval;
    z1[i] = zval;
    z2[i] = zval;
  }
}

__global__ __launch_bounds__(256) void k_absmax(const double* __restrict__ x, int64_t n,
                                                double* part) {
  double a[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[0] = fmax(a[0], fabs(x[i]));
  const int ops[1] = {1};
  store_partials<1>(part, a, ops);
}

// residuals and objective pieces: max|r_b|, max|r_c|, sum t'z, beta'Qbeta, beta'b, sum c
__global__ __launch_bounds__(256) void k_ipm_resid(const double* __restrict__ beta,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ z1,
                                                   const double* __restrict__ z2,
                                                   const double* __restrict__ Qb,
                                                   const double* __restrict__ b, double lam,
                                                   int64_t n, double* part) {
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
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
  store_partials<6>(part, a, ops);
}

// slack diagonals, preconditioner, and the predictor right-hand side
// [-r_b + g1 - g2; -r_c + g1 + g2] with g_i = -z_i
__global__ __launch_bounds__(256) void k_ipm_prep(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ Qb, const double* __restrict__ b,
    double lam, int64_t n, double* s1o, double* s2o, double* L1, double* L2, double* P11,
    double* P12, double* P22, double* rhs) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double s1 = y1 / t1, s2 = y2 / t2;
    double l1, l2, p11, p12, p22;
    precond_entries<double>(s1, s2, l1, l2, p11, p12, p22);
    s1o[i] = s1;
    s2o[i] = s2;
    L1[i] = l1;
    L2[i] = l2;
    P11[i] = p11;
    P12[i] = p12;
    P22[i] = p22;
    const double rb = Qb[i] - b[i] - y1 + y2;
    const double rc = lam - y1 - y2;
    rhs[i] = -rb - y1 + y2;
    rhs[n + i] = -rc - y1 - y2;
  }
}

// affine direction: dt, dz and the largest steps to the boundary
__global__ __launch_bounds__(256) void k_ipm_dir_aff(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ s1, const double* __restrict__ s2,
    const double* __restrict__ d, int64_t n, double* dt1o, double* dt2o, double* dz1o,
    double* dz2o, double* part) {
  double a[2] = {1e300, 1e300};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double db = d[i], dc = d[n + i];
    const double dt1 = dc + db, dt2 = dc - db;
    const double dz1 = -y1 - s1[i] * dt1;
    const double dz2 = -y2 - s2[i] * dt2;
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
  store_partials<2>(part, a, ops);
}

// complementarity after the affine step
__global__ __launch_bounds__(256) void k_ipm_muaff(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal, int64_t n,
    double* part) {
  const double ap = scal[S_AP_AFF], ad = scal[S_AD_AFF];
  double a[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double be = beta[i], ce = c[i];
    const double t1 = ce + be, t2 = ce - be;
    a[0] += (t1 + ap * dt1[i]) * (z1[i] + ad * dz1[i]) + (t2 + ap * dt2[i]) * (z2[i] + ad * dz2[i]);
  }
  const int ops[1] = {0};
  store_partials<1>(part, a, ops);
}

// corrector right-hand side: g_i = (sigma mu - t_i z_i - dt_i^aff dz_i^aff) / t_i
__global__ __launch_bounds__(256) void k_ipm_rhs_corr(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ Qb, const double* __restrict__ b,
    const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal,
    double lam, int64_t n, double* rhs) {
  const double smu = scal[S_SIGMA] * scal[S_MU];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double g1 = (smu - t1 * y1 - dt1[i] * dz1[i]) / t1;
    const double g2 = (smu - t2 * y2 - dt2[i] * dz2[i]) / t2;
    const double rb = Qb[i] - b[i] - y1 + y2;
    const double rc = lam - y1 - y2;
    rhs[i] = -rb + g1 - g2;
    rhs[n + i] = -rc + g1 + g2;
  }
}

// corrector direction (overwrites the affine dt, dz) and boundary steps
__global__ __launch_bounds__(256) void k_ipm_dir(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ s1, const double* __restrict__ s2,
    const double* __restrict__ d, const double* scal, int64_t n, double* dt1o, double* dt2o,
    double* dz1o, double* dz2o, double* part) {
  const double smu = scal[S_SIGMA] * scal[S_MU];
  double a[2] = {1e300, 1e300};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double g1 = (smu - t1 * y1 - dt1o[i] * dz1o[i]) / t1;
    const double g2 = (smu - t2 * y2 - dt2o[i] * dz2o[i]) / t2;
    const double db = d[i], dc = d[n + i];
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
  store_partials<2>(part, a, ops);
}

__global__ __launch_bounds__(256) void k_ipm_update(double* beta, double* c, double* z1,
                                                    double* z2, const double* __restrict__ d,
                                                    const double* __restrict__ dz1,
                                                    const double* __restrict__ dz2,
                                                    const double* scal, int64_t n) {
  const double ap = scal[S_AP], ad = scal[S_AD];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    beta[i] += ap * d[i];
    c[i] += ap * d[n + i];
    z1[i] += ad * dz1[i];
    z2[i] += ad * dz2[i];
  }
}


// ===========================================================================
// FFTNLPModel and the full KKT operators (elementwise, op precision T)
// ===========================================================================

// sum y^2, sum beta * Mtz0, sum c  (fft_model.jl:73)
template <typename T>
__global__ __launch_bounds__(256) void k_obj_parts(const T* __restrict__ y,
                                                   const T* __restrict__ x,
                                                   const T* __restrict__ mtz0, int64_t n,
                                                   double* part) {
  double a[3] = {0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double yi = y[i];
    a[0] += yi * yi;
    a[1] += (double)x[i] * (double)mtz0[i];
    a[2] += (double)x[n + i];
  }
  const int ops[3] = {0, 0, 0};
  store_partials<3>(part, a, ops);
}

// g_b = MtM beta - Mtz0 (in place on g_b), g_c = lambda  (fft_model.jl:77-89)
template <typename T>
__global__ __launch_bounds__(256) void k_grad_tail(T* g, const T* __restrict__ mtz0, T lam,
                                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    g[i] = g[i] - mtz0[i];
    g[n + i] = lam;
  }
}

// c_b = -beta - c, c_c = beta - c  (fft_model.jl:91-101)
template <typename T>
__global__ __launch_bounds__(256) void k_cons(const T* __restrict__ x, T* cv, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T b = x[i], c = x[n + i];
    cv[i] = -b - c;
    cv[n + i] = b - c;
  }
}

// full KKT product after M^T M x_beta is in Mb (fft_kkt.jl:239-246)
template <typename T>
__global__ __launch_bounds__(256) void k_kkt_full(const T* __restrict__ x, T* y,
                                                  const T* __restrict__ Mb, T alpha, T beta,
                                                  int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T xb = x[i], xz = x[n + i], xs1 = x[2 * n + i], xs2 = x[3 * n + i];
    const T xy1 = x[4 * n + i], xy2 = x[5 * n + i];
    T o0 = alpha * (Mb[i] - xy1 + xy2);
    T o1 = -alpha * (xy1 + xy2);
    T o2 = -alpha * xy1;
    T o3 = -alpha * xy2;
    T o4 = alpha * (-xb - xz - xs1);
    T o5 = alpha * (xb - xz - xs2);
    if (beta != (T)0) {
      o0 += beta * y[i];
      o1 += beta * y[n + i];
      o2 += beta * y[2 * n + i];
      o3 += beta * y[3 * n + i];
      o4 += beta * y[4 * n + i];
      o5 += beta * y[5 * n + i];
    }
    y[i] = o0;
    y[n + i] = o1;
    y[2 * n + i] = o2;
    y[3 * n + i] = o3;
    y[4 * n + i] = o4;
    y[5 * n + i] = o5;
  }
}

// J^T y (fft_kkt.jl:253-275)
template <typename T>
__global__ __launch_bounds__(256) void k_jtprod(const T* __restrict__ x, T* y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T y1 = x[i], y2 = x[n + i];
    y[i] = -y1 + y2;
    y[n + i] = -y1 - y2;
    y[2 * n + i] = -y1;
    y[3 * n + i] = -y2;
  }
}

// 6N KKT right-hand side -> condensed 2N right-hand side (fft_kkt.jl:333-335)
template <typename T>
__global__ __launch_bounds__(256) void k_condense(const T* __restrict__ w,
                                                  const T* __restrict__ s1,
                                                  const T* __restrict__ s2, T* b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T w1 = w[i], w2 = w[n + i], w3 = w[2 * n + i], w4 = w[3 * n + i];
    const T w5 = w[4 * n + i], w6 = w[5 * n + i], a = s1[i], c = s2[i];
    b[i] = w1 - w3 + w4 - a * w5 + c * w6;
    b[n + i] = w2 - w3 - w4 - a * w5 - c * w6;
  }
}

// condensed solution -> 6N solution, in place on w (fft_kkt.jl:344-350)
template <typename T>
__global__ __launch_bounds__(256) void k_unpack(T* w, const T* __restrict__ x,
                                                const T* __restrict__ s1,
                                                const T* __restrict__ s2, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T a = s1[i], c = s2[i];
    const T w1 = x[i], w2 = x[n + i];
    T w3 = w[2 * n + i], w4 = w[3 * n + i], w5 = w[4 * n + i], w6 = w[5 * n + i];
    w5 = -(w3 + a * (w1 + w2 + w5));
    w6 = -(w4 + c * (-w1 + w2 + w6));
    w3 = (w3 + w5) / a;
    w4 = (w4 + w6) / c;
    w[i] = w1;
    w[n + i] = w2;
    w[2 * n + i] = w3;
    w[3 * n + i] = w4;
    w[4 * n + i] = w5;
    w[5 * n + i] = w6;
  }
}

}  // namespace csipm
