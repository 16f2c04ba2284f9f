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
// Destinations of the forward passes that end in beta (or u) space
// ===========================================================================
template <typename T>
struct Epi {
  int kind;             // 0: out[i] = value          1: condensed KKT epilogue
                        // 2: reduced (Schur) epilogue y = alpha (m + D x) + beta y, D in L1
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
  if (e.kind == 2) {
    const T xb = e.xb[i];
    T qb = e.alpha * (m + e.L1[i] * xb);
    if (e.beta != (T)0) qb += e.beta * e.yb[i];
    e.yb[i] = qb;
    return (double)xb * (double)qb;
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

// the same on the two slots (re, im) of u-space element e, as one vector access
template <typename T>
__device__ __forceinline__ double epi_apply_u(const Epi<T>& e, int64_t el, cpx<T> m) {
  if (e.kind == 0) {
    reinterpret_cast<cpx<T>*>(e.out)[el] = m;
    return 0.0;
  }
  if (e.kind == 2) {
    const cpx<T> xb = reinterpret_cast<const cpx<T>*>(e.xb)[el];
    const cpx<T> dg = reinterpret_cast<const cpx<T>*>(e.L1)[el];
    cpx<T> qb = {e.alpha * (m.x + dg.x * xb.x), e.alpha * (m.y + dg.y * xb.y)};
    cpx<T>* yb = reinterpret_cast<cpx<T>*>(e.yb);
    if (e.beta != (T)0) {
      const cpx<T> ob = yb[el];
      qb.x += e.beta * ob.x;
      qb.y += e.beta * ob.y;
    }
    yb[el] = qb;
    return (double)xb.x * qb.x + (double)xb.y * qb.y;
  }
  const cpx<T> xb = reinterpret_cast<const cpx<T>*>(e.xb)[el];
  const cpx<T> xz = reinterpret_cast<const cpx<T>*>(e.xz)[el];
  const cpx<T> l1 = reinterpret_cast<const cpx<T>*>(e.L1)[el];
  const cpx<T> l2 = reinterpret_cast<const cpx<T>*>(e.L2)[el];
  cpx<T> qb = {e.alpha * (m.x + l1.x * xb.x + l2.x * xz.x),
               e.alpha * (m.y + l1.y * xb.y + l2.y * xz.y)};
  cpx<T> qz = {e.alpha * (l2.x * xb.x + l1.x * xz.x), e.alpha * (l2.y * xb.y + l1.y * xz.y)};
  cpx<T>* yb = reinterpret_cast<cpx<T>*>(e.yb);
  cpx<T>* yz = reinterpret_cast<cpx<T>*>(e.yz);
  if (e.beta != (T)0) {
    const cpx<T> ob = yb[el], oz = yz[el];
    qb.x += e.beta * ob.x;
    qb.y += e.beta * ob.y;
    qz.x += e.beta * oz.x;
    qz.y += e.beta * oz.y;
  }
  yb[el] = qb;
  yz[el] = qz;
  return (double)xb.x * qb.x + (double)xb.y * qb.y + (double)xz.x * qz.x + (double)xz.y * qz.y;
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
// FFT passes over the (H1, N2, N3) half spectrum.
//  k_lines: lines along dimension 2 (axis 1) or 3 (axis 2); a workgroup takes
//           2^lg lines with consecutive k1, so global accesses are contiguous.
//  k_dim1:  dimension 1; two real lines per complex FFT of length N1
//           (z = x_a + i x_b), 2^lg pairs per workgroup.
// Sources: SRC_BUF (scratch spectrum), SRC_BETA (gather from beta through the
// packing map), SRC_U (spectrum-ordered beta), SRC_REAL (real samples, dim 1).
// Destinations likewise, the beta / u ones with the epilogue Epi.
// ===========================================================================
enum { SRC_BUF = 0, SRC_BETA = 1, SRC_REAL = 2, SRC_U = 3 };
enum { DST_BUF = 0, DST_BETA = 1, DST_REAL = 2, DST_U = 3 };
enum { MODE_MTM = 0, MODE_C2R = 1, MODE_R2C = 2 };

template <typename T>
struct PassArgs {
  MapGeom g;
  FftPlan plan;
  const cpx<T>* tw;
  int axis;      // k_lines: 1 or 2
  int H1;
  int lg;        // log2 of lines (pairs) per workgroup
  int ps;        // LDS line stride
  int ntile1;    // k_lines: tiles along k1
  int nlines;    // k_dim1: real lines N2*N3
  int npairs;
  uint32_t h1m;  // ceil(2^32 / H1)
  uint32_t n1m;  // ceil(2^32 / N1)
  cpx<T>* buf;
  const T* beta_in;
  const cpx<T>* u_in;
  const T* real_in;
  T* real_out;
  Epi<T> epi;
  const uint32_t* miss;  // bit set = missing sample (nullable)
  T scale;
  const int* done;
};

__device__ __forceinline__ int udiv(int a, int d, uint32_t m) { return d == 1 ? a : fdiv(a, m); }

template <int D, typename T, int SRC>
__device__ __forceinline__ cpx<T> spec_load(const PassArgs<T>& a, int k1, int k2, int k3) {
  if (SRC == SRC_BETA) return beta_to_dft_elem<D, T>(a.beta_in, k1, k2, k3, a.g);
  return u_to_v<D, T>(a.u_in, k1, k2, k3, a.g, a.H1);
}

template <int D, typename T, int DST>
__device__ __forceinline__ double spec_store(const PassArgs<T>& a, int k1, int k2, int k3,
                                             cpx<T> v) {
  if (DST == DST_BETA) return scatter_elem<D, T>(a.epi, k1, k2, k3, v, a.g);
  const int64_t el = k1 + (int64_t)a.H1 * (k2 + (int64_t)a.g.n[1] * k3);
  return epi_apply_u(a.epi, el, v_to_u<D, T>(v, k1, k2, k3, a.g));
}

template <typename T, int NT, int E, int D, int SRC, int DST, bool FWD>
__global__ __launch_bounds__(NT) void k_lines(PassArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const int L = a.plan.L;
  const int lg = a.lg, nl = 1 << lg;
  const int tile1 = blockIdx.x % a.ntile1;
  const int o = blockIdx.x / a.ntile1;
  const int k10 = tile1 << lg;
  const int nv = min(nl, a.H1 - k10);
  const int64_t s2 = a.H1, s3 = (int64_t)a.H1 * a.g.n[1];
  const int64_t sl = a.axis == 1 ? s2 : s3;
  const int64_t base = k10 + (a.axis == 1 ? (int64_t)o * s3 : (int64_t)o * s2);
  const int total = L << lg;
  for (int idx = threadIdx.x; idx < total; idx += NT) {
    const int line = idx & (nl - 1);
    const int el = idx >> lg;
    cpx<T> val = {(T)0, (T)0};
    if (line < nv) {
      if (SRC == SRC_BUF) {
        val = a.buf[base + line + el * sl];
      } else {
        const int k2 = a.axis == 1 ? el : o;
        const int k3 = a.axis == 1 ? o : el;
        val = spec_load<D, T, SRC>(a, k10 + line, k2, k3);
      }
    }
    lds[idx] = val;
  }
  __syncthreads();
  lds_fft<T, E, NT, FWD>(lds, a.plan, lg, nl, a.tw);
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < total; idx += NT) {
    const int line = idx & (nl - 1);
    const int el = idx >> lg;
    if (line >= nv) continue;
    const cpx<T> val = lds[idx];
    if (DST == DST_BUF) {
      a.buf[base + line + el * sl] = val;
    } else {
      const int k2 = a.axis == 1 ? el : o;
      const int k3 = a.axis == 1 ? o : el;
      acc += spec_store<D, T, DST>(a, k10 + line, k2, k3, val);
    }
  }
  if (DST != DST_BUF && a.epi.dot_part) store_partial(a.epi.dot_part, acc);
}

__device__ __forceinline__ bool miss_bit(const uint32_t* m, int64_t s) {
  return m && ((m[s >> 5] >> (s & 31)) & 1u);
}

template <typename T, int NT, int E, int D, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void k_dim1(PassArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const int N1 = a.plan.L, H1 = a.H1;
  const int lg = a.lg, np = 1 << lg, ps = a.ps;
  const int p0 = blockIdx.x << lg;
  const bool even = (N1 & 1) == 0;
  if (MODE != MODE_R2C) {
    const int tot = H1 << lg;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      const int p = udiv(idx, H1, a.h1m);
      const int k = idx - p * H1;
      const int la = 2 * (p0 + p), lb = la + 1;
      cpx<T> A = {(T)0, (T)0}, B = {(T)0, (T)0};
      if (la < a.nlines) {
        if (SRC == SRC_BUF) {
          A = a.buf[(int64_t)la * H1 + k];
          if (lb < a.nlines) B = a.buf[(int64_t)lb * H1 + k];
        } else {
          A = spec_load<D, T, SRC>(a, k, 0, 0);   // 1D: a single line
        }
      }
      if (k == 0 || (even && k == H1 - 1)) {  // DC / Nyquist of a real line are real
        A.y = 0;
        B.y = 0;
      }
      lds[k * ps + p] = {A.x - B.y, A.y + B.x};
      if (k > 0 && N1 - k != k) lds[(N1 - k) * ps + p] = {A.x + B.y, B.x - A.y};
    }
    __syncthreads();
    lds_fft<T, E, NT, false>(lds, a.plan, lg, ps, a.tw);
    const int totr = N1 << lg;
    for (int idx = threadIdx.x; idx < totr; idx += NT) {
      const int p = idx & (np - 1);
      const int n = idx >> lg;
      const int la = 2 * (p0 + p), lb = la + 1;
      cpx<T> z = lds[n * ps + p];
      const T sa = (la < a.nlines && !miss_bit(a.miss, (int64_t)la * N1 + n)) ? a.scale : (T)0;
      const T sb = (lb < a.nlines && !miss_bit(a.miss, (int64_t)lb * N1 + n)) ? a.scale : (T)0;
      z.x *= sa;
      z.y *= sb;
      if (MODE == MODE_C2R) {
        if (la < a.nlines) a.real_out[(int64_t)la * N1 + n] = z.x;
        if (lb < a.nlines) a.real_out[(int64_t)lb * N1 + n] = z.y;
      } else {
        lds[n * ps + p] = z;
      }
    }
    if (MODE == MODE_C2R) return;
    __syncthreads();
  } else {
    const int totr = N1 << lg;
    for (int idx = threadIdx.x; idx < totr; idx += NT) {
      const int p = udiv(idx, N1, a.n1m);
      const int n = idx - p * N1;
      const int la = 2 * (p0 + p), lb = la + 1;
      const T xa = la < a.nlines ? a.real_in[(int64_t)la * N1 + n] * a.scale : (T)0;
      const T xb = lb < a.nlines ? a.real_in[(int64_t)lb * N1 + n] * a.scale : (T)0;
      lds[n * ps + p] = {xa, xb};
    }
    __syncthreads();
  }
  lds_fft<T, E, NT, true>(lds, a.plan, lg, ps, a.tw);
  double acc = 0.0;
  const int tot = H1 << lg;
  const T half = (T)0.5;
  for (int idx = threadIdx.x; idx < tot; idx += NT) {
    const int p = udiv(idx, H1, a.h1m);
    const int k = idx - p * H1;
    const int la = 2 * (p0 + p), lb = la + 1;
    if (la >= a.nlines) continue;
    const cpx<T> X = lds[k * ps + p];
    const cpx<T> Y = lds[(k == 0 ? 0 : N1 - k) * ps + p];
    const cpx<T> Va = {half * (X.x + Y.x), half * (X.y - Y.y)};
    const cpx<T> Vb = {half * (X.y + Y.y), half * (Y.x - X.x)};
    if (DST == DST_BUF) {
      a.buf[(int64_t)la * H1 + k] = Va;
      if (lb < a.nlines) a.buf[(int64_t)lb * H1 + k] = Vb;
    } else {
      acc += spec_store<D, T, DST>(a, k, 0, 0, Va);   // 1D: line 0 only
    }
  }
  if (DST != DST_BUF && a.epi.dot_part) store_partial(a.epi.dot_part, acc);
}

// beta <-> u (exact: a signed permutation), one thread per half-spectrum element.
// sign = 0 moves a diagonal (L1, L2, P..) instead: the same permutation, no signs.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_beta_to_u(const T* __restrict__ beta, cpx<T>* u,
                                                   MapGeom g, int H1, int64_t nc, int sign) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  const int k1 = (int)(e % H1);
  const int64_t r = e / H1;
  const int k2 = (int)(r % g.n[1]);
  const int k3 = (int)(r / g.n[1]);
  int64_t re, im;
  bool vertex, neg;
  cpx<T> w = {(T)0, (T)0};
  if (scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) {
    w.x = beta[re];
    if (!vertex) w.y = (neg && sign) ? -beta[im] : beta[im];
  }
  u[e] = w;
}

template <int D, typename T>
__global__ __launch_bounds__(256) void k_u_to_beta(const cpx<T>* __restrict__ u, T* beta,
                                                   MapGeom g, int H1, int64_t nc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  const int k1 = (int)(e % H1);
  const int64_t r = e / H1;
  const int k2 = (int)(r % g.n[1]);
  const int k3 = (int)(r / g.n[1]);
  int64_t re, im;
  bool vertex, neg;
  if (scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) {
    const cpx<T> w = u[e];
    beta[re] = w.x;
    if (!vertex) beta[im] = neg ? -w.y : w.y;
  }
}

// bit i set when u slot i holds a beta entry (32 slots = 16 elements per thread)
template <int D>
__global__ __launch_bounds__(256) void k_u_valid(uint32_t* bits, MapGeom g, int H1, int64_t nc) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w * 16 >= nc) return;
  uint32_t word = 0;
  for (int j = 0; j < 16; ++j) {
    const int64_t e = w * 16 + j;
    if (e >= nc) break;
    const int k1 = (int)(e % H1);
    const int64_t r = e / H1;
    const int k2 = (int)(r % g.n[1]);
    const int k3 = (int)(r / g.n[1]);
    int64_t re, im;
    bool vertex, neg;
    if (scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) {
      word |= 1u << (2 * j);
      if (!vertex) word |= 1u << (2 * j + 1);
    }
  }
  bits[w] = word;
}

__device__ __forceinline__ bool slot_valid(const uint32_t* v, int64_t i) {
  return (v[i >> 5] >> (i & 31)) & 1u;
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

// The CG vectors have nb blocks of n: nb = 2 for the condensed system (the
// block preconditioner [[P11, P12], [P12, P22]]), nb = 1 for the reduced
// system (diagonal preconditioner P11; P12, P22 unused).
// x = 0, r = b, z = P r, p = z; partial r'z
template <typename T>
__global__ __launch_bounds__(256) void k_cg_init(const T* __restrict__ b, T* x, T* r, T* z,
                                                 T* p, const T* __restrict__ P11,
                                                 const T* __restrict__ P12,
                                                 const T* __restrict__ P22, int64_t n, int nb,
                                                 double* part) {
  double acc = 0.0;
  const bool two = nb == 2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T rb = b[i];
    if (!two) {
      const T zb = P11[i] * rb;
      x[i] = 0;
      r[i] = rb;
      z[i] = zb;
      p[i] = zb;
      acc += (double)rb * zb;
      continue;
    }
    const T rz = b[n + i];
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
                                                   int nb, const CgState* st, double* part) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double acc = 0.0;
  const bool two = nb == 2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!two) {
      x[i] += alpha * p[i];
      const T rb = r[i] - alpha * q[i];
      const T zb = P11[i] * rb;
      r[i] = rb;
      z[i] = zb;
      acc += (double)rb * zb;
      continue;
    }
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

// reliable update: r = b - q with q = K x just recomputed, z = P r, partial r'z
template <typename T>
__global__ __launch_bounds__(256) void k_cg_residual(const T* __restrict__ b,
                                                     const T* __restrict__ q, T* r, T* z,
                                                     const T* __restrict__ P11,
                                                     const T* __restrict__ P12,
                                                     const T* __restrict__ P22, int64_t n,
                                                     int nb, const CgState* st, double* part) {
  if (st->done) return;
  double acc = 0.0;
  const bool two = nb == 2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!two) {
      const T rb = b[i] - q[i];
      const T zb = P11[i] * rb;
      r[i] = rb;
      z[i] = zb;
      acc += (double)rb * zb;
      continue;
    }
    const T rb = b[i] - q[i], rz = b[n + i] - q[n + i];
    const T zb = P11[i] * rb + P12[i] * rz;
    const T zz = P12[i] * rb + P22[i] * rz;
    r[i] = rb;
    r[n + i] = rz;
    z[i] = zb;
    z[n + i] = zz;
    acc += (double)rb * zb + (double)rz * zz;
  }
  store_partial(part, acc);
}

// after a reliable update: gamma from the true residual, restart p = z
__global__ void k_cg_regamma(CgState* st, const double* part, int nparts) {
  if (st->done) return;
  const double g = sum_partials(part, nparts);
  if (threadIdx.x == 0) {
    st->rnorm = sqrt(g > 0 ? g : 0.0);
    st->gamma = g;
    st->beta = 0.0;
    if (st->rnorm <= st->eps) {
      st->done = 1;
      st->status = 0;
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

// slots with valid == null or bit set hold a beta entry; the others (padding
// of the spectrum-ordered layout) stay inert: beta = 0, c = 1, z = 0
#define FOR_SLOTS(i, n)                                                          \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n);      \
       i += (int64_t)gridDim.x * blockDim.x)
#define VALID(v, i) (!(v) || slot_valid((v), (i)))

__global__ __launch_bounds__(256) void k_ipm_init(double* beta, double* c, double* z1, double* z2,
                                                  int64_t n, const double* scal, double lam,
                                                  const uint32_t* valid) {
  const double cval = 1.0 + scal[S_BNORM];
  const double zval = lam > 0 ? 0.5 * lam : 1.0;
  FOR_SLOTS(i, n) {
    const bool ok = VALID(valid, i);
    beta[i] = 0.0;
    c[i] = ok ? cval : 1.0;
    z1[i] = ok ? zval : 0.0;
    z2[i] = ok ? zval : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_absmax(const double* __restrict__ x, int64_t n,
                                                double* part) {
  double a[1] = {0.0};
  FOR_SLOTS(i, n) a[0] = fmax(a[0], fabs(x[i]));
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
                                                   int64_t n, double* part,
                                                   const uint32_t* valid) {
  double a[6] = {0, 0, 0, 0, 0, 0};
  FOR_SLOTS(i, n) {
    if (!VALID(valid, i)) continue;
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

// slack diagonals, preconditioner (in the CG precision TC) and the predictor
// right-hand side [-r_b + g1 - g2; -r_c + g1 + g2] with g_i = -z_i
template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_prep(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ Qb, const double* __restrict__ b,
    double lam, int64_t n, double* s1o, double* s2o, TC* L1, TC* L2, TC* P11, TC* P12, TC* P22,
    TC* rhs, const uint32_t* valid) {
  FOR_SLOTS(i, n) {
    double s1 = 0, s2 = 0, rb = 0, rc = 0, y1 = 0, y2 = 0;
    double l1 = 0, l2 = 0, p11 = 1, p12 = 0, p22 = 1;
    const bool ok = VALID(valid, i);
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
    rhs[i] = (TC)(ok ? -rb - y1 + y2 : 0.0);
    rhs[n + i] = (TC)(ok ? -rc - y1 - y2 : 0.0);
  }
}

// The same for the reduced system: L1 = S1 + S2, L2 = S1 - S2 (FP64),
// D = 4 S1 S2 / (S1 + S2) (FP64, and in TC with the preconditioner 1/(1 + D)),
// rhsd = [h_b - (L2/L1) h_c; h_c] (FP64) and rhs = h_b - (L2/L1) h_c in TC.
template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_prep_red(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ Qb, const double* __restrict__ b,
    double lam, int64_t n, double* s1o, double* s2o, TC* D, TC* P, TC* rhs, double* L1d,
    double* L2d, double* Dd, double* rhsd, const uint32_t* valid) {
  FOR_SLOTS(i, n) {
    double s1 = 0, s2 = 0, l1 = 0, l2 = 0, dg = 0, hb = 0, hc = 0;
    if (VALID(valid, i)) {
      const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
      const double t1 = ce + be, t2 = ce - be;
      s1 = y1 / t1;
      s2 = y2 / t2;
      l1 = s1 + s2;
      l2 = s1 - s2;
      dg = 4.0 * s1 * s2 / l1;
      const double rb = Qb[i] - b[i] - y1 + y2;
      const double rc = lam - y1 - y2;
      hb = -rb - y1 + y2;
      hc = -rc - y1 - y2;
      hb -= (l2 / l1) * hc;
    }
    s1o[i] = s1;
    s2o[i] = s2;
    L1d[i] = l1;
    L2d[i] = l2;
    Dd[i] = dg;
    D[i] = (TC)dg;
    P[i] = (TC)(1.0 / (1.0 + dg));
    rhsd[i] = hb;
    rhsd[n + i] = hc;
    rhs[i] = (TC)hb;
  }
}

// dc = (h_c - L2 db) / L1 into the second block of d (FP64)
__global__ __launch_bounds__(256) void k_red_back(double* d, const double* __restrict__ rhsd,
                                                  const double* __restrict__ L1d,
                                                  const double* __restrict__ L2d, int64_t n,
                                                  const uint32_t* valid) {
  FOR_SLOTS(i, n) d[n + i] = VALID(valid, i) ? (rhsd[n + i] - L2d[i] * d[i]) / L1d[i] : 0.0;
}

// affine direction: dt, dz and the largest steps to the boundary
template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_dir_aff(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ s1, const double* __restrict__ s2,
    const TC* __restrict__ d, int64_t n, double* dt1o, double* dt2o, double* dz1o, double* dz2o,
    double* part, const uint32_t* valid) {
  double a[2] = {1e300, 1e300};
  FOR_SLOTS(i, n) {
    if (!VALID(valid, i)) {
      dt1o[i] = dt2o[i] = dz1o[i] = dz2o[i] = 0.0;
      continue;
    }
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double db = (double)d[i], dc = (double)d[n + i];
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
    double* part, const uint32_t* valid) {
  const double ap = scal[S_AP_AFF], ad = scal[S_AD_AFF];
  double a[1] = {0.0};
  FOR_SLOTS(i, n) {
    if (!VALID(valid, i)) continue;
    const double be = beta[i], ce = c[i];
    const double t1 = ce + be, t2 = ce - be;
    a[0] += (t1 + ap * dt1[i]) * (z1[i] + ad * dz1[i]) + (t2 + ap * dt2[i]) * (z2[i] + ad * dz2[i]);
  }
  const int ops[1] = {0};
  store_partials<1>(part, a, ops);
}

// corrector right-hand side: g_i = (sigma mu - t_i z_i - dt_i^aff dz_i^aff) / t_i
template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_rhs_corr(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ Qb, const double* __restrict__ b,
    const double* __restrict__ dt1, const double* __restrict__ dt2,
    const double* __restrict__ dz1, const double* __restrict__ dz2, const double* scal,
    double lam, int64_t n, TC* rhs, double* rhsd, const double* __restrict__ L1d,
    const double* __restrict__ L2d, const uint32_t* valid) {
  const double smu = scal[S_SIGMA] * scal[S_MU];
  FOR_SLOTS(i, n) {
    double hb = 0, hc = 0;
    if (VALID(valid, i)) {
      const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
      const double t1 = ce + be, t2 = ce - be;
      const double g1 = (smu - t1 * y1 - dt1[i] * dz1[i]) / t1;
      const double g2 = (smu - t2 * y2 - dt2[i] * dz2[i]) / t2;
      const double rb = Qb[i] - b[i] - y1 + y2;
      const double rc = lam - y1 - y2;
      hb = -rb + g1 - g2;
      hc = -rc + g1 + g2;
    }
    if (rhsd) {   // reduced system (see k_ipm_prep_red)
      if (VALID(valid, i)) hb -= (L2d[i] / L1d[i]) * hc;
      rhsd[i] = hb;
      rhsd[n + i] = hc;
      rhs[i] = (TC)hb;
    } else {
      rhs[i] = (TC)hb;
      rhs[n + i] = (TC)hc;
    }
  }
}

// corrector direction (overwrites the affine dt, dz) and boundary steps
template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_dir(
    const double* __restrict__ beta, const double* __restrict__ c, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ s1, const double* __restrict__ s2,
    const TC* __restrict__ d, const double* scal, int64_t n, double* dt1o, double* dt2o,
    double* dz1o, double* dz2o, double* part, const uint32_t* valid) {
  const double smu = scal[S_SIGMA] * scal[S_MU];
  double a[2] = {1e300, 1e300};
  FOR_SLOTS(i, n) {
    if (!VALID(valid, i)) {
      dt1o[i] = dt2o[i] = dz1o[i] = dz2o[i] = 0.0;
      continue;
    }
    const double be = beta[i], ce = c[i], y1 = z1[i], y2 = z2[i];
    const double t1 = ce + be, t2 = ce - be;
    const double g1 = (smu - t1 * y1 - dt1o[i] * dz1o[i]) / t1;
    const double g2 = (smu - t2 * y2 - dt2o[i] * dz2o[i]) / t2;
    const double db = (double)d[i], dc = (double)d[n + i];
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

template <typename TC>
__global__ __launch_bounds__(256) void k_ipm_update(double* beta, double* c, double* z1,
                                                    double* z2, const TC* __restrict__ d,
                                                    const double* __restrict__ dz1,
                                                    const double* __restrict__ dz2,
                                                    const double* scal, int64_t n,
                                                    const uint32_t* valid) {
  const double ap = scal[S_AP], ad = scal[S_AD];
  FOR_SLOTS(i, n) {
    if (!VALID(valid, i)) continue;
    beta[i] += ap * (double)d[i];
    c[i] += ap * (double)d[n + i];
    z1[i] += ad * dz1[i];
    z2[i] += ad * dz2[i];
  }
}

// mixed precision helpers: x_lo = (TC) x_hi;  x_hi += (double) x_lo
template <typename TC>
__global__ __launch_bounds__(256) void k_narrow(const double* __restrict__ x, TC* y, int64_t n) {
  FOR_SLOTS(i, n) y[i] = (TC)x[i];
}
template <typename TC>
__global__ __launch_bounds__(256) void k_widen_add(double* y, const TC* __restrict__ x, int64_t n) {
  FOR_SLOTS(i, n) y[i] += (double)x[i];
}
template <typename TC>
__global__ __launch_bounds__(256) void k_widen(double* y, const TC* __restrict__ x, int64_t n) {
  FOR_SLOTS(i, n) y[i] = (double)x[i];
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
