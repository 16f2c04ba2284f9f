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
  T* out;               // kind 0: the value; kind 2: m itself as well, when not null
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
    if (e.out) e.out[i] = m;
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
    if (e.out) reinterpret_cast<cpx<T>*>(e.out)[el] = m;
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

// The epilogue of a u-space element in two steps, so that a thread can issue
// the operand loads of several elements before it consumes any of them:
// epi_fetch_u loads what kind `e.kind` reads (always in range for el < U),
// epi_commit_u forms and stores the result and returns its x'y term.
template <typename T>
struct EpiOps {
  cpx<T> xb, xz, l1, l2, ob, oz;
};

template <typename T>
__device__ __forceinline__ EpiOps<T> epi_fetch_u(const Epi<T>& e, int64_t el) {
  EpiOps<T> o;
  const cpx<T> z = {(T)0, (T)0};
  o.xb = o.xz = o.l1 = o.l2 = o.ob = o.oz = z;
  if (e.kind == 0) return o;
  o.xb = reinterpret_cast<const cpx<T>*>(e.xb)[el];
  o.l1 = reinterpret_cast<const cpx<T>*>(e.L1)[el];
  if (e.beta != (T)0) o.ob = reinterpret_cast<const cpx<T>*>(e.yb)[el];
  if (e.kind == 2) return o;
  o.xz = reinterpret_cast<const cpx<T>*>(e.xz)[el];
  o.l2 = reinterpret_cast<const cpx<T>*>(e.L2)[el];
  if (e.beta != (T)0) o.oz = reinterpret_cast<const cpx<T>*>(e.yz)[el];
  return o;
}

// the arithmetic of epi_apply_u, on fetched operands
template <typename T>
__device__ __forceinline__ double epi_commit_u(const Epi<T>& e, int64_t el, cpx<T> m,
                                               const EpiOps<T>& o) {
  if (e.kind == 0) {
    reinterpret_cast<cpx<T>*>(e.out)[el] = m;
    return 0.0;
  }
  if (e.kind == 2) {
    if (e.out) reinterpret_cast<cpx<T>*>(e.out)[el] = m;
    cpx<T> qb = {e.alpha * (m.x + o.l1.x * o.xb.x), e.alpha * (m.y + o.l1.y * o.xb.y)};
    if (e.beta != (T)0) {
      qb.x += e.beta * o.ob.x;
      qb.y += e.beta * o.ob.y;
    }
    reinterpret_cast<cpx<T>*>(e.yb)[el] = qb;
    return (double)o.xb.x * qb.x + (double)o.xb.y * qb.y;
  }
  cpx<T> qb = {e.alpha * (m.x + o.l1.x * o.xb.x + o.l2.x * o.xz.x),
               e.alpha * (m.y + o.l1.y * o.xb.y + o.l2.y * o.xz.y)};
  cpx<T> qz = {e.alpha * (o.l2.x * o.xb.x + o.l1.x * o.xz.x),
               e.alpha * (o.l2.y * o.xb.y + o.l1.y * o.xz.y)};
  if (e.beta != (T)0) {
    qb.x += e.beta * o.ob.x;
    qb.y += e.beta * o.ob.y;
    qz.x += e.beta * o.oz.x;
    qz.y += e.beta * o.oz.y;
  }
  reinterpret_cast<cpx<T>*>(e.yb)[el] = qb;
  reinterpret_cast<cpx<T>*>(e.yz)[el] = qz;
  return (double)o.xb.x * qb.x + (double)o.xb.y * qb.y + (double)o.xz.x * qz.x +
         (double)o.xz.y * qz.y;
}

// Global-memory phases of the pass kernels run in groups of stage_k<T>()
// elements per thread: fetch(i) issues the loads of element i (branch-free,
// on an index clamped into range), put(i, r) consumes them.  All fetches of
// a group come before its first put, so a thread has that many loads in
// flight instead of one.
template <typename T>
constexpr int stage_k() {
  return sizeof(T) == 4 ? 8 : 4;
}

template <int NT, int K, typename R, typename F, typename P>
__device__ __forceinline__ void staged(int total, F&& fetch, P&& put) {
  for (int i0 = threadIdx.x; i0 < total; i0 += K * NT) {
    R r[K];
#pragma unroll
    for (int j = 0; j < K; ++j) r[j] = fetch(min(i0 + j * NT, total - 1));
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (i0 + j * NT < total) put(i0 + j * NT, r[j]);
  }
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
__device__ __forceinline__ void store_partial(double* part, double acc, int idx) {
  const double s = block_reduce(acc, OpSum());
  if (threadIdx.x == 0 && part) part[idx] = s;
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
  const FftPlan* plan;  // device-resident (read with uniform loads)
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
  // line transform: length len; bs = 1 selects Bluestein's chirp-z form, in
  // which case plan is the convolution FFT of length M >= 2 len - 1; bs = 2
  // selects Rader's form for a prime len, plan being the FFT of length len - 1
  int len;
  int bs;
  const cpx<T>* chirp;  // w_n = exp(-i pi n^2 / len), n < len
  const cpx<T>* bfw;    // FFT_M(b) / M of the forward kernel b = conj(w) (wrapped);
                        // Rader: FFT_{L-1}(b) / (L-1) of b_j = exp(-2 pi i g^-j / L)
  const cpx<T>* binv;   // the same for the inverse kernel b = w (Rader: conj b)
  const int* rperm;     // Rader: [e] row of element e in the convolution order
                        // (e = 0: row L), [L + m] element g^-m mod L of output m
  // long 1D lines (four-step): the length-F transform (F = len, or the
  // Bluestein length when obs) as A columns x B rows, position col + A*row
  int A, B, F;
  int obs;              // outer Bluestein: chirp / bfw / binv are of length len / F / F
  uint32_t am;          // ceil(2^32 / A)
  const cpx<T>* twF;    // exp(-2 pi i j / F), j < F
};

// DFT of length a.len of the 2^lg lines in LDS ([element][line], stride ps),
// in place.  Direct: the Stockham plan.  Bluestein (lengths with a prime
// factor above 17): X_k = w_k sum_n (x_n w_n) conj(w_{k-n}) (forward; the
// inverse conjugates w), the convolution by FFT_M, pointwise product with the
// precomputed kernel spectrum, inverse FFT_M.  LDS must hold M rows.
// Rader (prime L = p, bs = 2), with g a primitive root mod p:
//   X_0 = x_0 + sum_{n>0} x_n,   X_{g^-m} = x_0 + sum_q x_{g^q} w^{g^(q-m)}
// (w = exp(-+2 pi i / p)), a cyclic convolution of length p - 1 of
// a_q = x_{g^q} with b_j = w^{g^-j}, done by FFT_{p-1}, the pointwise product
// with the precomputed FFT_{p-1}(b) / (p - 1) and the inverse FFT_{p-1}.
// LDS rows of a tile: 0..p-1 the line in natural order, p..2p-2 the
// convolution (row p + q), 2p-1 x_0 and 2p X_0 (= x_0 + A_0, the DC term of
// FFT_{p-1}(a)).  The two permutations move between disjoint rows, one
// element at a time.
template <typename T, int E, int NT, bool FWD>
__device__ __forceinline__ void rader_dft(cpx<T>* lds, const PassArgs<T>& a, int lg, int ps) {
  const int L = a.len, Lm = L - 1, nl = 1 << lg;
  const int* __restrict__ perm = a.rperm;
  cpx<T>* conv = lds + L * ps;
  cpx<T>* x0row = lds + (2 * L - 1) * ps;
  cpx<T>* X0row = lds + 2 * L * ps;
  // the table loads of each phase are issued in groups (staged) so that a
  // thread has several in flight instead of waiting for each before its move
  staged<NT, 8, int>(
      L << lg, [&](int idx) { return perm[idx >> lg]; },   // e = 0: q = p - 1, the x_0 row
      [&](int idx, int q) {
        const int line = idx & (nl - 1), e = idx >> lg;
        conv[q * ps + line] = lds[e * ps + line];
      });
  __syncthreads();
  lds_fft<T, E, NT, true>(conv, *a.plan, lg, ps, a.tw);
  const cpx<T>* __restrict__ B = FWD ? a.bfw : a.binv;
  staged<NT, stage_k<T>(), cpx<T>>(
      Lm << lg, [&](int idx) { return B[idx >> lg]; },
      [&](int idx, const cpx<T>& bk) {
        const int line = idx & (nl - 1), e = idx >> lg;
        cpx<T>* q = conv + e * ps + line;
        const cpx<T> v = *q;
        if (e == 0) X0row[line] = cadd(x0row[line], v);
        *q = cmul(v, bk);
      });
  __syncthreads();
  lds_fft<T, E, NT, false>(conv, *a.plan, lg, ps, a.tw);
  // X_{g^-m} = x_0 + c_m into row g^-m (1..p-1); the last index: X_0 into row 0
  staged<NT, 8, int>(
      L << lg, [&](int idx) { return perm[L + min(idx >> lg, Lm - 1)]; },
      [&](int idx, int row) {
        const int line = idx & (nl - 1), m = idx >> lg;
        if (m < Lm)
          lds[row * ps + line] = cadd(conv[m * ps + line], x0row[line]);
        else
          lds[line] = X0row[line];
      });
  __syncthreads();
}

template <typename T, int E, int NT, bool FWD>
__device__ __forceinline__ void line_dft(cpx<T>* lds, const PassArgs<T>& a, int lg, int ps) {
  if (!a.bs) {
    lds_fft<T, E, NT, FWD>(lds, *a.plan, lg, ps, a.tw);
    return;
  }
  if (a.bs == 2) {
    rader_dft<T, E, NT, FWD>(lds, a, lg, ps);
    return;
  }
  const int L = a.len, M = uniform(a.plan->L), nl = 1 << lg;
  constexpr int K = stage_k<T>();
  // chirp and kernel-spectrum loads issued in groups (staged), as in the
  // tile loads
  staged<NT, K, cpx<T>>(
      M << lg, [&](int idx) { return a.chirp[min(idx >> lg, L - 1)]; },
      [&](int idx, const cpx<T>& wc) {
        const int line = idx & (nl - 1), e = idx >> lg;
        cpx<T>* q = lds + e * ps + line;
        if (e < L) {
          cpx<T> w = wc;
          if (!FWD) w.y = -w.y;
          *q = cmul(*q, w);
        } else {
          *q = {(T)0, (T)0};
        }
      });
  __syncthreads();
  lds_fft<T, E, NT, true>(lds, *a.plan, lg, ps, a.tw);
  const cpx<T>* B = FWD ? a.bfw : a.binv;
  staged<NT, K, cpx<T>>(
      M << lg, [&](int idx) { return B[idx >> lg]; },
      [&](int idx, const cpx<T>& bk) {
        const int line = idx & (nl - 1), e = idx >> lg;
        cpx<T>* q = lds + e * ps + line;
        *q = cmul(*q, bk);
      });
  __syncthreads();
  lds_fft<T, E, NT, false>(lds, *a.plan, lg, ps, a.tw);
  staged<NT, K, cpx<T>>(
      L << lg, [&](int idx) { return a.chirp[idx >> lg]; },
      [&](int idx, const cpx<T>& wc) {
        const int line = idx & (nl - 1), e = idx >> lg;
        cpx<T>* q = lds + e * ps + line;
        cpx<T> w = wc;
        if (!FWD) w.y = -w.y;
        *q = cmul(*q, w);
      });
  __syncthreads();
}

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

// Workgroups are dispatched to the 8 XCDs round robin (block b on XCD b % 8),
// each XCD with its own L2.  xcd_block() numbers the blocks so that blocks
// b, b + 8, b + 16, .. (one XCD) take consecutive numbers: adjacent tiles,
// which share the cache lines at their edges, then meet in one L2.  A
// bijection of [0, gridDim.x); the last gridDim.x % 8 blocks keep their own.
__device__ __forceinline__ int xcd_block() {
  const int b = (int)blockIdx.x, q = (int)gridDim.x >> 3;
  return b >= (q << 3) ? b : (b & 7) * q + (b >> 3);
}

// a fetched source value: mode -1 = outside the tile (zero), otherwise the
// u_src map of an SRC_U load (other sources: the value as loaded)
template <typename T>
struct Raw {
  cpx<T> w;
  int mode;
};

template <typename T, int SRC>
__device__ __forceinline__ cpx<T> raw_value(const Raw<T>& r) {
  if (r.mode < 0) return {(T)0, (T)0};
  return SRC == SRC_U ? u_map(r.w, r.mode) : r.w;
}

// a spectrum value read from LDS with its epilogue operands
template <typename T>
struct Fetched {
  cpx<T> m;
  EpiOps<T> ops;
};

// one workgroup's tile of a line pass (tile = its index within the pass)
template <typename T, int NT, int E, int D, int SRC, int DST, bool FWD>
__device__ __forceinline__ void lines_body(const PassArgs<T>& a, int tile) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  // finished instances: the flag is read beside the loads, not before them
  const int dflag = a.done ? *a.done : 0;
  const int L = a.len;
  const int lg = a.lg, nl = 1 << lg;
  const int tile1 = tile % a.ntile1;
  const int o = tile / a.ntile1;
  const int k10 = tile1 << lg;
  const int nv = min(nl, a.H1 - k10);
  const int64_t s2 = a.H1, s3 = (int64_t)a.H1 * a.g.n[1];
  const int64_t sl = a.axis == 1 ? s2 : s3;
  const int64_t base = k10 + (a.axis == 1 ? (int64_t)o * s3 : (int64_t)o * s2);
  const int total = L << lg;
  constexpr int K = stage_k<T>();
  // a tile whose k1 are all representatives (1 <= k1 <= P1) reads and writes
  // u at the scratch index with the plain sqrt2 scalings of u_to_v / v_to_u
  const bool interior = k10 >= 1 && k10 + nv - 1 <= a.g.P[0];
  if (SRC == SRC_U && interior) {
    staged<NT, K, Raw<T>>(
        total,
        [&](int idx) {
          const int line = idx & (nl - 1), el = idx >> lg;
          Raw<T> r;
          r.mode = line < nv ? 0 : -1;
          r.w = a.u_in[base + (line < nv ? line : 0) + el * sl];
          return r;
        },
        [&](int idx, const Raw<T>& r) { lds[idx] = raw_value<T, SRC>(r); });
  } else {
  staged<NT, K, Raw<T>>(
      total,
      [&](int idx) {
        const int line = idx & (nl - 1), el = idx >> lg;
        const bool ok = line < nv;
        const int k1 = k10 + (ok ? line : 0);
        Raw<T> r;
        r.mode = 0;
        if (SRC == SRC_BUF) {
          r.w = a.buf[base + (k1 - k10) + el * sl];
        } else {
          const int k2 = a.axis == 1 ? el : o;
          const int k3 = a.axis == 1 ? o : el;
          if (SRC == SRC_U)
            r.w = a.u_in[u_src<D>(k1, k2, k3, a.g, a.H1, r.mode)];
          else
            r.w = spec_load<D, T, SRC>(a, k1, k2, k3);
        }
        if (!ok) r.mode = -1;
        return r;
      },
      [&](int idx, const Raw<T>& r) { lds[idx] = raw_value<T, SRC>(r); });
  }
  if (dflag) return;
  __syncthreads();
  line_dft<T, E, NT, FWD>(lds, a, lg, nl);
  double acc = 0.0;
  if (DST == DST_BUF) {
    staged<NT, K, cpx<T>>(
        total, [&](int idx) { return lds[idx]; },
        [&](int idx, const cpx<T>& v) {
          const int line = idx & (nl - 1), el = idx >> lg;
          if (line < nv) a.buf[base + line + el * sl] = v;
        });
  } else if (DST == DST_U && interior) {
    staged<NT, (K > 2 ? K / 2 : 1), Fetched<T>>(
        total,
        [&](int idx) {
          const int line = idx & (nl - 1), el = idx >> lg;
          Fetched<T> f;
          f.m = lds[idx];
          f.ops = epi_fetch_u(a.epi, base + (line < nv ? line : 0) + el * sl);
          return f;
        },
        [&](int idx, const Fetched<T>& f) {
          const int line = idx & (nl - 1), el = idx >> lg;
          if (line >= nv) return;
          acc += epi_commit_u(a.epi, base + line + el * sl, scale_sq2(f.m), f.ops);
        });
  } else if (DST == DST_U) {
    staged<NT, (K > 2 ? K / 2 : 1), Fetched<T>>(
        total,
        [&](int idx) {
          const int line = idx & (nl - 1), el = idx >> lg;
          const int k1 = k10 + (line < nv ? line : 0);
          const int k2 = a.axis == 1 ? el : o;
          const int k3 = a.axis == 1 ? o : el;
          Fetched<T> f;
          f.m = lds[idx];
          f.ops = epi_fetch_u(a.epi, k1 + (int64_t)a.H1 * (k2 + (int64_t)a.g.n[1] * k3));
          return f;
        },
        [&](int idx, const Fetched<T>& f) {
          const int line = idx & (nl - 1), el = idx >> lg;
          if (line >= nv) return;
          const int k1 = k10 + line;
          const int k2 = a.axis == 1 ? el : o;
          const int k3 = a.axis == 1 ? o : el;
          const int64_t e = k1 + (int64_t)a.H1 * (k2 + (int64_t)a.g.n[1] * k3);
          acc += epi_commit_u(a.epi, e, v_to_u<D, T>(f.m, k1, k2, k3, a.g), f.ops);
        });
  } else {
    for (int idx = threadIdx.x; idx < total; idx += NT) {
      const int line = idx & (nl - 1);
      const int el = idx >> lg;
      if (line >= nv) continue;
      const int k2 = a.axis == 1 ? el : o;
      const int k3 = a.axis == 1 ? o : el;
      acc += spec_store<D, T, DST>(a, k10 + line, k2, k3, lds[idx]);
    }
  }
  if (DST != DST_BUF && a.epi.dot_part) store_partial(a.epi.dot_part, acc, tile);
}

template <typename T, int NT, int E, int D, int SRC, int DST, bool FWD>
__global__ __launch_bounds__(NT) void k_lines(PassArgs<T> a) {
  lines_body<T, NT, E, D, SRC, DST, FWD>(a, xcd_block());
}

// block -> (instance, tile) map of a launch over several instances
struct BSeg {
  const int* inst;   // instance of each block
  const int* first;  // first block of each instance (n + 1 entries)
};

template <typename T, int NT, int E, int D, int SRC, int DST, bool FWD>
__global__ __launch_bounds__(NT) void kb_lines(const PassArgs<T>* __restrict__ tab, BSeg seg) {
  const int b = xcd_block();
  const int i = uniform(seg.inst[b]);
  const PassArgs<T> a = tab[i];  // loaded before any store: stays in scalar registers
  lines_body<T, NT, E, D, SRC, DST, FWD>(a, b - uniform(seg.first[i]));
}

__device__ __forceinline__ bool miss_bit(const uint32_t* m, int64_t s) {
  return m && ((m[s >> 5] >> (s & 31)) & 1u);
}

// staged values of the dim-1 pass
template <typename T>
struct Pair {
  cpx<T> A, B;
  int mode;
};

template <typename T>
struct Masked {
  cpx<T> z;
  uint32_t wa, wb;
};

template <typename T>
struct Fetched2 {
  cpx<T> X, Y;
  EpiOps<T> ops;
};

template <typename T, int NT, int E, int D, int SRC, int DST, int MODE>
__device__ __forceinline__ void dim1_body(const PassArgs<T>& a, int tile) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  const int dflag = a.done ? *a.done : 0;   // read beside the loads
  const int N1 = a.len, H1 = a.H1;
  const int lg = a.lg, np = 1 << lg, ps = a.ps;
  const int p0 = tile << lg;
  const bool even = (N1 & 1) == 0;
  constexpr int K = stage_k<T>();
  const int nl1 = a.nlines - 1;
  if constexpr (MODE != MODE_R2C) {
    // the mask words of the tile's lines are loaded with the tile (into LDS)
    // rather than per element between the transforms: a line's N1 samples
    // span at most W words
    constexpr int kMaskWords = 640;
    constexpr int kMR = (kMaskWords + NT - 1) / NT;
    __shared__ uint32_t mw[kMR * NT];
    const int W = (N1 + 31) / 32 + 1;
    const int nlt = 2 * np;
    const bool stage_mask = a.miss && nlt * W <= kMaskWords;
    uint32_t mreg[kMR];
    if (stage_mask) {
      const int64_t nwords = ((int64_t)a.nlines * N1 + 31) >> 5;
#pragma unroll
      for (int r = 0; r < kMR; ++r) {
        const int i = threadIdx.x + r * NT;
        const int ll = i / W, j = i - ll * W;
        const int64_t l = 2 * (int64_t)p0 + ll;
        const int64_t w = ((l * N1) >> 5) + j;
        const bool ok = i < nlt * W && l < a.nlines && w < nwords;
        const uint32_t v = a.miss[ok ? w : 0];
        mreg[r] = ok ? v : 0u;
      }
    }
    const int tot = H1 << lg;
    staged<NT, K, Pair<T>>(
        tot,
        [&](int idx) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          const int la = 2 * (p0 + p);
          Pair<T> r;
          r.mode = 0;
          r.B = {(T)0, (T)0};
          if (SRC == SRC_BUF) {
            r.A = a.buf[(int64_t)min(la, nl1) * H1 + k];
            r.B = a.buf[(int64_t)min(la + 1, nl1) * H1 + k];
          } else if (SRC == SRC_U) {
            r.A = a.u_in[u_src<D>(k, 0, 0, a.g, a.H1, r.mode)];   // 1D: a single line
          } else {
            r.A = spec_load<D, T, SRC>(a, k, 0, 0);
          }
          return r;
        },
        [&](int idx, const Pair<T>& r) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          const int la = 2 * (p0 + p), lb = la + 1;
          const cpx<T> zero = {(T)0, (T)0};
          cpx<T> A = la < a.nlines ? (SRC == SRC_U ? u_map(r.A, r.mode) : r.A) : zero;
          cpx<T> B = (SRC == SRC_BUF && lb < a.nlines) ? r.B : zero;
          if (k == 0 || (even && k == H1 - 1)) {  // DC / Nyquist of a real line are real
            A.y = 0;
            B.y = 0;
          }
          lds[k * ps + p] = {A.x - B.y, A.y + B.x};
          if (k > 0 && N1 - k != k) lds[(N1 - k) * ps + p] = {A.x + B.y, B.x - A.y};
        });
    if (stage_mask) {
#pragma unroll
      for (int r = 0; r < kMR; ++r) mw[threadIdx.x + r * NT] = mreg[r];
    }
    if (dflag) return;
    __syncthreads();
    line_dft<T, E, NT, false>(lds, a, lg, ps);
    const int totr = N1 << lg;
    if (MODE == MODE_MTM && (stage_mask || !a.miss) && np <= NT) {
      // the mask in place from LDS: a thread keeps one line pair p (NT is a
      // multiple of np), so its lines, their staged mask rows and the bit
      // offset of sample 0 in their first word are fixed; per sample only
      // the word and bit of o + n are formed
      const int p = (int)threadIdx.x & (np - 1);
      const int dn = NT >> lg;
      const int la = 2 * (p0 + p), lb = la + 1;
      const T sa0 = la < a.nlines ? a.scale : (T)0;
      const T sb0 = lb < a.nlines ? a.scale : (T)0;
      const int oa = (int)(((int64_t)la * N1) & 31), ob = (int)(((int64_t)lb * N1) & 31);
      const uint32_t* ra = mw + (2 * p) * W;
      const uint32_t* rb = ra + W;
      cpx<T>* q = lds + ((int)threadIdx.x >> lg) * ps + p;
      for (int n = (int)threadIdx.x >> lg; n < N1; n += dn, q += dn * ps) {
        T sa = sa0, sb = sb0;
        if (stage_mask) {
          const int ta = oa + n, tb = ob + n;
          if ((ra[ta >> 5] >> (ta & 31)) & 1u) sa = (T)0;
          if ((rb[tb >> 5] >> (tb & 31)) & 1u) sb = (T)0;
        }
        cpx<T> z = *q;
        z.x *= sa;
        z.y *= sb;
        *q = z;
      }
    } else {
    staged<NT, K, Masked<T>>(
        totr,
        [&](int idx) {
          const int p = idx & (np - 1);
          const int n = idx >> lg;
          const int la = 2 * (p0 + p);
          Masked<T> r;
          r.z = lds[n * ps + p];
          r.wa = r.wb = 0u;
          if (stage_mask) {
            const int64_t sa = (int64_t)la * N1 + n, sb = sa + N1;
            r.wa = mw[(2 * p) * W + (int)((sa >> 5) - (((int64_t)la * N1) >> 5))];
            r.wb = mw[(2 * p + 1) * W + (int)((sb >> 5) - (((int64_t)(la + 1) * N1) >> 5))];
          } else if (a.miss) {
            r.wa = a.miss[((int64_t)min(la, nl1) * N1 + n) >> 5];
            r.wb = a.miss[((int64_t)min(la + 1, nl1) * N1 + n) >> 5];
          }
          return r;
        },
        [&](int idx, const Masked<T>& r) {
          const int p = idx & (np - 1);
          const int n = idx >> lg;
          const int la = 2 * (p0 + p), lb = la + 1;
          const int64_t sa_i = (int64_t)la * N1 + n, sb_i = (int64_t)lb * N1 + n;
          const bool ma = (r.wa >> (sa_i & 31)) & 1u;
          const bool mb = (r.wb >> (sb_i & 31)) & 1u;
          const T sa = (la < a.nlines && !ma) ? a.scale : (T)0;
          const T sb = (lb < a.nlines && !mb) ? a.scale : (T)0;
          cpx<T> z = r.z;
          z.x *= sa;
          z.y *= sb;
          if (MODE == MODE_C2R) {
            if (la < a.nlines) a.real_out[sa_i] = z.x;
            if (lb < a.nlines) a.real_out[sb_i] = z.y;
          } else {
            lds[n * ps + p] = z;
          }
        });
    }
    if (MODE == MODE_C2R) return;
    __syncthreads();
  } else {
    const int totr = N1 << lg;
    staged<NT, K, cpx<T>>(
        totr,
        [&](int idx) {
          const int p = udiv(idx, N1, a.n1m);
          const int n = idx - p * N1;
          const int la = 2 * (p0 + p);
          return cpx<T>{a.real_in[(int64_t)min(la, nl1) * N1 + n],
                        a.real_in[(int64_t)min(la + 1, nl1) * N1 + n]};
        },
        [&](int idx, const cpx<T>& r) {
          const int p = udiv(idx, N1, a.n1m);
          const int n = idx - p * N1;
          const int la = 2 * (p0 + p), lb = la + 1;
          const T xa = la < a.nlines ? r.x * a.scale : (T)0;
          const T xb = lb < a.nlines ? r.y * a.scale : (T)0;
          lds[n * ps + p] = {xa, xb};
        });
    if (dflag) return;
    __syncthreads();
  }
  line_dft<T, E, NT, true>(lds, a, lg, ps);
  double acc = 0.0;
  const int tot = H1 << lg;
  const T half = (T)0.5;
  if (DST == DST_BUF) {
    staged<NT, K, Pair<T>>(
        tot,
        [&](int idx) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          Pair<T> r;
          r.A = lds[k * ps + p];
          r.B = lds[(k == 0 ? 0 : N1 - k) * ps + p];
          r.mode = 0;
          return r;
        },
        [&](int idx, const Pair<T>& r) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          const int la = 2 * (p0 + p), lb = la + 1;
          if (la >= a.nlines) return;
          const cpx<T> X = r.A, Y = r.B;
          const cpx<T> Va = {half * (X.x + Y.x), half * (X.y - Y.y)};
          const cpx<T> Vb = {half * (X.y + Y.y), half * (Y.x - X.x)};
          a.buf[(int64_t)la * H1 + k] = Va;
          if (lb < a.nlines) a.buf[(int64_t)lb * H1 + k] = Vb;
        });
  } else if (DST == DST_U) {   // 1D: line 0 only
    staged<NT, (K > 2 ? K / 2 : 1), Fetched2<T>>(
        tot,
        [&](int idx) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          Fetched2<T> f;
          f.X = lds[k * ps + p];
          f.Y = lds[(k == 0 ? 0 : N1 - k) * ps + p];
          f.ops = epi_fetch_u(a.epi, k);
          return f;
        },
        [&](int idx, const Fetched2<T>& f) {
          const int p = udiv(idx, H1, a.h1m);
          const int k = idx - p * H1;
          if (2 * (p0 + p) >= a.nlines) return;
          const cpx<T> Va = {half * (f.X.x + f.Y.x), half * (f.X.y - f.Y.y)};
          acc += epi_commit_u(a.epi, k, v_to_u<D, T>(Va, k, 0, 0, a.g), f.ops);
        });
  } else {
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      const int p = udiv(idx, H1, a.h1m);
      const int k = idx - p * H1;
      const int la = 2 * (p0 + p);
      if (la >= a.nlines) continue;
      const cpx<T> X = lds[k * ps + p];
      const cpx<T> Y = lds[(k == 0 ? 0 : N1 - k) * ps + p];
      const cpx<T> Va = {half * (X.x + Y.x), half * (X.y - Y.y)};
      acc += spec_store<D, T, DST>(a, k, 0, 0, Va);   // 1D: line 0 only
    }
  }
  if (DST != DST_BUF && a.epi.dot_part) store_partial(a.epi.dot_part, acc, tile);
}

template <typename T, int NT, int E, int D, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void k_dim1(PassArgs<T> a) {
  dim1_body<T, NT, E, D, SRC, DST, MODE>(a, xcd_block());
}

template <typename T, int NT, int E, int D, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void kb_dim1(const PassArgs<T>* __restrict__ tab, BSeg seg) {
  const int b = xcd_block();
  const int i = uniform(seg.inst[b]);
  const PassArgs<T> a = tab[i];  // loaded before any store: stays in scalar registers
  dim1_body<T, NT, E, D, SRC, DST, MODE>(a, b - uniform(seg.first[i]));
}

// ===========================================================================
// Long 1D lines: the length-F DFT in four steps over the (B rows x A columns)
// view of the scratch spectrum (position col + A*row).  Column passes
// transform along rows (length B, stride A, 2^lg adjacent columns per
// workgroup), row passes along columns (length A, contiguous, 2^lg rows).
// Direct (F = N):
//   MTM   col FIRST (inverse, twiddle)  row MTM (inverse, mask at the time
//         index t = row + B*col, forward, twiddle)  col LAST (forward, epilogue)
// Bluestein (F = M >= 2N - 1, chirp w_n = exp(-i pi n^2/N)):
//   MTM   col FIRST (x conj(w), forward, twiddle)  row CONV_INV (forward,
//         x FFT_M(w)/M, inverse, twiddle)  col MID (inverse, x conj(w), mask,
//         x w, forward, twiddle)  row CONV_FWD  col LAST (inverse, x w, epilogue)
// ===========================================================================
enum { C_FIRST = 0, C_REAL_IN = 1, C_MID = 2, C_LAST = 3, C_MID_C2R = 4 };
enum { R_MTM = 0, R_C2R = 1, R_R2C = 2, R_CONV_INV = 3, R_CONV_FWD = 4 };

template <typename T>
__device__ __forceinline__ cpx<T> twf(const PassArgs<T>& a, int64_t j, bool inv) {
  cpx<T> w = a.twF[j];
  if (inv) w.y = -w.y;
  return w;
}

template <typename T>
__device__ __forceinline__ cpx<T> chirp_at(const PassArgs<T>& a, int64_t n, bool conj_) {
  cpx<T> w = a.chirp[n];
  if (conj_) w.y = -w.y;
  return w;
}

// spectrum value k (< N) of a Hermitian line held as its half (k < H1)
template <typename T, int SRC>
__device__ __forceinline__ cpx<T> half_load(const PassArgs<T>& a, int64_t k) {
  if (k < a.H1) return spec_load<1, T, SRC>(a, (int)k, 0, 0);
  const cpx<T> v = spec_load<1, T, SRC>(a, (int)(a.len - k), 0, 0);
  return {v.x, -v.y};
}

template <typename T, int NT, int E, int SRC, int DST, int MODE>
__device__ __forceinline__ void col_body(const PassArgs<T>& a, int tile) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const int A = a.A, B = a.B, lg = a.lg, nl = 1 << lg;
  const int64_t N = a.len;
  const bool bs = a.obs != 0;
  const int c0 = tile << lg;
  const int nv = min(nl, A - c0);
  for (int idx = threadIdx.x; idx < (B << lg); idx += NT) {
    const int line = idx & (nl - 1), r = idx >> lg;
    const int64_t n = c0 + line + (int64_t)A * r;
    cpx<T> v = {(T)0, (T)0};
    if (line < nv) {
      if (MODE == C_FIRST) {
        if (n < N) {
          v = half_load<T, SRC>(a, n);
          if (bs) v = cmul(v, chirp_at(a, n, true));
        }
      } else if (MODE == C_REAL_IN) {
        if (n < N) {
          v = {a.real_in[n] * a.scale, (T)0};
          if (bs) v = cmul(v, chirp_at(a, n, false));
        }
      } else {
        v = a.buf[n];
      }
    }
    lds[r * nl + line] = v;
  }
  __syncthreads();
  const bool fwd1 = MODE == C_FIRST ? bs : (MODE == C_REAL_IN ? true : (MODE == C_LAST ? !bs : false));
  if (fwd1) lds_fft<T, E, NT, true>(lds, *a.plan, lg, nl, a.tw);
  else lds_fft<T, E, NT, false>(lds, *a.plan, lg, nl, a.tw);
  if (MODE == C_MID || MODE == C_MID_C2R) {
    // conv -> x_n = conj(w_n) conv_n, masked real part; then (MTM) x w for the forward
    for (int idx = threadIdx.x; idx < (B << lg); idx += NT) {
      const int line = idx & (nl - 1), r = idx >> lg;
      const int64_t n = c0 + line + (int64_t)A * r;
      cpx<T>* q = lds + r * nl + line;
      T xr = (T)0;
      if (line < nv && n < N && !miss_bit(a.miss, n)) xr = cmul(*q, chirp_at(a, n, true)).x * a.scale;
      if (MODE == C_MID_C2R) {
        if (line < nv && n < N) a.real_out[n] = xr;
      } else {
        const cpx<T> w = n < N ? chirp_at(a, n, false) : cpx<T>{(T)0, (T)0};
        *q = {xr * w.x, xr * w.y};
      }
    }
    if (MODE == C_MID_C2R) return;
    __syncthreads();
    lds_fft<T, E, NT, true>(lds, *a.plan, lg, nl, a.tw);
  }
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < (B << lg); idx += NT) {
    const int line = idx & (nl - 1), r = idx >> lg;
    if (line >= nv) continue;
    const int col = c0 + line;
    cpx<T> v = lds[r * nl + line];
    if (MODE != C_LAST) {
      const bool inv = MODE == C_FIRST && !bs;
      v = cmul(v, twf(a, (int64_t)col * r, inv));
      a.buf[col + (int64_t)A * r] = v;
    } else {
      const int64_t k = col + (int64_t)A * r;
      if (k < a.H1) {
        if (bs) v = cmul(v, chirp_at(a, k, false));
        acc += spec_store<1, T, DST>(a, (int)k, 0, 0, v);
      }
    }
  }
  if (MODE == C_LAST && a.epi.dot_part) store_partial(a.epi.dot_part, acc, tile);
}

template <typename T, int NT, int E, int MODE>
__device__ __forceinline__ void row_body(const PassArgs<T>& a, int tile) {
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  if (a.done && *a.done) return;
  const int A = a.A, B = a.B, lg = a.lg, nl = 1 << lg, ps = a.ps;
  const int r0 = tile << lg;
  const int nv = min(nl, B - r0);
  const int tot = A << lg;
  for (int idx = threadIdx.x; idx < tot; idx += NT) {
    const int line = udiv(idx, A, a.am);
    const int c = idx - line * A;
    const int row = r0 + line;
    cpx<T> v = {(T)0, (T)0};
    if (line < nv) {
      if (MODE == R_R2C) v = {a.real_in[row + (int64_t)B * c] * a.scale, (T)0};
      else v = a.buf[c + (int64_t)A * row];
    }
    lds[c * ps + line] = v;
  }
  __syncthreads();
  const bool fwd1 = MODE == R_R2C || MODE == R_CONV_INV || MODE == R_CONV_FWD;
  if (fwd1) lds_fft<T, E, NT, true>(lds, *a.plan, lg, ps, a.tw);
  else lds_fft<T, E, NT, false>(lds, *a.plan, lg, ps, a.tw);
  if (MODE == R_MTM || MODE == R_C2R) {
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      const int line = udiv(idx, A, a.am);
      const int c = idx - line * A;
      const int64_t t = r0 + line + (int64_t)B * c;
      cpx<T>* q = lds + c * ps + line;
      const T xr = (line < nv && !miss_bit(a.miss, t)) ? q->x * a.scale : (T)0;
      if (MODE == R_C2R) {
        if (line < nv) a.real_out[t] = xr;
      } else {
        *q = {xr, (T)0};
      }
    }
    if (MODE == R_C2R) return;
    __syncthreads();
    lds_fft<T, E, NT, true>(lds, *a.plan, lg, ps, a.tw);
  } else if (MODE == R_CONV_INV || MODE == R_CONV_FWD) {
    const cpx<T>* Bh = MODE == R_CONV_FWD ? a.bfw : a.binv;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      const int line = udiv(idx, A, a.am);
      const int c = idx - line * A;
      cpx<T>* q = lds + c * ps + line;
      if (line < nv) *q = cmul(*q, Bh[r0 + line + (int64_t)B * c]);
    }
    __syncthreads();
    lds_fft<T, E, NT, false>(lds, *a.plan, lg, ps, a.tw);
  }
  const bool inv_tw = MODE == R_CONV_INV || MODE == R_CONV_FWD;
  for (int idx = threadIdx.x; idx < tot; idx += NT) {
    const int line = udiv(idx, A, a.am);
    const int c = idx - line * A;
    if (line >= nv) continue;
    const int row = r0 + line;
    a.buf[c + (int64_t)A * row] = cmul(lds[c * ps + line], twf(a, (int64_t)row * c, inv_tw));
  }
}

template <typename T, int NT, int E, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void k_col(PassArgs<T> a) {
  col_body<T, NT, E, SRC, DST, MODE>(a, xcd_block());
}
template <typename T, int NT, int E, int SRC, int DST, int MODE>
__global__ __launch_bounds__(NT) void kb_col(const PassArgs<T>* __restrict__ tab, BSeg seg) {
  const int b = xcd_block();
  const int i = uniform(seg.inst[b]);
  const PassArgs<T> a = tab[i];  // loaded before any store: stays in scalar registers
  col_body<T, NT, E, SRC, DST, MODE>(a, b - uniform(seg.first[i]));
}
template <typename T, int NT, int E, int MODE>
__global__ __launch_bounds__(NT) void k_row(PassArgs<T> a) {
  row_body<T, NT, E, MODE>(a, xcd_block());
}
template <typename T, int NT, int E, int MODE>
__global__ __launch_bounds__(NT) void kb_row(const PassArgs<T>* __restrict__ tab, BSeg seg) {
  const int b = xcd_block();
  const int i = uniform(seg.inst[b]);
  const PassArgs<T> a = tab[i];  // loaded before any store: stays in scalar registers
  row_body<T, NT, E, MODE>(a, b - uniform(seg.first[i]));
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
                                                   MapGeom g, int H1, int64_t nc, int sign) {
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
    if (!vertex) beta[im] = (neg && sign) ? -w.y : w.y;
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
