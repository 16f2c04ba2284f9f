// Device-side building blocks of libcsipm: complex arithmetic, the Hermitian
// beta <-> DFT packing geometry, Stockham butterflies and block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radix_consts.hpp"

namespace csipm {

// ---------------------------------------------------------------------------
// complex numbers (interleaved re, im; 16 B for double, 8 B for float)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (FWD) or +i (inverse)
template <typename T, bool FWD>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) {
  return FWD ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
}

// ---------------------------------------------------------------------------
// Packing geometry.  Per dimension d of length n: M = n/2, P = (n-1)/2,
// self-conjugate indices S = {0, M} (n even) or {0} (n odd), representatives
// R = 1..P, second half H = M+1..M+P, conjugate partner (n - k) mod n.
// The beta vector is the reference's block sequence (vertices, edges, faces,
// interior; Re block then Im block), src/mapping_cpu_kernels.jl:1-351.
// ---------------------------------------------------------------------------
struct MapGeom {
  int d;
  int n[3], M[3], P[3], S[3];
  int64_t N;
  int64_t P12, P13, P23, P123;
  // block offsets; 3D: V, Ek, Ej, Ei, Fi, Fj, Fk, I.  2D: V, Er, Ec, C.
  // 1D: V, R.
  int64_t off[8];
};

// Geometry of the packing for sizes n[0..dim-1] (trailing sizes 1).
inline MapGeom make_map_geom(int dim, const int n_in[3]) {
  MapGeom g{};
  g.d = dim;
  for (int i = 0; i < 3; ++i) {
    const int n = i < dim ? n_in[i] : 1;
    g.n[i] = n;
    g.M[i] = n / 2;
    g.P[i] = (n - 1) / 2;
    g.S[i] = (n % 2 == 0) ? 2 : 1;
  }
  g.N = (int64_t)g.n[0] * g.n[1] * g.n[2];
  g.P12 = (int64_t)g.P[0] * g.P[1];
  g.P13 = (int64_t)g.P[0] * g.P[2];
  g.P23 = (int64_t)g.P[1] * g.P[2];
  g.P123 = g.P12 * g.P[2];
  const int64_t S1 = g.S[0], S2 = g.S[1], S3 = g.S[2];
  const int64_t P1 = g.P[0], P2 = g.P[1], P3 = g.P[2];
  for (int i = 0; i < 8; ++i) g.off[i] = 0;
  if (dim == 1) {
    g.off[1] = S1;
  } else if (dim == 2) {
    g.off[1] = S1 * S2;
    g.off[2] = g.off[1] + 2 * S1 * P2;
    g.off[3] = g.off[2] + 2 * S2 * P1;
  } else {
    g.off[1] = S1 * S2 * S3;
    g.off[2] = g.off[1] + 2 * S1 * S2 * P3;
    g.off[3] = g.off[2] + 2 * S1 * S3 * P2;
    g.off[4] = g.off[3] + 2 * S2 * S3 * P1;
    g.off[5] = g.off[4] + 4 * S1 * g.P23;
    g.off[6] = g.off[5] + 4 * S2 * g.P13;
    g.off[7] = g.off[6] + 4 * S3 * g.P12;
  }
  return g;
}

// category of index k: c = 0 (S, v = slot 0/1), 1 (R, v = k-1), 2 (H, v = k-M-1)
struct Cat {
  int c;
  int v;
};

__host__ __device__ __forceinline__ Cat cat_of(int k, const MapGeom& g, int dim) {
  const int n = g.n[dim], M = g.M[dim], P = g.P[dim];
  if (k == 0) return {0, 0};
  if (((n & 1) == 0) && k == M) return {0, 1};
  if (k <= P) return {1, k - 1};
  return {2, k - M - 1};
}

__host__ __device__ __forceinline__ Cat cat_mirror(Cat a, int P) {
  if (a.c == 1) return {2, P - 1 - a.v};
  if (a.c == 2) return {1, P - 1 - a.v};
  return a;
}

// Is the class a representative read directly by DFT_to_beta?
//  3D: the last non-self dimension is in R; 2D: the first non-self dimension
//  is in R; 1D: not in H.  (All-self classes are vertices.)
template <int D>
__host__ __device__ __forceinline__ bool is_rep(Cat c1, Cat c2, Cat c3) {
  if (D == 1) return c1.c != 2;
  if (D == 2) {
    if (c1.c != 0) return c1.c == 1;
    return c2.c != 2;
  }
  if (c3.c != 0) return c3.c == 1;
  if (c2.c != 0) return c2.c == 1;
  return c1.c != 2;
}

// beta index of the Re part of a representative class; bs = distance to the
// Im part; vertex = no Im part and no sqrt(2).
template <int D>
__host__ __device__ __forceinline__ int64_t rep_index(Cat c1, Cat c2, Cat c3, const MapGeom& g,
                                             int64_t& bs, bool& vertex) {
  vertex = false;
  if (D == 1) {
    if (c1.c == 0) { vertex = true; bs = 0; return c1.v; }
    bs = g.P[0];
    return g.off[1] + c1.v;
  }
  if (D == 2) {
    if (c1.c == 0 && c2.c == 0) { vertex = true; bs = 0; return g.off[0] + c1.v * g.S[1] + c2.v; }
    if (c1.c == 0) { bs = g.P[1]; return g.off[1] + 2 * (int64_t)c1.v * g.P[1] + c2.v; }
    if (c2.c == 0) { bs = g.P[0]; return g.off[2] + 2 * (int64_t)c2.v * g.P[0] + c1.v; }
    bs = g.P12;
    return g.off[3] + (c2.c == 2 ? 2 * g.P12 : 0) + c1.v + (int64_t)g.P[0] * c2.v;
  }
  const int S2 = g.S[1], S3 = g.S[2];
  const int64_t P1 = g.P[0], P2 = g.P[1], P3 = g.P[2];
  if (c1.c == 0 && c2.c == 0 && c3.c == 0) {
    vertex = true; bs = 0;
    return g.off[0] + (c1.v * S2 + c2.v) * S3 + c3.v;
  }
  if (c1.c == 0 && c2.c == 0) { bs = P3; return g.off[1] + 2 * (int64_t)(c1.v * S2 + c2.v) * P3 + c3.v; }
  if (c1.c == 0 && c3.c == 0) { bs = P2; return g.off[2] + 2 * (int64_t)(c1.v * S3 + c3.v) * P2 + c2.v; }
  if (c2.c == 0 && c3.c == 0) { bs = P1; return g.off[3] + 2 * (int64_t)(c2.v * S3 + c3.v) * P1 + c1.v; }
  if (c1.c == 0) {
    bs = g.P23;
    return g.off[4] + 4 * (int64_t)c1.v * g.P23 + (c2.c == 2 ? 2 * g.P23 : 0) + c2.v + P2 * c3.v;
  }
  if (c2.c == 0) {
    bs = g.P13;
    return g.off[5] + 4 * (int64_t)c2.v * g.P13 + (c1.c == 2 ? 2 * g.P13 : 0) + c1.v + P1 * c3.v;
  }
  if (c3.c == 0) {
    bs = g.P12;
    return g.off[6] + 4 * (int64_t)c3.v * g.P12 + (c1.c == 2 ? 2 * g.P12 : 0) + c1.v + P1 * c2.v;
  }
  bs = g.P123;
  const int blk = (c1.c == 2 ? 1 : 0) + (c2.c == 2 ? 2 : 0);
  return g.off[7] + 2 * (int64_t)blk * g.P123 + c1.v + P1 * (c2.v + P2 * c3.v);
}

// beta -> spectrum element (k1, k2, k3): the per-element formula of
// src/mapping_cpu_kernels.jl:363-738.  Representatives read (b_re + i b_im)/sqrt2,
// other elements are conjugates of their partner, (b_re - i b_im)/sqrt2;
// vertices are real.  Each part is one IEEE division by the double sqrt(2).
template <int D, typename T, typename BT>
__host__ __device__ __forceinline__ cpx<T> beta_to_dft_elem(const BT* __restrict__ beta, int k1, int k2,
                                                  int k3, const MapGeom& g) {
#pragma clang fp contract(off)
  Cat c1 = cat_of(k1, g, 0);
  Cat c2 = D >= 2 ? cat_of(k2, g, 1) : Cat{0, 0};
  Cat c3 = D >= 3 ? cat_of(k3, g, 2) : Cat{0, 0};
  bool neg = false;
  if (!is_rep<D>(c1, c2, c3)) {
    c1 = cat_mirror(c1, g.P[0]);
    c2 = cat_mirror(c2, g.P[1]);
    c3 = cat_mirror(c3, g.P[2]);
    neg = true;
  }
  int64_t bs;
  bool vertex;
  const int64_t re = rep_index<D>(c1, c2, c3, g, bs, vertex);
  const T sq2 = (T)1.4142135623730951;
  if (vertex) return {(T)beta[re], (T)0};
  const T a = (T)beta[re];
  const T b = (T)beta[re + bs];
  return {a / sq2, (neg ? -b : b) / sq2};
}

// Targets of a half-spectrum element in DFT_to_beta: beta indices reading its
// real and imaginary parts.  Returns false if no beta entry reads it.
//   scale_im = +1 (direct read) or -1 (3D mirror blocks read -Im of this element)
template <int D>
__host__ __device__ __forceinline__ bool scatter_targets(int k1, int k2, int k3, const MapGeom& g,
                                                int64_t& re, int64_t& im, bool& vertex,
                                                bool& neg_im) {
  Cat c1 = cat_of(k1, g, 0);
  Cat c2 = D >= 2 ? cat_of(k2, g, 1) : Cat{0, 0};
  Cat c3 = D >= 3 ? cat_of(k3, g, 2) : Cat{0, 0};
  int64_t bs;
  neg_im = false;
  if (is_rep<D>(c1, c2, c3)) {
    re = rep_index<D>(c1, c2, c3, g, bs, vertex);
    im = re + bs;
    return true;
  }
  if (c1.c != 1) return false;  // redundant conjugate inside a self-conjugate plane
  c1 = cat_mirror(c1, g.P[0]);
  c2 = cat_mirror(c2, g.P[1]);
  c3 = cat_mirror(c3, g.P[2]);
  re = rep_index<D>(c1, c2, c3, g, bs, vertex);
  im = re + bs;
  neg_im = true;
  return true;
}

// ---------------------------------------------------------------------------
// Radix-R DFT butterflies on registers.  FWD: exp(-2 pi i jk/R).
// ---------------------------------------------------------------------------
template <typename T, bool FWD>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft4(cpx<T>* v) {
  cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cpx<T> t2 = cadd(v[1], v[3]), t3 = mul_mi<T, FWD>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft8(cpx<T>* v) {
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, FWD>(e);
  dft4<T, FWD>(o);
  const T h = (T)0.70710678118654752440;
  // w^1 = (h, -+h), w^2 = -+i, w^3 = (-h, -+h)
  cpx<T> w1o = FWD ? cpx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)}
                   : cpx<T>{h * (o[1].x - o[1].y), h * (o[1].y + o[1].x)};
  cpx<T> w2o = mul_mi<T, FWD>(o[2]);
  cpx<T> w3o = FWD ? cpx<T>{h * (-o[3].x + o[3].y), h * (-o[3].y - o[3].x)}
                   : cpx<T>{h * (-o[3].x - o[3].y), h * (-o[3].y + o[3].x)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], w1o);
  v[5] = csub(e[1], w1o);
  v[2] = cadd(e[2], w2o);
  v[6] = csub(e[2], w2o);
  v[3] = cadd(e[3], w3o);
  v[7] = csub(e[3], w3o);
}

template <typename T, bool FWD>
__device__ __forceinline__ void dft16(cpx<T>* v) {
  // 4 x 4: columns of stride 4, twiddle, rows
  cpx<T> a[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int p = 0; p < 4; ++p) a[q][p] = v[q + 4 * p];
    dft4<T, FWD>(a[q]);
  }
  // twiddle a[q][p] *= w16^(q*p)
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  const T h = (T)0.70710678118654752440;
  const T cs[10] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, 0};
  (void)cs;
#pragma unroll
  for (int q = 1; q < 4; ++q) {
#pragma unroll
    for (int p = 1; p < 4; ++p) {
      const int m = q * p;  // 1..9
      T c, s;
      switch (m) {
        case 1: c = c1; s = s1; break;
        case 2: c = h; s = h; break;
        case 3: c = s1; s = c1; break;
        case 4: c = 0; s = 1; break;
        case 6: c = -h; s = h; break;
        default: c = -c1; s = -s1; break;  // m == 9: angle 9*pi/8 -> cos=-c1, sin=-s1
      }
      const T sg = FWD ? -s : s;
      cpx<T> w = {c, sg};
      a[q][p] = cmul(a[q][p], w);
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    cpx<T> r[4] = {a[0][p], a[1][p], a[2][p], a[3][p]};
    dft4<T, FWD>(r);
#pragma unroll
    for (int q = 0; q < 4; ++q) v[p + 4 * q] = r[q];
  }
}

// odd R = 2H+1: X[k] = x0 + sum_j a_j cos(2pi jk/R) -/+ i sum_j b_j sin(2pi jk/R)
// with a_j = x_j + x_{R-j}, b_j = x_j - x_{R-j}
template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_odd(cpx<T>* v) {
  constexpr int H = (R - 1) / 2;
  cpx<T> a[H], b[H];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    a[j - 1] = cadd(v[j], v[R - j]);
    b[j - 1] = csub(v[j], v[R - j]);
  }
  const cpx<T> x0 = v[0];
  cpx<T> y0 = x0;
#pragma unroll
  for (int j = 0; j < H; ++j) y0 = cadd(y0, a[j]);
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    T re = x0.x, im = x0.y, sre = 0, sim = 0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const int m = (j * k) % R;
      const T c = (T)OddRadix<R>::c[m];
      const T s = (T)OddRadix<R>::s[m];
      re += a[j - 1].x * c;
      im += a[j - 1].y * c;
      sre += b[j - 1].y * s;
      sim += b[j - 1].x * s;
    }
    if (FWD) {
      v[k] = {re + sre, im - sim};
      v[R - k] = {re - sre, im + sim};
    } else {
      v[k] = {re - sre, im + sim};
      v[R - k] = {re + sre, im - sim};
    }
  }
  v[0] = y0;
}

// the same, writing X[k] to dst[k * stride] as it is formed: a_j and b_j
// overwrite the inputs in place and no output array is held, which keeps the
// large odd radices (11, 13, 15, 17) within the register budget of the small
// ones
template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_odd_store(cpx<T>* v, cpx<T>* __restrict__ dst, int stride) {
  constexpr int H = (R - 1) / 2;
  const cpx<T> x0 = v[0];
  cpx<T> y0 = x0;
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    const cpx<T> a = cadd(v[j], v[R - j]);
    const cpx<T> b = csub(v[j], v[R - j]);
    v[j] = a;
    v[R - j] = b;
    y0 = cadd(y0, a);
  }
  dst[0] = y0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    T re = x0.x, im = x0.y, sre = 0, sim = 0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const int m = (j * k) % R;
      const T c = (T)OddRadix<R>::c[m];
      const T sn = (T)OddRadix<R>::s[m];
      re += v[j].x * c;
      im += v[j].y * c;
      sre += v[R - j].y * sn;
      sim += v[R - j].x * sn;
    }
    if (FWD) {
      dst[k * stride] = {re + sre, im - sim};
      dst[(R - k) * stride] = {re - sre, im + sim};
    } else {
      dst[k * stride] = {re - sre, im + sim};
      dst[(R - k) * stride] = {re + sre, im - sim};
    }
  }
}

template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (R == 2) dft2<T, FWD>(v);
  else if constexpr (R == 4) dft4<T, FWD>(v);
  else if constexpr (R == 8) dft8<T, FWD>(v);
  else if constexpr (R == 16) dft16<T, FWD>(v);
  else dft_odd<T, R, FWD>(v);
}

// ---------------------------------------------------------------------------
// Stockham autosort FFT of 2^lg lines of length L held in LDS in
// [element][line] order: element e of line l at lds[e * ps + l] (ps >= 2^lg,
// the line stride, padded where the load loops would otherwise conflict).
// Stage s with radix R after Ns = product of earlier radices: butterfly j
// reads elements j + r*L/R, multiplies input r by w^(r*(j mod Ns)*L/(Ns*R)),
// and writes elements (j - j mod Ns)*R + j mod Ns + r*Ns.  Butterflies are
// numbered b = j * 2^lg + l, so the lines of a tile run on adjacent lanes and
// share twiddles.  Each thread stages at most E elements per stage in
// registers; the caller guarantees 2^lg * L <= NT * E.
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 20;

struct FftPlan {
  int L;
  int nstages;
  int radix[kMaxStages];
  int ns[kMaxStages];        // product of the radices of earlier stages
  int tws[kMaxStages];       // twiddle index stride L / (ns * radix)
  uint32_t nsm[kMaxStages];  // ceil(2^32 / ns), ns >= 2: j / ns = umulhi(j, nsm)
};

// a wave-uniform value into a scalar register (the value must be uniform)
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// exact for a, d <= 2^16 (host checks the plan limits)
__device__ __forceinline__ int fdiv(int a, uint32_t m) { return (int)__umulhi((uint32_t)a, m); }

// 15-point DFT by the prime-factor (Good-Thomas) split 15 = 3 x 5, which
// needs no twiddles: with x read at n = (5a + 3b) mod 15 and X written at
// k = (10 k1 + 6 k2) mod 15, X = DFT3 over a of DFT5 over b (W15^(nk) =
// W3^(a k1) W5^(b k2) for these maps).  3 five-point and 5 three-point
// transforms instead of the 15-point pair form's 7 x 7 products.
template <typename T, bool FWD>
__device__ __forceinline__ void dft15_store(cpx<T>* v, cpx<T>* __restrict__ dst, int stride) {
  cpx<T> y[3][5];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 5; ++b) y[a][b] = v[(5 * a + 3 * b) % 15];
    dft_odd<T, 5, FWD>(y[a]);
  }
#pragma unroll
  for (int k2 = 0; k2 < 5; ++k2) {
    cpx<T> z[3] = {y[0][k2], y[1][k2], y[2][k2]};
    dft_odd<T, 3, FWD>(z);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) dst[((10 * k1 + 6 * k2) % 15) * stride] = z[k1];
  }
}

template <typename T, int R, int E, int NT, bool FWD>
__device__ __forceinline__ void stockham_stage(cpx<T>* __restrict__ lds, const FftPlan& p, int s,
                                               int lg, int ps, const cpx<T>* __restrict__ tw) {
  constexpr int NQ = (E + R - 1) / R;
  const int L = uniform(p.L);
  const int nb = L / R;
  const int nl = 1 << lg;
  const int total = nb << lg;
  const int Ns = uniform(p.ns[s]);
  const int tws = uniform(p.tws[s]);
  const uint32_t nsm = (uint32_t)uniform((int)p.nsm[s]);
  // an opaque copy of the lane index: keeps the compiler from hoisting every
  // stage's per-lane address arithmetic out of the caller's stage loop (the
  // hoisted values of all radix cases would stay live across the whole FFT)
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  cpx<T> v[NQ][R];
  // radices up to 17: twiddle loads first (index clamped into the table for
  // idle lanes), so that their latency overlaps the LDS reads and the barrier
  // instead of following them.  Radices 19..31 load each twiddle where it is
  // used: holding R of them beside the R inputs would exceed the registers
  // of the FP64 kernels.
  constexpr bool kAhead = R <= 17;
  cpx<T> w[NQ][kAhead ? R : 1];
  if (kAhead && Ns > 1) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = min(tid + q * NT, total - 1);
      const int j = b >> lg;
      const int step = (j - fdiv(j, nsm) * Ns) * tws;
#pragma unroll
      for (int r = 1; r < R; ++r) w[q][kAhead ? r : 0] = tw[r * step];
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int b = tid + q * NT;
    if (b < total) {
      const int j = b >> lg;
      const int line = b & (nl - 1);
      const cpx<T>* src = lds + j * ps + line;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * nb * ps];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int b = tid + q * NT;
    if (b < total) {
      const int j = b >> lg;
      const int line = b & (nl - 1);
      int k = 0;
      if (Ns > 1) {
        k = j - fdiv(j, nsm) * Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> t = kAhead ? w[q][kAhead ? r : 0] : tw[r * k * tws];
          if (!FWD) t.y = -t.y;
          v[q][r] = cmul(v[q][r], t);
        }
      }
      cpx<T>* dst = lds + ((j - k) * R + k) * ps + line;
      if constexpr (R == 2 || R == 4 || R == 8 || R == 16) {
        dft<T, R, FWD>(v[q]);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns * ps] = v[q][r];
      } else {
        if constexpr (R == 15)
          dft15_store<T, FWD>(v[q], dst, Ns * ps);
        else
          dft_odd_store<T, R, FWD>(v[q], dst, Ns * ps);
      }
    }
  }
  __syncthreads();
}

template <typename T, int E, int NT, bool FWD>
__device__ __forceinline__ void lds_fft(cpx<T>* lds, const FftPlan& plan, int lg, int ps,
                        const cpx<T>* __restrict__ tw) {
  const int nst = uniform(plan.nstages);
  for (int s = 0; s < nst; ++s) {
    switch (uniform(plan.radix[s])) {
      case 2: stockham_stage<T, 2, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 3: stockham_stage<T, 3, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 4: stockham_stage<T, 4, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 5: stockham_stage<T, 5, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 7: stockham_stage<T, 7, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 8: stockham_stage<T, 8, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 9: stockham_stage<T, 9, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 11: stockham_stage<T, 11, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 13: stockham_stage<T, 13, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 15: stockham_stage<T, 15, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 16: stockham_stage<T, 16, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 17: stockham_stage<T, 17, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 19: stockham_stage<T, 19, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 23: stockham_stage<T, 23, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 29: stockham_stage<T, 29, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      case 31: stockham_stage<T, 31, E, NT, FWD>(lds, plan, s, lg, ps, tw); break;
      default: break;  // the host planner only emits the radices above
    }
  }
}

// ---------------------------------------------------------------------------
// Spectrum-ordered copy of beta ("u"): one complex slot per half-spectrum
// element e = k1 + H1*(k2 + N2*k3); stored elements hold
// u = sqrt(2) * v (v.x for vertices, whose Im slot is 0), redundant
// conjugates in the self-conjugate dim-1 planes hold 0.  Relative to beta it
// is a permutation with sign flips (the 3D mirror blocks' Im parts), so dot
// products, the diagonal KKT terms and the CG iterates carry over exactly,
// while the FFT passes read and write it along contiguous lines.
// ---------------------------------------------------------------------------
template <int D, typename T>
__host__ __device__ __forceinline__ cpx<T> u_to_v(const cpx<T>* __restrict__ u, int k1, int k2,
                                                  int k3, const MapGeom& g, int H1) {
#pragma clang fp contract(off)
  const T sq2 = (T)1.4142135623730951;
  const Cat c1 = cat_of(k1, g, 0);
  const int64_t e = k1 + (int64_t)H1 * (k2 + (int64_t)g.n[1] * k3);
  if (c1.c == 1) {
    const cpx<T> w = u[e];
    return {w.x / sq2, w.y / sq2};
  }
  const Cat c2 = D >= 2 ? cat_of(k2, g, 1) : Cat{0, 0};
  const Cat c3 = D >= 3 ? cat_of(k3, g, 2) : Cat{0, 0};
  if (is_rep<D>(c1, c2, c3)) {
    const cpx<T> w = u[e];
    if (c2.c == 0 && c3.c == 0) return {w.x, (T)0};
    return {w.x / sq2, w.y / sq2};
  }
  const int m2 = k2 == 0 ? 0 : g.n[1] - k2;
  const int m3 = k3 == 0 ? 0 : g.n[2] - k3;
  const cpx<T> w = u[k1 + (int64_t)H1 * (m2 + (int64_t)g.n[1] * m3)];
  return {w.x / sq2, -(w.y / sq2)};
}

// u_to_v in two steps, without branches around the load: u_src returns the
// slot u_to_v reads and how its value maps (0: w/sqrt2, 1: (w.x, 0),
// 2: conj(w)/sqrt2); u_map applies that map with u_to_v's arithmetic.
template <int D>
__host__ __device__ __forceinline__ int64_t u_src(int k1, int k2, int k3, const MapGeom& g,
                                                  int H1, int& mode) {
  const Cat c1 = cat_of(k1, g, 0);
  const Cat c2 = D >= 2 ? cat_of(k2, g, 1) : Cat{0, 0};
  const Cat c3 = D >= 3 ? cat_of(k3, g, 2) : Cat{0, 0};
  const bool rep = c1.c == 1 || is_rep<D>(c1, c2, c3);
  const bool vert = c1.c != 1 && rep && c2.c == 0 && c3.c == 0;
  const int m2 = (rep || k2 == 0) ? k2 : g.n[1] - k2;
  const int m3 = (rep || k3 == 0) ? k3 : g.n[2] - k3;
  mode = vert ? 1 : (rep ? 0 : 2);
  return k1 + (int64_t)H1 * (m2 + (int64_t)g.n[1] * m3);
}

template <typename T>
__host__ __device__ __forceinline__ cpx<T> u_map(cpx<T> w, int mode) {
#pragma clang fp contract(off)
  const T sq2 = (T)1.4142135623730951;
  const T x = w.x / sq2, y = w.y / sq2;
  if (mode == 1) return {w.x, (T)0};
  return {x, mode == 2 ? -y : y};
}

// sqrt2 * v, v_to_u of a representative element (kept apart from any
// following add, as in v_to_u)
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scale_sq2(cpx<T> v) {
#pragma clang fp contract(off)
  const T sq2 = (T)1.4142135623730951;
  return {sq2 * v.x, sq2 * v.y};
}

// u slot value of spectrum element (k1,k2,k3) holding v
template <int D, typename T>
__host__ __device__ __forceinline__ cpx<T> v_to_u(cpx<T> v, int k1, int k2, int k3,
                                                  const MapGeom& g) {
#pragma clang fp contract(off)
  const T sq2 = (T)1.4142135623730951;
  const Cat c1 = cat_of(k1, g, 0);
  if (c1.c == 1) return {sq2 * v.x, sq2 * v.y};
  const Cat c2 = D >= 2 ? cat_of(k2, g, 1) : Cat{0, 0};
  const Cat c3 = D >= 3 ? cat_of(k3, g, 2) : Cat{0, 0};
  if (!is_rep<D>(c1, c2, c3)) return {(T)0, (T)0};
  if (c2.c == 0 && c3.c == 0) return {v.x, (T)0};
  return {sq2 * v.x, sq2 * v.y};
}

// ---------------------------------------------------------------------------
// Deterministic block reduction (wave64 shuffles, then LDS across waves).
// The result is valid in thread 0.
// ---------------------------------------------------------------------------
template <typename Op>
__device__ __forceinline__ double block_reduce(double v, Op op) {
  __shared__ double red_smem[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_smem[w] = v;
  __syncthreads();
  double r = v;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    r = red_smem[0];
    for (int i = 1; i < nw; ++i) r = op(r, red_smem[i]);
  }
  return r;
}

struct OpSum {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct OpMax {
  __device__ double operator()(double a, double b) const { return a > b ? a : b; }
};
struct OpMin {
  __device__ double operator()(double a, double b) const { return a < b ? a : b; }
};

}  // namespace csipm
