// Host-side entry points over the SAME packing-geometry functions the GPU
// kernels use (beta_to_dft_elem, dft_to_beta_elem, scatter_targets), so the
// CPU tests can compare the device index arithmetic with the oracle and check
// every index stays inside its buffer.  No kernels; never used on the
// product path.
#include <vector>

#include "csipm_kernels.hpp"

using namespace csipm;

namespace {
MapGeom geom(int dim, const int64_t* sizes) {
  int n[3] = {1, 1, 1};
  for (int i = 0; i < dim; ++i) n[i] = (int)sizes[i];
  return make_map_geom(dim, n);
}
template <int D>
void b2d(const MapGeom& g, const double* beta, double* v, int ld1) {
  cpx<double>* out = (cpx<double>*)v;
  for (int k3 = 0; k3 < g.n[2]; ++k3)
    for (int k2 = 0; k2 < g.n[1]; ++k2)
      for (int k1 = 0; k1 < ld1; ++k1)
        out[k1 + (int64_t)ld1 * (k2 + (int64_t)g.n[1] * k3)] =
            beta_to_dft_elem<D, double>(beta, k1, k2, k3, g);
}
template <int D>
void d2b(const MapGeom& g, const double* v, double* beta, int ld1) {
  for (int64_t i = 0; i < g.N; ++i) beta[i] = dft_to_beta_elem<D>((const cpx<double>*)v, i, g, ld1);
}
template <int D>
void sc(const MapGeom& g, int64_t* re, int64_t* im, int8_t* flags, int H1) {
  for (int k3 = 0; k3 < g.n[2]; ++k3)
    for (int k2 = 0; k2 < g.n[1]; ++k2)
      for (int k1 = 0; k1 < H1; ++k1) {
        const int64_t t = k1 + (int64_t)H1 * (k2 + (int64_t)g.n[1] * k3);
        int64_t a = -1, b = -1;
        bool vertex = false, neg = false;
        const bool hit = scatter_targets<D>(k1, k2, k3, g, a, b, vertex, neg);
        re[t] = hit ? a : -1;
        im[t] = (hit && !vertex) ? b : -1;
        flags[t] = (int8_t)((hit ? 1 : 0) | (vertex ? 2 : 0) | (neg ? 4 : 0));
      }
}
}  // namespace

extern "C" {
// the geometry fields, for comparison with the oracle: n, M, P, S (3 each),
// N, P12, P13, P23, P123, off[8]
void geomchk_info(int dim, const int64_t* sizes, int64_t* out) {
  const MapGeom g = geom(dim, sizes);
  int k = 0;
  for (int i = 0; i < 3; ++i) out[k++] = g.n[i];
  for (int i = 0; i < 3; ++i) out[k++] = g.M[i];
  for (int i = 0; i < 3; ++i) out[k++] = g.P[i];
  for (int i = 0; i < 3; ++i) out[k++] = g.S[i];
  out[k++] = g.N;
  out[k++] = g.P12;
  out[k++] = g.P13;
  out[k++] = g.P23;
  out[k++] = g.P123;
  for (int i = 0; i < 8; ++i) out[k++] = g.off[i];
}
void geomchk_beta_to_dft(int dim, const int64_t* sizes, const double* beta, double* v, int full) {
  const MapGeom g = geom(dim, sizes);
  const int ld1 = full ? g.n[0] : g.n[0] / 2 + 1;
  if (dim == 1) b2d<1>(g, beta, v, ld1);
  else if (dim == 2) b2d<2>(g, beta, v, ld1);
  else b2d<3>(g, beta, v, ld1);
}
void geomchk_dft_to_beta(int dim, const int64_t* sizes, const double* v, double* beta, int full) {
  const MapGeom g = geom(dim, sizes);
  const int ld1 = full ? g.n[0] : g.n[0] / 2 + 1;
  if (dim == 1) d2b<1>(g, v, beta, ld1);
  else if (dim == 2) d2b<2>(g, v, beta, ld1);
  else d2b<3>(g, v, beta, ld1);
}
void geomchk_scatter(int dim, const int64_t* sizes, int64_t* re, int64_t* im, int8_t* flags) {
  const MapGeom g = geom(dim, sizes);
  const int H1 = g.n[0] / 2 + 1;
  if (dim == 1) sc<1>(g, re, im, flags, H1);
  else if (dim == 2) sc<2>(g, re, im, flags, H1);
  else sc<3>(g, re, im, flags, H1);
}
}

namespace {
template <int D>
void via_u(const MapGeom& g, const double* beta, double* v, double* beta_back, const double* vin) {
  const int H1 = g.n[0] / 2 + 1;
  const int64_t nc = (int64_t)H1 * g.n[1] * g.n[2];
  std::vector<cpx<double>> u(nc), u2(nc);
  // beta -> u (the k_beta_to_u rule)
  for (int64_t e = 0; e < nc; ++e) {
    const int k1 = (int)(e % H1);
    const int64_t r = e / H1;
    const int k2 = (int)(r % g.n[1]), k3 = (int)(r / g.n[1]);
    int64_t re, im;
    bool vertex, neg;
    cpx<double> w = {0, 0};
    if (scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) {
      w.x = beta[re];
      if (!vertex) w.y = neg ? -beta[im] : beta[im];
    }
    u[e] = w;
  }
  // u -> v (pass A's load) and v -> u -> beta (pass E's store + k_u_to_beta)
  for (int64_t e = 0; e < nc; ++e) {
    const int k1 = (int)(e % H1);
    const int64_t r = e / H1;
    const int k2 = (int)(r % g.n[1]), k3 = (int)(r / g.n[1]);
    const cpx<double> vv = u_to_v<D, double>(u.data(), k1, k2, k3, g, H1);
    v[2 * e] = vv.x;
    v[2 * e + 1] = vv.y;
    const cpx<double> in = {vin[2 * e], vin[2 * e + 1]};
    u2[e] = v_to_u<D, double>(in, k1, k2, k3, g);
  }
  for (int64_t e = 0; e < nc; ++e) {
    const int k1 = (int)(e % H1);
    const int64_t r = e / H1;
    const int k2 = (int)(r % g.n[1]), k3 = (int)(r / g.n[1]);
    int64_t re, im;
    bool vertex, neg;
    if (scatter_targets<D>(k1, k2, k3, g, re, im, vertex, neg)) {
      beta_back[re] = u2[e].x;
      if (!vertex) beta_back[im] = neg ? -u2[e].y : u2[e].y;
    }
  }
}
}  // namespace

extern "C" {
// v = u_to_v(beta_to_u(beta)) (must equal beta_to_DFT(beta));
// beta_back = u_to_beta(v_to_u(vin)) (must equal DFT_to_beta(vin))
void geomchk_via_u(int dim, const int64_t* sizes, const double* beta, double* v,
                   double* beta_back, const double* vin) {
  const MapGeom g = geom(dim, sizes);
  if (dim == 1) via_u<1>(g, beta, v, beta_back, vin);
  else if (dim == 2) via_u<2>(g, beta, v, beta_back, vin);
  else via_u<3>(g, beta, v, beta_back, vin);
}
}
