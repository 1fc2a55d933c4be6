// GLL quadrature, Lagrange interpolants and control-node shape functions for
// NGLL = 5.
//
//   GLL nodes/weights ..... src/quadrature/gll/gll_library.cpp:102-148
//   hprime(i, j) = l_j'(xi_i) src/quadrature/gll/lagrange_poly.cpp:161-191
//   interpolants .......... src/quadrature/gll/lagrange_poly.cpp:12-43
//   shape functions ....... src/jacobian/shape_functions.cpp:17-152
//
// Arithmetic follows the reference's type_real (float) expressions so that
// GLL point coordinates, and therefore the numbering and source/receiver
// location, come out the same.
#pragma once

#include <cmath>
#include <stdexcept>

namespace sfx {

constexpr int NGLL = 5;

template <class T> inline T legendre(T z, int n) {
  T p1 = 1.0, p2 = z, p3 = p2;
  for (int k = 1; k < n; k++) {
    const double dk = static_cast<double>(k);
    p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

template <class T> inline T legendre_derivative(T z, int n) {
  T p1 = 1.0, p2 = z, d1 = 0.0, d2 = 1.0, d3 = 1.0;
  for (int k = 1; k < n; k++) {
    const double dk = static_cast<double>(k);
    const T p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    d3 = ((2.0 * dk + 1.0) * p2 + (2.0 * dk + 1.0) * z * d2 - dk * d1) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
    d1 = d2;
    d2 = d3;
  }
  return d3;
}

struct Gll {
  float xi[NGLL];
  float w[NGLL];
  float hprime[NGLL][NGLL];  // [i][j] = l_j'(xi_i)
  double xi_d[NGLL], w_d[NGLL];
};

// Interior nodes are the roots of P'_{N-1} (Newton in double, rounded once).
inline Gll make_gll() {
  Gll g;
  const int n = NGLL - 1;
  g.xi_d[0] = -1.0;
  g.xi_d[n] = 1.0;
  for (int i = 1; i < n; i++) {
    double z = -std::cos(M_PI * i / n);
    for (int it = 0; it < 100; it++) {
      const double p = legendre<double>(z, n), dp = legendre_derivative<double>(z, n);
      const double d2p = (2.0 * z * dp - n * (n + 1.0) * p) / (1.0 - z * z);
      const double dz = dp / d2p;
      z -= dz;
      if (std::abs(dz) < 1e-17) break;
    }
    g.xi_d[i] = z;
  }
  g.xi_d[n / 2] = 0.0;
  for (int i = 0; i < NGLL; i++) {
    const double p = legendre<double>(g.xi_d[i], n);
    g.w_d[i] = 2.0 / (n * (n + 1.0) * p * p);
    g.xi[i] = static_cast<float>(g.xi_d[i]);
    g.w[i] = static_cast<float>(g.w_d[i]);
  }
  for (int i = 0; i < NGLL; i++)
    for (int j = 0; j < NGLL; j++) {
      float v;
      if (i == 0 && j == 0)
        v = -1.0f * static_cast<float>(n) * (static_cast<float>(n) + 1.0) * 0.25;
      else if (i == n && j == n)
        v = 1.0f * static_cast<float>(n) * (static_cast<float>(n) + 1.0) * 0.25;
      else if (i == j)
        v = 0.0f;
      else
        v = legendre<float>(g.xi[i], n) / (legendre<float>(g.xi[j], n) * (g.xi[i] - g.xi[j])) +
            (1.0 - g.xi[i] * g.xi[i]) * legendre_derivative<float>(g.xi[i], n) /
                (static_cast<float>(n) * (static_cast<float>(n) + 1.0) * legendre<float>(g.xi[j], n) *
                 (g.xi[i] - g.xi[j]) * (g.xi[i] - g.xi[j]));
      g.hprime[i][j] = v;
    }
  return g;
}

template <class T> inline void lagrange(T xi, const T *nodes, T *h, T *hp) {
  for (int d = 0; d < NGLL; d++) {
    T prod1 = 1.0, prod2 = 1.0, sum = 0.0;
    const T x0 = nodes[d];
    for (int i = 0; i < NGLL; i++) {
      if (i == d) continue;
      prod1 = prod1 * (xi - nodes[i]);
      prod2 = prod2 * (x0 - nodes[i]);
      T prod3 = 1.0;
      for (int j = 0; j < NGLL; j++)
        if (j != d && j != i) prod3 = prod3 * (xi - nodes[j]);
      sum = sum + prod3;
    }
    const T inv = 1.0 / prod2;
    h[d] = prod1 * inv;
    if (hp) hp[d] = sum * inv;
  }
}

// 2-D control-node shape functions and derivatives (4 or 9 nodes).
inline void shape2d(float xi, float gamma, int ngnod, float *s) {
  const float one = 1.0f, quarter = 0.25f, half = 0.5f;
  const float sp = xi + one, sm = xi - one, tp = gamma + one, tm = gamma - one;
  if (ngnod == 4) {
    s[0] = quarter * sm * tm;
    s[1] = -quarter * sp * tm;
    s[2] = quarter * sp * tp;
    s[3] = -quarter * sm * tp;
  } else if (ngnod == 9) {
    const float ss = xi * xi, tt = gamma * gamma, st = xi * gamma;
    s[0] = quarter * sm * st * tm;
    s[1] = quarter * sp * st * tm;
    s[2] = quarter * sp * st * tp;
    s[3] = quarter * sm * st * tp;
    s[4] = half * tm * gamma * (one - ss);
    s[5] = half * sp * xi * (one - tt);
    s[6] = half * tp * gamma * (one - ss);
    s[7] = half * sm * xi * (one - tt);
    s[8] = (one - ss) * (one - tt);
  } else {
    throw std::invalid_argument("ngnod must be 4 or 9");
  }
}

inline void dshape2d(float xi, float gamma, int ngnod, float *d0, float *d1) {
  const float one = 1.0f, two = 2.0f, quarter = 0.25f, half = 0.5f;
  const float sp = xi + one, sm = xi - one, tp = gamma + one, tm = gamma - one;
  if (ngnod == 4) {
    d0[0] = quarter * tm;
    d0[1] = -quarter * tm;
    d0[2] = quarter * tp;
    d0[3] = -quarter * tp;
    d1[0] = quarter * sm;
    d1[1] = -quarter * sp;
    d1[2] = quarter * sp;
    d1[3] = -quarter * sm;
  } else if (ngnod == 9) {
    const float s2 = xi * two, t2 = gamma * two, ss = xi * xi, tt = gamma * gamma, st = xi * gamma;
    d0[0] = quarter * tm * gamma * (s2 - one);
    d0[1] = quarter * tm * gamma * (s2 + one);
    d0[2] = quarter * tp * gamma * (s2 + one);
    d0[3] = quarter * tp * gamma * (s2 - one);
    d1[0] = quarter * sm * xi * (t2 - one);
    d1[1] = quarter * sp * xi * (t2 - one);
    d1[2] = quarter * sp * xi * (t2 + one);
    d1[3] = quarter * sm * xi * (t2 + one);
    d0[4] = -one * st * tm;
    d0[5] = half * (one - tt) * (s2 + one);
    d0[6] = -one * st * tp;
    d0[7] = half * (one - tt) * (s2 - one);
    d1[4] = half * (one - ss) * (t2 - one);
    d1[5] = -one * st * sp;
    d1[6] = half * (one - ss) * (t2 + one);
    d1[7] = -one * st * sm;
    d0[8] = -one * s2 * (one - tt);
    d1[8] = -one * t2 * (one - ss);
  } else {
    throw std::invalid_argument("ngnod must be 4 or 9");
  }
}

// Ricker wavelet, src/source_time_function/ricker.cpp:39-50 with
// d2gaussian of utilities.cpp:27-32.
inline float ricker(float t, float f0, float factor) {
  const float pi = static_cast<float>(M_PI);
  const float a = pi * pi * f0 * f0;
  const float d2 = (1.0 - 2.0 * a * t * t) * std::exp(-a * t * t);
  return -1.0 * factor * d2;
}

// Dirac: -factor times the Gaussian of utilities.cpp:6-12 (the second
// integral of a Ricker) at f0 = 1 / (10 dt), src/source_time_function/
// dirac.cpp:20-48.
inline float dirac(float t, float f0, float factor) {
  const float pi = static_cast<float>(M_PI);
  const float a = pi * pi * f0 * f0;
  const float g = -1.0 * std::exp(-a * t * t) / (2.0 * a);
  return -1.0 * factor * g;
}

}  // namespace sfx
