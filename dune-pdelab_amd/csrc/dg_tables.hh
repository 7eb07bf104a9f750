// dg_tables.hh -- host-side 1D tables of the tensor-product Q_k DG operator.
//
// On an axis-aligned hexahedral grid with cell-constant A = kappa*I the volume and
// face integrals of ConvectionDiffusionDG (R:dune/pdelab/localoperator/
// convectiondiffusiondg.hh:106-188, 271-471, 684-879) factor exactly into
// Kronecker products of 1D matrices, because the Gauss rule with k+1 points per
// direction (quadratureRule(geo, 2k), R:dune/pdelab/common/quadraturerules.hh:110-120)
// integrates the 1D integrands (degree <= 2k) exactly:
//   M_ij = int phi_i phi_j,  K_ij = int phi_i' phi_j',  G_ij = int phi_i' phi_j
// on [0,1], plus the traces e0 = phi(0), e1 = phi(1) and d0 = phi'(0), d1 = phi'(1).
// phi are the equidistant Lagrange polynomials of
// R:dune/pdelab/finiteelement/qkdglagrange.hh:55-79 (nodes j/k, so e0, e1 are unit
// vectors). Everything is computed in long double and rounded once.
#ifndef PDELAB_DG_TABLES_HH
#define PDELAB_DG_TABLES_HH

#include <cmath>
#include <vector>

namespace pdelab_amd {

constexpr int kMaxN1 = 4;   // k <= 3
constexpr int kMaxQ = 8;    // load-vector quadrature points per direction

struct Tables1D {
  int k = 1, n1 = 2;
  long double M[kMaxN1][kMaxN1], K[kMaxN1][kMaxN1], G[kMaxN1][kMaxN1], Minv[kMaxN1][kMaxN1];
  long double e0[kMaxN1], e1[kMaxN1], d0[kMaxN1], d1[kMaxN1];
  long double wint[kMaxN1];  // int_0^1 phi_i
  // load-vector rule: m points, phi_i at those points
  int m = 2;
  long double qx[kMaxQ], qw[kMaxQ], P[kMaxQ][kMaxN1];
};

inline long double lagrange_p(int k, int i, long double x) {
  long double r = 1.0L;
  for (int j = 0; j <= k; ++j)
    if (j != i) r *= (k * x - j) / (long double)(i - j);
  return r;
}
inline long double lagrange_dp(int k, int i, long double x) {
  long double r = 0.0L;
  for (int j = 0; j <= k; ++j) {
    if (j == i) continue;
    long double prod = (long double)k / (long double)(i - j);
    for (int l = 0; l <= k; ++l)
      if (l != i && l != j) prod *= (k * x - l) / (long double)(i - l);
    r += prod;
  }
  return r;
}

// Gauss-Legendre with m points on [0,1], ascending
inline void gauss01(int m, long double* x, long double* w) {
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < m; ++i) {
    long double t = std::cos(pi * (i + 0.75L) / (m + 0.5L)), p0 = 1, p1 = t, dp = 1;
    for (int it = 0; it < 100; ++it) {
      p0 = 1.0L; p1 = t;
      for (int j = 1; j < m; ++j) { long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1); p0 = p1; p1 = p2; }
      if (m == 1) { p1 = t; p0 = 1; }
      dp = m * (t * p1 - p0) / (t * t - 1.0L);
      long double dt = p1 / dp;
      t -= dt;
      if (std::fabs((double)dt) < 1e-30) break;
    }
    p0 = 1.0L; p1 = t;
    for (int j = 1; j < m; ++j) { long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1); p0 = p1; p1 = p2; }
    if (m == 1) { p1 = t; p0 = 1; }
    dp = m * (t * p1 - p0) / (t * t - 1.0L);
    x[m - 1 - i] = (1.0L + t) / 2.0L;
    w[m - 1 - i] = 1.0L / ((1.0L - t * t) * dp * dp);  // = (2/((1-t^2)P'^2))/2
  }
}

inline bool build_tables(int k, int intorderadd, Tables1D& T) {
  if (k < 1 || k > 3) return false;
  T.k = k;
  T.n1 = k + 1;
  const int n1 = T.n1;
  long double gx[kMaxN1 + 1], gw[kMaxN1 + 1];
  gauss01(n1, gx, gw);  // exact for degree 2k+1
  for (int i = 0; i < n1; ++i) {
    T.e0[i] = lagrange_p(k, i, 0.0L);
    T.e1[i] = lagrange_p(k, i, 1.0L);
    T.d0[i] = lagrange_dp(k, i, 0.0L);
    T.d1[i] = lagrange_dp(k, i, 1.0L);
    T.wint[i] = 0;
    for (int j = 0; j < n1; ++j) T.M[i][j] = T.K[i][j] = T.G[i][j] = 0;
  }
  for (int q = 0; q < n1; ++q) {
    long double p[kMaxN1], dp[kMaxN1];
    for (int i = 0; i < n1; ++i) { p[i] = lagrange_p(k, i, gx[q]); dp[i] = lagrange_dp(k, i, gx[q]); }
    for (int i = 0; i < n1; ++i) {
      T.wint[i] += gw[q] * p[i];
      for (int j = 0; j < n1; ++j) {
        T.M[i][j] += gw[q] * p[i] * p[j];
        T.K[i][j] += gw[q] * dp[i] * dp[j];
        T.G[i][j] += gw[q] * dp[i] * p[j];
      }
    }
  }
  // Minv by Gauss-Jordan
  long double a[kMaxN1][2 * kMaxN1];
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < 2 * n1; ++j) a[i][j] = j < n1 ? T.M[i][j] : (j - n1 == i ? 1.0L : 0.0L);
  for (int c = 0; c < n1; ++c) {
    int piv = c;
    for (int r = c + 1; r < n1; ++r)
      if (std::fabs((double)a[r][c]) > std::fabs((double)a[piv][c])) piv = r;
    for (int j = 0; j < 2 * n1; ++j) std::swap(a[c][j], a[piv][j]);
    long double d = a[c][c];
    for (int j = 0; j < 2 * n1; ++j) a[c][j] /= d;
    for (int r = 0; r < n1; ++r) {
      if (r == c) continue;
      long double f = a[r][c];
      for (int j = 0; j < 2 * n1; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) T.Minv[i][j] = a[i][n1 + j];
  // load-vector rule of order 2k + intorderadd (the rule lambda_volume uses,
  // convectiondiffusiondg.hh:1060-1062)
  T.m = (2 * k + intorderadd) / 2 + 1;
  if (T.m > kMaxQ) return false;
  gauss01(T.m, T.qx, T.qw);
  for (int q = 0; q < T.m; ++q)
    for (int i = 0; i < n1; ++i) T.P[q][i] = lagrange_p(k, i, T.qx[q]);
  return true;
}

}  // namespace pdelab_amd

#endif
