// pdelab_oracle.hh -- THE ORACLE (test infrastructure, not product code).
//
// A CPU restatement of dune-pdelab's FP64 algorithm for the matrix-free SIPG DG hot
// path: quadrature-point loops of ConvectionDiffusionDG, the DefaultAssembler
// traversal (each interior face once, from the cell with the larger index), the
// residual / jacobian-apply / jacobian / point-diagonal engines, and the dune-istl
// style CG / BiCGSTAB + Jacobi loops with StationaryLinearProblemSolver on top.
// It deliberately does NOT use the tensor-product (Kronecker) factorisation of the
// GPU kernels, so that agreement between the two is evidence, not a tautology.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this code, and only as the checker or the timed CPU baseline.
//
// Pins (what checks each part; tests/test_oracle.py):
//   basis / quadrature  exact rational 1D tables from sympy (tests/golden) and the
//                       closed-form 3- and 4-point Gauss rules
//   operator            assembled block matrix times x == jacobian_apply(x);
//                       symmetry of SIPG, positivity; polynomial reproduction
//                       R(I_h u) = 0 for u in Q_k; L2 convergence order k+1;
//                       point diagonal == diag of the assembled matrix
//   solvers             residual reduction <= tolerance; solution vs a direct solve
//   gradient            central finite differences of J
// Parity unpinned: bit-level YaspGrid summation order and dune-istl's exact
// iteration counts (the reference stack is not available to compare with).
#ifndef PDELAB_ORACLE_HH
#define PDELAB_ORACLE_HH

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/pdelab_gpu.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Forward-mode dual number, used to differentiate the local operator with respect
// to log(kappa) of one cell (gradient oracle).
struct Dual {
  double v, d;
  Dual(double v_ = 0.0, double d_ = 0.0) : v(v_), d(d_) {}
  Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
  Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
inline double value(double a) { return a; }
inline double value(const Dual& a) { return a.v; }

// ---------------------------------------------------------------------------
// 1D equidistant Lagrange polynomials of degree k on [0,1], nodes j/k.
// R:dune/pdelab/finiteelement/qkdglagrange.hh:55-79 (p, dp).
inline double lagrange_p(int k, int i, double x) {
  double r = 1.0;
  for (int j = 0; j <= k; ++j)
    if (j != i) r *= (k * x - j) / (i - j);
  return r;
}
inline double lagrange_dp(int k, int i, double x) {
  double r = 0.0;
  for (int j = 0; j <= k; ++j) {
    if (j == i) continue;
    double prod = (k * 1.0) / (i - j);
    for (int l = 0; l <= k; ++l)
      if (l != i && l != j) prod *= (k * x - l) / (i - l);
    r += prod;
  }
  return r;
}

// Gauss-Legendre rule on [0,1] exact for polynomials of degree `order`
// (dune-geometry's GaussLegendre family used by quadratureRule(geo, order),
// R:dune/pdelab/common/quadraturerules.hh:110-120): order/2+1 points.
// Nodes by Newton iteration on P_m in long double.
struct Rule1D {
  std::vector<double> x, w;
};
inline Rule1D gauss_legendre(int order) {
  const int m = order / 2 + 1;
  Rule1D r;
  r.x.resize(m);
  r.w.resize(m);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < m; ++i) {
    long double t = std::cos(pi * (i + 0.75L) / (m + 0.5L));
    long double dp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = t;
      for (int j = 1; j < m; ++j) {
        long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1);
        p0 = p1;
        p1 = p2;
      }
      if (m == 1) { p1 = t; p0 = 1.0L; }
      dp = m * (t * p1 - p0) / (t * t - 1.0L);
      long double dt = p1 / dp;
      t -= dt;
      if (std::fabs((double)dt) < 1e-30) break;
    }
    // recompute derivative at the converged node
    long double p0 = 1.0L, p1 = t;
    for (int j = 1; j < m; ++j) {
      long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1);
      p0 = p1;
      p1 = p2;
    }
    if (m == 1) { p1 = t; p0 = 1.0L; }
    dp = m * (t * p1 - p0) / (t * t - 1.0L);
    long double w = 2.0L / ((1.0L - t * t) * dp * dp);
    // map [-1,1] -> [0,1]; store ascending
    r.x[m - 1 - i] = (double)((1.0L + t) / 2.0L);
    r.w[m - 1 - i] = (double)(w / 2.0L);
  }
  return r;
}

// ---------------------------------------------------------------------------
// Structured hexahedral grid on [0,lx]x[0,ly]x[0,lz] (YaspGrid):
// lexicographic cells, x fastest; intersections in reference-face order
// 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z.
struct Grid {
  int n[3];
  double L[3];
  double h[3];
  size_t ncells;
  explicit Grid(const pdelab_config& c) {
    n[0] = c.nx; n[1] = c.ny; n[2] = c.nz;
    L[0] = c.lx; L[1] = c.ly; L[2] = c.lz;
    for (int d = 0; d < 3; ++d) h[d] = L[d] / n[d];
    ncells = (size_t)n[0] * n[1] * n[2];
  }
  void coords(size_t c, int ijk[3]) const {
    ijk[0] = (int)(c % n[0]);
    ijk[1] = (int)((c / n[0]) % n[1]);
    ijk[2] = (int)(c / ((size_t)n[0] * n[1]));
  }
  size_t index(const int ijk[3]) const { return (size_t)ijk[0] + (size_t)n[0] * ((size_t)ijk[1] + (size_t)n[1] * ijk[2]); }
  // neighbour across face f; returns false on the domain boundary
  bool neighbor(size_t c, int f, size_t& cn) const {
    int ijk[3];
    coords(c, ijk);
    int d = f / 2;
    ijk[d] += (f % 2) ? 1 : -1;
    if (ijk[d] < 0 || ijk[d] >= n[d]) return false;
    cn = index(ijk);
    return true;
  }
  // cell geometry: lower corner and edge lengths as YaspGrid's axis-aligned
  // geometry computes them (upper - lower)
  void cell_box(size_t c, double lo[3], double ext[3]) const {
    int ijk[3];
    coords(c, ijk);
    for (int d = 0; d < 3; ++d) {
      lo[d] = ijk[d] * h[d];
      double hi = (ijk[d] + 1) * h[d];
      ext[d] = hi - lo[d];
    }
  }
};

// ---------------------------------------------------------------------------
// Parameter class: A = kappa*I, or a diagonal A = diag(a_x, a_y, a_z), constant per cell
// (permeabilityIsConstantPerCell, convectiondiffusionparameter.hh:28-62,142-157),
// c constant, b constant or a velocity field tabulated at the quadrature points,
// f / g / bctype per the problem selector of pdelab_config.
struct Param {
  pdelab_config cfg;
  std::vector<double> kappa;        // per cell, or a_x, a_y, a_z per cell (aniso)
  bool aniso = false;
  // n.A n for the faces normal to direction d (and A's diagonal entry d)
  double kappa_d(size_t c, int d) const { return aniso ? kappa[3 * c + d] : kappa[c]; }
  std::vector<double> cell_source;  // per cell, may be empty
  std::vector<double> f_tab;        // TABULATED: ncells * m^3
  std::vector<double> g_tab;        // TABULATED: boundary faces * m^2
  // Velocity field b(x) at the quadrature points (empty: the constant cfg.b), the values
  // param.b(cell, x) returns at the points where convectiondiffusiondg.hh evaluates it
  // (:178 volume, :426 skeleton, :797 boundary):
  //   b_vol:  per cell, the m^3 volume points (x fastest), (b_x, b_y, b_z) per point;
  //   b_face: per face, the m^2 face points (tangential t1 < t2, t1 fastest), the component
  //           b_d along +e_d of a face normal to d; faces of direction 0, then 1, then 2,
  //           within a direction lexicographic in the face position (x fastest; the
  //           position along d runs 0..n_d). On an interior face the skeleton evaluates b from
  //           the inside cell, which the assembler's visit rule makes the higher-index cell.
  std::vector<double> b_vol, b_face;
  size_t bface_dir[4] = {0, 0, 0, 0};
  int grid_n[3] = {0, 0, 0};
  int bm = 0;  // points per direction of the tabulation
  bool varb() const { return !b_vol.empty(); }
  void set_velocity(const double* bv, const double* bf, const int n[3], int m, size_t ncells) {
    if (!bv) { b_vol.clear(); b_face.clear(); return; }
    for (int d = 0; d < 3; ++d) grid_n[d] = n[d];
    bm = m;
    bface_dir[0] = 0;
    for (int d = 0; d < 3; ++d) {
      size_t nfd = 1;
      for (int e = 0; e < 3; ++e) nfd *= (size_t)(n[e] + (e == d ? 1 : 0));
      bface_dir[d + 1] = bface_dir[d] + nfd;
    }
    b_vol.assign(bv, bv + ncells * (size_t)m * m * m * 3);
    b_face.assign(bf, bf + bface_dir[3] * (size_t)m * m);
  }
  // b at volume point q of cell c (tabq = c m^3 + q), or the constant b
  const double* b_at(size_t tabq) const { return varb() ? &b_vol[3 * tabq] : cfg.b; }
  // b . n at face point q of the face of cell `cell` on side `side` (n = the side's outer
  // normal of that cell)
  double b_normal(const int cell[3], int side, int q) const {
    const int d = side / 2;
    const double ns = (side % 2) ? 1.0 : -1.0;
    if (!varb()) return ns * cfg.b[d];
    int pos[3] = {cell[0], cell[1], cell[2]};
    pos[d] += side % 2;
    const size_t e0 = (size_t)grid_n[0] + (d == 0), e1 = (size_t)grid_n[1] + (d == 1);
    const size_t f = bface_dir[d] + (size_t)pos[0] + e0 * ((size_t)pos[1] + e1 * (size_t)pos[2]);
    return ns * b_face[f * (size_t)bm * bm + q];
  }
  int bctype(int side) const {
    int b = cfg.bctype[side];
    return b == 0 ? PDELAB_BC_DIRICHLET : b;
  }
  double exact(const double x[3]) const {  // analytic solution where the problem has one
    if (cfg.problem == PDELAB_PROBLEM_EXP) {
      double r2 = 0;
      for (int d = 0; d < 3; ++d) r2 += (x[d] - 0.5) * (x[d] - 0.5);
      return std::exp(-r2);
    }
    if (cfg.problem == PDELAB_PROBLEM_QUADRATIC) {
      const double* q = cfg.quad;
      return q[0] + q[1] * x[0] + q[2] * x[1] + q[3] * x[2] + q[4] * x[0] * x[0] + q[5] * x[1] * x[1] + q[6] * x[2] * x[2];
    }
    if (cfg.problem == PDELAB_PROBLEM_CONSTANT) return cfg.g0;
    return 0.0;
  }
  // f(x) (ConvectionDiffusionModelProblem::f); tabq = index into f_tab
  double f(size_t cell, const double x[3], size_t tabq) const {
    double val = 0.0;
    switch (cfg.problem) {
      case PDELAB_PROBLEM_ZERO: val = 0.0; break;
      case PDELAB_PROBLEM_EXP: {
        double r2 = 0, bg = 0;
        for (int d = 0; d < 3; ++d) r2 += (x[d] - 0.5) * (x[d] - 0.5);
        double u = std::exp(-r2);
        const double* bb = b_at(tabq);
        for (int d = 0; d < 3; ++d) bg += bb[d] * (-2.0 * (x[d] - 0.5) * u);
        val = cfg.kappa0 * (6.0 - 4.0 * r2) * u + bg + cfg.c * u;
        break;
      }
      case PDELAB_PROBLEM_CONSTANT: val = cfg.f0; break;
      case PDELAB_PROBLEM_GAUSSIANS: {
        for (int i = 0; i < cfg.n_gaussians; ++i) {
          const double* gs = cfg.gaussians[i];
          double r2 = 0;
          for (int d = 0; d < 3; ++d) r2 += (x[d] - gs[d]) * (x[d] - gs[d]);
          val += gs[4] * std::exp(-r2 / (2.0 * gs[3] * gs[3]));
        }
        break;
      }
      case PDELAB_PROBLEM_QUADRATIC: {
        const double* q = cfg.quad;
        double lap = 2.0 * (q[4] + q[5] + q[6]);
        double grad[3] = {q[1] + 2 * q[4] * x[0], q[2] + 2 * q[5] * x[1], q[3] + 2 * q[6] * x[2]};
        const double* bb = b_at(tabq);
        double bg = bb[0] * grad[0] + bb[1] * grad[1] + bb[2] * grad[2];
        val = -cfg.kappa0 * lap + bg + cfg.c * exact(x);
        break;
      }
      case PDELAB_PROBLEM_TABULATED: val = f_tab.empty() ? 0.0 : f_tab[tabq]; break;
      default: val = 0.0;
    }
    if (!cell_source.empty()) val += cell_source[cell];
    return val;
  }
  // gradient of the analytic solution (EXP, QUADRATIC)
  void grad_exact(const double x[3], double gr[3]) const {
    if (cfg.problem == PDELAB_PROBLEM_EXP) {
      const double u = exact(x);
      for (int d = 0; d < 3; ++d) gr[d] = -2.0 * (x[d] - 0.5) * u;
    } else if (cfg.problem == PDELAB_PROBLEM_QUADRATIC) {
      const double* q = cfg.quad;
      for (int d = 0; d < 3; ++d) gr[d] = q[1 + d] + 2.0 * q[4 + d] * x[d];
    } else {
      gr[0] = gr[1] = gr[2] = 0.0;
    }
  }
  // Neumann flux j(x) with (b u - A grad u).n = j and outflow data o(x) with
  // -(A grad u).n = o (convectiondiffusionparameter.hh:121-124, 196-208): the fluxes of the
  // analytic solution (A = kappa0 I) for EXP and QUADRATIC, the constants j0 / o0 otherwise,
  // the tabulated boundary value for TABULATED. n is the unit outer normal, bn = b.n at x.
  double j(const double x[3], const double n[3], size_t tabq, double bn) const {
    if (cfg.problem == PDELAB_PROBLEM_TABULATED) return g_tab.empty() ? 0.0 : g_tab[tabq];
    if (cfg.problem == PDELAB_PROBLEM_EXP || cfg.problem == PDELAB_PROBLEM_QUADRATIC) {
      double gr[3];
      grad_exact(x, gr);
      double gn = 0.0;
      for (int d = 0; d < 3; ++d) gn += gr[d] * n[d];
      return bn * exact(x) - cfg.kappa0 * gn;
    }
    return cfg.j0;
  }
  double o(const double x[3], const double n[3], size_t tabq) const {
    if (cfg.problem == PDELAB_PROBLEM_TABULATED) return g_tab.empty() ? 0.0 : g_tab[tabq];
    if (cfg.problem == PDELAB_PROBLEM_EXP || cfg.problem == PDELAB_PROBLEM_QUADRATIC) {
      double gr[3];
      grad_exact(x, gr);
      return -cfg.kappa0 * (gr[0] * n[0] + gr[1] * n[1] + gr[2] * n[2]);
    }
    return cfg.o0;
  }
  // g(x) (ConvectionDiffusionModelProblem::g); tabq = index into g_tab
  double g(const double x[3], size_t tabq) const {
    switch (cfg.problem) {
      case PDELAB_PROBLEM_EXP:
      case PDELAB_PROBLEM_QUADRATIC: return exact(x);
      case PDELAB_PROBLEM_CONSTANT:
      case PDELAB_PROBLEM_GAUSSIANS: return cfg.g0;
      case PDELAB_PROBLEM_TABULATED: return g_tab.empty() ? 0.0 : g_tab[tabq];
      default: return 0.0;
    }
  }
};

// Index of the first boundary face of each side in the side-major boundary-face
// numbering used for tabulated g (pdelab_gpu.h).
struct BoundaryNumbering {
  size_t offset[7];
  int n[3];
  explicit BoundaryNumbering(const Grid& g) {
    for (int d = 0; d < 3; ++d) n[d] = g.n[d];
    offset[0] = 0;
    for (int s = 0; s < 6; ++s) {
      int d = s / 2;
      int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      offset[s + 1] = offset[s] + (size_t)n[t1] * n[t2];
    }
  }
  size_t face(int side, const int ijk[3]) const {
    int d = side / 2;
    int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    return offset[side] + (size_t)ijk[t1] + (size_t)n[t1] * ijk[t2];
  }
};

// ---------------------------------------------------------------------------
// Basis tables at the quadrature points (the role of LocalBasisCache,
// R:dune/pdelab/finiteelement/localbasiscache.hh:17-72): values and reference
// gradients of the Q_k basis (qkdglagrange.hh:142-200, multiindex :118-128).
struct Tables {
  int k, n1, nloc, m;  // degree, 1D size, local size, points per direction
  Rule1D rule;
  // volume: m^3 points, x fastest
  std::vector<double> vw;      // weight
  std::vector<double> vxi;     // [q][3]
  std::vector<double> vphi;    // [q][nloc]
  std::vector<double> vgrad;   // [q][nloc][3]
  // faces: side s (0..5) with m^2 points each; tangential t1 < t2, t1 fastest
  std::vector<double> fw;      // [q]
  std::vector<double> fxi;     // [s][q][3]
  std::vector<double> fphi;    // [s][q][nloc]
  std::vector<double> fgrad;   // [s][q][nloc][3]

  static void eval(int k, const double xi[3], double* phi, double* grad) {
    int n1 = k + 1, nloc = n1 * n1 * n1;
    for (int i = 0; i < nloc; ++i) {
      int a[3] = {i % n1, (i / n1) % n1, i / (n1 * n1)};
      double v = 1.0;
      for (int j = 0; j < 3; ++j) v *= lagrange_p(k, a[j], xi[j]);
      phi[i] = v;
      for (int j = 0; j < 3; ++j) {
        double g = lagrange_dp(k, a[j], xi[j]);
        for (int l = 0; l < 3; ++l)
          if (l != j) g *= lagrange_p(k, a[l], xi[l]);
        grad[3 * i + j] = g;
      }
    }
  }
  Tables(int k_, int order) : k(k_), n1(k_ + 1), nloc((k_ + 1) * (k_ + 1) * (k_ + 1)) {
    rule = gauss_legendre(order);
    m = (int)rule.x.size();
    int nq = m * m * m;
    vw.resize(nq); vxi.resize(3 * nq); vphi.resize((size_t)nq * nloc); vgrad.resize((size_t)nq * nloc * 3);
    for (int q = 0; q < nq; ++q) {
      int qa[3] = {q % m, (q / m) % m, q / (m * m)};
      double xi[3];
      double w = 1.0;
      for (int d = 0; d < 3; ++d) { xi[d] = rule.x[qa[d]]; w *= rule.w[qa[d]]; }
      vw[q] = w;
      for (int d = 0; d < 3; ++d) vxi[3 * q + d] = xi[d];
      eval(k, xi, &vphi[(size_t)q * nloc], &vgrad[(size_t)q * nloc * 3]);
    }
    int nf = m * m;
    fw.resize(nf); fxi.resize(6 * nf * 3); fphi.resize((size_t)6 * nf * nloc); fgrad.resize((size_t)6 * nf * nloc * 3);
    for (int q = 0; q < nf; ++q) fw[q] = rule.w[q % m] * rule.w[q / m];
    for (int s = 0; s < 6; ++s) {
      int d = s / 2;
      int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      for (int q = 0; q < nf; ++q) {
        double xi[3];
        xi[d] = (s % 2) ? 1.0 : 0.0;
        xi[t1] = rule.x[q % m];
        xi[t2] = rule.x[q / m];
        size_t o = (size_t)s * nf + q;
        for (int e = 0; e < 3; ++e) fxi[3 * o + e] = xi[e];
        eval(k, xi, &fphi[o * nloc], &fgrad[o * nloc * 3]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// ConvectionDiffusionDG local operator,
// R:dune/pdelab/localoperator/convectiondiffusiondg.hh:54-1134, restated for the
// axis-aligned hexahedral grid. RF is double, or Dual for the kappa gradient.
template <class RF>
struct LocalOperator {
  const Grid& grid;
  const Param& param;
  const Tables& tab;
  int method, weights, intorderadd;
  double alpha, theta;
  size_t seed = (size_t)-1;  // Dual: cell whose log(kappa) is differentiated

  LocalOperator(const Grid& g, const Param& p, const Tables& t)
      : grid(g), param(p), tab(t), method(p.cfg.method), weights(p.cfg.weights),
        intorderadd(p.cfg.intorderadd), alpha(p.cfg.alpha) {
    // convectiondiffusiondg.hh:99-101
    theta = 1.0;
    if (method == PDELAB_SIPG) theta = -1.0;
    if (method == PDELAB_IIPG) theta = 0.0;
  }

  // entry d of the diagonal tensor A(cell, center) (kappa_c for A = kappa_c I). The faces are
  // normal to a coordinate direction d, so A n = a_d n there and n.A n = a_d: the reference's
  // A_s.mv(n_F, An_F_s), delta_s = An_F_s . n_F and An_F_s . grad u (convectiondiffusiondg.hh:316-333,
  // 418-424) reduce to the scalar a_d below.
  RF kap(size_t c, int d) const;

  // alpha_volume, convectiondiffusiondg.hh:106-188 (also jacobian_apply_volume :192-195)
  void alpha_volume(size_t c, const double* x, RF* r) const {
    const int nloc = tab.nloc;
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const double det = ext[0] * ext[1] * ext[2];
    const double jit[3] = {1.0 / ext[0], 1.0 / ext[1], 1.0 / ext[2]};
    const RF A[3] = {kap(c, 0), kap(c, 1), kap(c, 2)};
    const double cc = param.cfg.c;
    std::vector<double> gradphi(3 * nloc);
    const int nq = tab.m * tab.m * tab.m;
    for (int q = 0; q < nq; ++q) {
      const double* b = param.b_at(c * (size_t)nq + q);  // :178
      const double* phi = &tab.vphi[(size_t)q * nloc];
      const double* js = &tab.vgrad[(size_t)q * nloc * 3];
      double u = 0.0;
      for (int i = 0; i < nloc; ++i) u += x[i] * phi[i];
      for (int i = 0; i < nloc; ++i)
        for (int d = 0; d < 3; ++d) gradphi[3 * i + d] = jit[d] * js[3 * i + d];
      double gradu[3] = {0, 0, 0};
      for (int i = 0; i < nloc; ++i)
        for (int d = 0; d < 3; ++d) gradu[d] += x[i] * gradphi[3 * i + d];
      RF Agradu[3] = {A[0] * gradu[0], A[1] * gradu[1], A[2] * gradu[2]};
      const double factor = tab.vw[q] * det;
      for (int i = 0; i < nloc; ++i) {
        const double* gp = &gradphi[3 * i];
        RF ag = Agradu[0] * gp[0] + Agradu[1] * gp[1] + Agradu[2] * gp[2];
        double bg = b[0] * gp[0] + b[1] * gp[1] + b[2] * gp[2];
        r[i] += (ag - u * bg + cc * u * phi[i]) * factor;
      }
    }
  }

  // jacobian_volume, convectiondiffusiondg.hh:199-266; mat is nloc x nloc, row i (test), col j
  void jacobian_volume(size_t c, RF* mat) const {
    const int nloc = tab.nloc;
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const double det = ext[0] * ext[1] * ext[2];
    const double jit[3] = {1.0 / ext[0], 1.0 / ext[1], 1.0 / ext[2]};
    const RF A[3] = {kap(c, 0), kap(c, 1), kap(c, 2)};
    const double cc = param.cfg.c;
    std::vector<double> gradphi(3 * nloc);
    const int nq = tab.m * tab.m * tab.m;
    for (int q = 0; q < nq; ++q) {
      const double* b = param.b_at(c * (size_t)nq + q);  // :255
      const double* phi = &tab.vphi[(size_t)q * nloc];
      const double* js = &tab.vgrad[(size_t)q * nloc * 3];
      for (int i = 0; i < nloc; ++i)
        for (int d = 0; d < 3; ++d) gradphi[3 * i + d] = jit[d] * js[3 * i + d];
      const double factor = tab.vw[q] * det;
      for (int j = 0; j < nloc; ++j) {
        const double* gj = &gradphi[3 * j];
        RF Agj[3] = {A[0] * gj[0], A[1] * gj[1], A[2] * gj[2]};
        for (int i = 0; i < nloc; ++i) {
          const double* gi = &gradphi[3 * i];
          RF ag = Agj[0] * gi[0] + Agj[1] * gi[1] + Agj[2] * gi[2];
          double bg = b[0] * gi[0] + b[1] * gi[1] + b[2] * gi[2];
          mat[(size_t)i * nloc + j] += (ag - phi[j] * bg + cc * phi[j] * phi[i]) * factor;
        }
      }
    }
  }

  // Common face data of an interior intersection seen from the inside cell `cin`
  // through its face `fin`; the outside cell sees it through fin^1.
  struct FaceCoeffs {
    RF omega_s, omega_n, harmonic, penalty, As, An;
    double nF[3], factor_area, hF;
  };
  FaceCoeffs skeleton_coeffs(size_t cin, int fin, size_t cout) const {
    FaceCoeffs fc;
    double lo[3], ext_in[3], ext_out[3];
    grid.cell_box(cin, lo, ext_in);
    grid.cell_box(cout, lo, ext_out);
    const int d = fin / 2;
    const double vol_in = ext_in[0] * ext_in[1] * ext_in[2];
    const double vol_out = ext_out[0] * ext_out[1] * ext_out[2];
    const double area = vol_in / ext_in[d];
    fc.factor_area = area;
    // face diameter h_F = min(V_s, V_n)/|F|, convectiondiffusiondg.hh:313
    fc.hF = std::min(vol_in, vol_out) / area;
    for (int e = 0; e < 3; ++e) fc.nF[e] = 0.0;
    fc.nF[d] = (fin % 2) ? 1.0 : -1.0;
    fc.As = kap(cin, d);
    fc.An = kap(cout, d);
    // A n, delta = n.A n (A = kappa I): :316-333
    RF An_s[3], An_n[3];
    for (int e = 0; e < 3; ++e) { An_s[e] = fc.As * fc.nF[e]; An_n[e] = fc.An * fc.nF[e]; }
    if (weights == PDELAB_WEIGHTS_ON) {
      RF delta_s = An_s[0] * fc.nF[0] + An_s[1] * fc.nF[1] + An_s[2] * fc.nF[2];
      RF delta_n = An_n[0] * fc.nF[0] + An_n[1] * fc.nF[1] + An_n[2] * fc.nF[2];
      fc.omega_s = delta_n / (delta_s + delta_n + 1e-20);
      fc.omega_n = delta_s / (delta_s + delta_n + 1e-20);
      fc.harmonic = 2.0 * delta_s * delta_n / (delta_s + delta_n + 1e-20);
    } else {
      fc.omega_s = RF(0.5);
      fc.omega_n = RF(0.5);
      fc.harmonic = RF(1.0);
    }
    const int degree = tab.k;
    // penalty, :346
    fc.penalty = (alpha / fc.hF) * fc.harmonic * (double)(degree * (degree + 3 - 1));
    return fc;
  }

  // alpha_skeleton, convectiondiffusiondg.hh:271-471
  void alpha_skeleton(size_t cin, int fin, size_t cout, const double* xs, const double* xn, RF* rs, RF* rn) const {
    const int nloc = tab.nloc;
    const int fout = fin ^ 1;
    FaceCoeffs fc = skeleton_coeffs(cin, fin, cout);
    double lo[3], ei[3], eo[3];
    grid.cell_box(cin, lo, ei);
    grid.cell_box(cout, lo, eo);
    const double jis[3] = {1.0 / ei[0], 1.0 / ei[1], 1.0 / ei[2]};
    const double jin[3] = {1.0 / eo[0], 1.0 / eo[1], 1.0 / eo[2]};
    const double* n = fc.nF;
    const int nf = tab.m * tab.m;
    int cijk[3];
    grid.coords(cin, cijk);
    std::vector<double> tg_s(3 * nloc), tg_n(3 * nloc);
    for (int q = 0; q < nf; ++q) {
      size_t os = (size_t)fin * nf + q, on = (size_t)fout * nf + q;
      const double* phi_s = &tab.fphi[os * nloc];
      const double* phi_n = &tab.fphi[on * nloc];
      const double* gs = &tab.fgrad[os * nloc * 3];
      const double* gn = &tab.fgrad[on * nloc * 3];
      double u_s = 0.0, u_n = 0.0;
      for (int i = 0; i < nloc; ++i) u_s += xs[i] * phi_s[i];
      for (int i = 0; i < nloc; ++i) u_n += xn[i] * phi_n[i];
      for (int i = 0; i < nloc; ++i)
        for (int e = 0; e < 3; ++e) { tg_s[3 * i + e] = jis[e] * gs[3 * i + e]; tg_n[3 * i + e] = jin[e] * gn[3 * i + e]; }
      double gu_s[3] = {0, 0, 0}, gu_n[3] = {0, 0, 0};
      for (int i = 0; i < nloc; ++i)
        for (int e = 0; e < 3; ++e) { gu_s[e] += xs[i] * tg_s[3 * i + e]; gu_n[e] += xn[i] * tg_n[3 * i + e]; }
      // upwinding, :425-438
      const double normalflux = param.b_normal(cijk, fin, q);
      double up_s, up_n;
      if (normalflux >= 0.0) { up_s = 1.0; up_n = 0.0; } else { up_s = 0.0; up_n = 1.0; }
      const double factor = tab.fw[q] * fc.factor_area;
      // A n . grad u  (A = kappa I)
      RF Angu_s = fc.As * (n[0] * gu_s[0] + n[1] * gu_s[1] + n[2] * gu_s[2]);
      RF Angu_n = fc.An * (n[0] * gu_n[0] + n[1] * gu_n[1] + n[2] * gu_n[2]);
      double term1 = (up_s * u_s + up_n * u_n) * normalflux * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += term1 * phi_s[i];
      for (int i = 0; i < nloc; ++i) rn[i] += -term1 * phi_n[i];
      RF term2 = -(fc.omega_s * Angu_s + fc.omega_n * Angu_n) * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += term2 * phi_s[i];
      for (int i = 0; i < nloc; ++i) rn[i] += -term2 * phi_n[i];
      double term3 = (u_s - u_n) * factor;
      for (int i = 0; i < nloc; ++i) {
        RF Angp = fc.As * (n[0] * tg_s[3 * i] + n[1] * tg_s[3 * i + 1] + n[2] * tg_s[3 * i + 2]);
        rs[i] += term3 * theta * fc.omega_s * Angp;
      }
      for (int i = 0; i < nloc; ++i) {
        RF Angp = fc.An * (n[0] * tg_n[3 * i] + n[1] * tg_n[3 * i + 1] + n[2] * tg_n[3 * i + 2]);
        rn[i] += term3 * theta * fc.omega_n * Angp;
      }
      RF term4 = fc.penalty * (u_s - u_n) * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += term4 * phi_s[i];
      for (int i = 0; i < nloc; ++i) rn[i] += -term4 * phi_n[i];
    }
  }

  // jacobian_skeleton, convectiondiffusiondg.hh:484-669; blocks row-major nloc x nloc
  void jacobian_skeleton(size_t cin, int fin, size_t cout, RF* mss, RF* msn, RF* mns, RF* mnn) const {
    const int nloc = tab.nloc;
    const int fout = fin ^ 1;
    FaceCoeffs fc = skeleton_coeffs(cin, fin, cout);
    double lo[3], ei[3], eo[3];
    grid.cell_box(cin, lo, ei);
    grid.cell_box(cout, lo, eo);
    const double jis[3] = {1.0 / ei[0], 1.0 / ei[1], 1.0 / ei[2]};
    const double jin[3] = {1.0 / eo[0], 1.0 / eo[1], 1.0 / eo[2]};
    const double* n = fc.nF;
    const int nf = tab.m * tab.m;
    int cijk[3];
    grid.coords(cin, cijk);
    std::vector<double> ngs(nloc), ngn(nloc);  // n . grad phi (transformed)
    for (int q = 0; q < nf; ++q) {
      size_t os = (size_t)fin * nf + q, on = (size_t)fout * nf + q;
      const double* ps = &tab.fphi[os * nloc];
      const double* pn = &tab.fphi[on * nloc];
      const double* gs = &tab.fgrad[os * nloc * 3];
      const double* gn = &tab.fgrad[on * nloc * 3];
      for (int i = 0; i < nloc; ++i) {
        ngs[i] = n[0] * jis[0] * gs[3 * i] + n[1] * jis[1] * gs[3 * i + 1] + n[2] * jis[2] * gs[3 * i + 2];
        ngn[i] = n[0] * jin[0] * gn[3 * i] + n[1] * jin[1] * gn[3 * i + 1] + n[2] * jin[2] * gn[3 * i + 2];
      }
      const double normalflux = param.b_normal(cijk, fin, q);  // :613
      double up_s, up_n;
      if (normalflux >= 0.0) { up_s = 1.0; up_n = 0.0; } else { up_s = 0.0; up_n = 1.0; }
      const double factor = tab.fw[q] * fc.factor_area;
      RF ipfactor = fc.penalty * factor;
      for (int j = 0; j < nloc; ++j) {
        RF temp1 = -(fc.As * ngs[j]) * fc.omega_s * factor;
        for (int i = 0; i < nloc; ++i) {
          RF v = up_s * ps[j] * normalflux * factor * ps[i] + temp1 * ps[i] +
                 ps[j] * factor * theta * fc.omega_s * (fc.As * ngs[i]) + ps[j] * ipfactor * ps[i];
          mss[(size_t)i * nloc + j] += v;
        }
      }
      for (int j = 0; j < nloc; ++j) {
        RF temp1 = -(fc.An * ngn[j]) * fc.omega_n * factor;
        for (int i = 0; i < nloc; ++i) {
          RF v = up_n * pn[j] * normalflux * factor * ps[i] + temp1 * ps[i] -
                 pn[j] * factor * theta * fc.omega_s * (fc.As * ngs[i]) - pn[j] * ipfactor * ps[i];
          msn[(size_t)i * nloc + j] += v;
        }
      }
      for (int j = 0; j < nloc; ++j) {
        RF temp1 = -(fc.As * ngs[j]) * fc.omega_s * factor;
        for (int i = 0; i < nloc; ++i) {
          RF v = -up_s * ps[j] * normalflux * factor * pn[i] - temp1 * pn[i] +
                 ps[j] * factor * theta * fc.omega_n * (fc.An * ngn[i]) - ps[j] * ipfactor * pn[i];
          mns[(size_t)i * nloc + j] += v;
        }
      }
      for (int j = 0; j < nloc; ++j) {
        RF temp1 = -(fc.An * ngn[j]) * fc.omega_n * factor;
        for (int i = 0; i < nloc; ++i) {
          RF v = -up_n * pn[j] * normalflux * factor * pn[i] - temp1 * pn[i] -
                 pn[j] * factor * theta * fc.omega_n * (fc.An * ngn[i]) + pn[j] * ipfactor * pn[i];
          mnn[(size_t)i * nloc + j] += v;
        }
      }
    }
  }

  struct BoundaryCoeffs {
    RF As, harmonic, penalty;
    double nF[3], factor_area, hF;
  };
  BoundaryCoeffs boundary_coeffs(size_t c, int side) const {
    BoundaryCoeffs bc;
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const int d = side / 2;
    const double vol = ext[0] * ext[1] * ext[2];
    const double area = vol / ext[d];
    bc.factor_area = area;
    bc.hF = vol / area;  // :717
    for (int e = 0; e < 3; ++e) bc.nF[e] = 0.0;
    bc.nF[d] = (side % 2) ? 1.0 : -1.0;
    bc.As = kap(c, d);
    RF An[3] = {bc.As * bc.nF[0], bc.As * bc.nF[1], bc.As * bc.nF[2]};
    if (weights == PDELAB_WEIGHTS_ON)
      bc.harmonic = An[0] * bc.nF[0] + An[1] * bc.nF[1] + An[2] * bc.nF[2];
    else
      bc.harmonic = RF(1.0);
    const int degree = tab.k;
    bc.penalty = (alpha / bc.hF) * bc.harmonic * (double)(degree * (degree + 3 - 1));
    return bc;
  }

  // residual_boundary_integral, convectiondiffusiondg.hh:684-879
  // (alpha_boundary :884-889 with jac_apply=false, jacobian_apply_boundary :893-899 with true)
  void boundary_integral(size_t c, int side, const double* xs, RF* rs, bool jac_apply) const {
    const int nloc = tab.nloc;
    BoundaryCoeffs bc = boundary_coeffs(c, side);
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const double ji[3] = {1.0 / ext[0], 1.0 / ext[1], 1.0 / ext[2]};
    const double* n = bc.nF;
    const int bct = param.bctype(side);
    const int nf = tab.m * tab.m;
    BoundaryNumbering bn(grid);
    int ijk[3];
    grid.coords(c, ijk);
    const size_t bface = bn.face(side, ijk);
    std::vector<double> tg(3 * nloc);
    for (int q = 0; q < nf; ++q) {
      if (bct == PDELAB_BC_NONE) continue;
      size_t o = (size_t)side * nf + q;
      const double* phi = &tab.fphi[o * nloc];
      const double factor = tab.fw[q] * bc.factor_area;
      double xg[3];
      const double* xi = &tab.fxi[3 * o];
      for (int e = 0; e < 3; ++e) xg[e] = lo[e] + ext[e] * xi[e];
      const double normalflux = param.b_normal(ijk, side, q);  // :797
      if (bct == PDELAB_BC_NEUMANN) {
        // :776-783, the flux data only
        if (!jac_apply) {
          const double jv = param.j(xg, n, bface * nf + q, normalflux);
          for (int i = 0; i < nloc; ++i) rs[i] += jv * phi[i] * factor;
        }
        continue;
      }
      double u = 0.0;
      for (int i = 0; i < nloc; ++i) u += xs[i] * phi[i];
      if (bct == PDELAB_BC_OUTFLOW) {
        // :794-816
        if (normalflux < -1e-30) throw std::runtime_error("Outflow boundary condition on inflow!");
        double term1 = u * normalflux * factor;
        for (int i = 0; i < nloc; ++i) rs[i] += term1 * phi[i];
        if (!jac_apply) {
          const double ov = param.o(xg, n, bface * nf + q);
          for (int i = 0; i < nloc; ++i) rs[i] += ov * phi[i] * factor;
        }
        continue;
      }
      const double* gr = &tab.fgrad[o * nloc * 3];
      for (int i = 0; i < nloc; ++i)
        for (int e = 0; e < 3; ++e) tg[3 * i + e] = ji[e] * gr[3 * i + e];
      double gu[3] = {0, 0, 0};
      for (int i = 0; i < nloc; ++i)
        for (int e = 0; e < 3; ++e) gu[e] += xs[i] * tg[3 * i + e];
      double g = param.g(xg, bface * nf + q);
      if (jac_apply) g = 0.0;
      double up_s, up_n;
      if (normalflux >= 0.0) { up_s = 1.0; up_n = 0.0; } else { up_s = 0.0; up_n = 1.0; }
      double term1 = (up_s * u + up_n * g) * normalflux * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += term1 * phi[i];
      RF term2 = (bc.As * (n[0] * gu[0] + n[1] * gu[1] + n[2] * gu[2])) * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += -term2 * phi[i];
      double term3 = (u - g) * factor;
      for (int i = 0; i < nloc; ++i) {
        RF Angp = bc.As * (n[0] * tg[3 * i] + n[1] * tg[3 * i + 1] + n[2] * tg[3 * i + 2]);
        rs[i] += term3 * theta * Angp;
      }
      RF term4 = bc.penalty * (u - g) * factor;
      for (int i = 0; i < nloc; ++i) rs[i] += term4 * phi[i];
    }
  }

  // jacobian_boundary, convectiondiffusiondg.hh:902-1044
  void jacobian_boundary(size_t c, int side, RF* mat) const {
    const int nloc = tab.nloc;
    BoundaryCoeffs bc = boundary_coeffs(c, side);
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const double ji[3] = {1.0 / ext[0], 1.0 / ext[1], 1.0 / ext[2]};
    const double* n = bc.nF;
    const int bct = param.bctype(side);
    const int nf = tab.m * tab.m;
    int ijk[3];
    grid.coords(c, ijk);
    std::vector<double> ng(nloc);
    for (int q = 0; q < nf; ++q) {
      if (bct == PDELAB_BC_NONE || bct == PDELAB_BC_NEUMANN) continue;
      size_t o = (size_t)side * nf + q;
      const double* phi = &tab.fphi[o * nloc];
      const double factor = tab.fw[q] * bc.factor_area;
      const double normalflux = param.b_normal(ijk, side, q);  // :995
      if (bct == PDELAB_BC_OUTFLOW) {
        if (normalflux < -1e-30) throw std::runtime_error("Outflow boundary condition on inflow!");
        for (int j = 0; j < nloc; ++j)
          for (int i = 0; i < nloc; ++i) mat[(size_t)i * nloc + j] += phi[j] * normalflux * factor * phi[i];
        continue;
      }
      const double* gr = &tab.fgrad[o * nloc * 3];
      for (int i = 0; i < nloc; ++i)
        ng[i] = n[0] * ji[0] * gr[3 * i] + n[1] * ji[1] * gr[3 * i + 1] + n[2] * ji[2] * gr[3 * i + 2];
      double up_s = normalflux >= 0.0 ? 1.0 : 0.0;
      for (int j = 0; j < nloc; ++j)
        for (int i = 0; i < nloc; ++i) {
          RF v = up_s * phi[j] * normalflux * factor * phi[i] - (bc.As * ng[j]) * factor * phi[i] +
                 phi[j] * factor * theta * (bc.As * ng[i]) + bc.penalty * phi[j] * phi[i] * factor;
          mat[(size_t)i * nloc + j] += v;
        }
    }
  }

  // lambda_volume, convectiondiffusiondg.hh:1048-1075
  void lambda_volume(size_t c, RF* r) const {
    const int nloc = tab.nloc;
    double lo[3], ext[3];
    grid.cell_box(c, lo, ext);
    const double det = ext[0] * ext[1] * ext[2];
    const int nq = tab.m * tab.m * tab.m;
    for (int q = 0; q < nq; ++q) {
      const double* phi = &tab.vphi[(size_t)q * nloc];
      double xg[3];
      for (int e = 0; e < 3; ++e) xg[e] = lo[e] + ext[e] * tab.vxi[3 * q + e];
      double f = param.f(c, xg, c * (size_t)nq + q);
      const double factor = tab.vw[q] * det;
      for (int i = 0; i < nloc; ++i) r[i] += -f * phi[i] * factor;
    }
  }
};

template <>
inline double LocalOperator<double>::kap(size_t c, int d) const { return param.kappa_d(c, d); }
template <>
inline Dual LocalOperator<Dual>::kap(size_t c, int d) const {
  double k = param.kappa_d(c, d);
  return Dual(k, c == seed ? k : 0.0);  // d kappa / d log kappa = kappa
}

// ---------------------------------------------------------------------------
// ScalarL2 mass operator (scaling 1), R:dune/pdelab/localoperator/l2.hh:45-83
inline void mass_volume(const Grid& grid, const Tables& tab, size_t c, const double* x, double* r) {
  const int nloc = tab.nloc;
  double lo[3], ext[3];
  grid.cell_box(c, lo, ext);
  const double det = ext[0] * ext[1] * ext[2];
  const int nq = tab.m * tab.m * tab.m;
  for (int q = 0; q < nq; ++q) {
    const double* phi = &tab.vphi[(size_t)q * nloc];
    double u = 0.0;
    for (int i = 0; i < nloc; ++i) u += x[i] * phi[i];
    const double factor = tab.vw[q] * det;
    for (int i = 0; i < nloc; ++i) r[i] += u * phi[i] * factor;
  }
}

}  // namespace oracle

#endif
