// oracle.cc -- THE ORACLE (test infrastructure, not product code).
//
// Global assembly engines, dune-istl style Krylov solvers and the
// StationaryLinearProblemSolver driver on top of the local operator restated in
// pdelab_oracle.hh, exported through a small C API for the Python tests and for
// bench.py's CPU baseline. See pdelab_oracle.hh for what pins each part.
#include <chrono>
#include <cstring>
#include <memory>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "pdelab_oracle.hh"

namespace oracle {

struct Oracle {
  pdelab_config cfg;
  Grid grid;
  Param param;
  Tables tab;      // operator quadrature (order 2k + intorderadd)
  LocalOperator<double> lop;
  size_t nloc, N;
  std::string err;

  static Param make_param(const pdelab_config& c) {
    Param p;
    p.cfg = c;
    return p;
  }

  Oracle(const pdelab_config& c, const double* kappa)
      : cfg(c), grid(c), param(make_param(c)), tab(c.degree, 2 * c.degree + c.intorderadd), lop(grid, param, tab) {
    param.aniso = c.kappa_per_cell == 2;
    param.kappa.assign(grid.ncells * (param.aniso ? 3 : 1), c.kappa0);
    if (c.kappa_per_cell && kappa) param.kappa.assign(kappa, kappa + grid.ncells * (param.aniso ? 3 : 1));
    nloc = tab.nloc;
    N = grid.ncells * nloc;
  }

  // ---- DefaultAssembler traversal, R:dune/pdelab/gridoperator/default/assembler.hh:84-279
  // residual engine (R:dune/pdelab/gridoperator/default/residualengine.hh): r += R(x)
  void residual(const double* x, double* r) const {
    std::vector<double> rl(nloc), rn(nloc);
    for (size_t c = 0; c < grid.ncells; ++c) {
      std::fill(rl.begin(), rl.end(), 0.0);
      const double* xc = x + c * nloc;
      lop.lambda_volume(c, rl.data());       // assembleVVolume
      lop.alpha_volume(c, xc, rl.data());    // assembleUVVolume
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          if (c > cn) {  // visit_face = ids > idn, assembler.hh:181
            std::fill(rn.begin(), rn.end(), 0.0);
            lop.alpha_skeleton(c, f, cn, xc, x + cn * nloc, rl.data(), rn.data());
            for (size_t i = 0; i < nloc; ++i) r[cn * nloc + i] += rn[i];
          }
        } else {
          lop.boundary_integral(c, f, xc, rl.data(), false);
        }
      }
      for (size_t i = 0; i < nloc; ++i) r[c * nloc + i] += rl[i];
    }
  }

  // jacobian-apply engine (default/jacobianapplyengine.hh): y += J z
  void jacobian_apply(const double* z, double* y) const {
    std::vector<double> yl(nloc), yn(nloc);
    for (size_t c = 0; c < grid.ncells; ++c) {
      std::fill(yl.begin(), yl.end(), 0.0);
      const double* zc = z + c * nloc;
      lop.alpha_volume(c, zc, yl.data());
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          if (c > cn) {
            std::fill(yn.begin(), yn.end(), 0.0);
            lop.alpha_skeleton(c, f, cn, zc, z + cn * nloc, yl.data(), yn.data());
            for (size_t i = 0; i < nloc; ++i) y[cn * nloc + i] += yn[i];
          }
        } else {
          lop.boundary_integral(c, f, zc, yl.data(), true);
        }
      }
      for (size_t i = 0; i < nloc; ++i) y[c * nloc + i] += yl[i];
    }
  }

  // Same operator with two-sided skeleton visits (assembler.hh:181 with
  // require_skeleton_two_sided): every cell only writes its own block, so cells can
  // be processed in parallel. Used for the all-cores CPU baseline.
  void jacobian_apply_twosided(const double* z, double* y, int nthreads) const {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
      std::vector<double> yl(nloc), dummy(nloc);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (long long cc = 0; cc < (long long)grid.ncells; ++cc) {
        size_t c = (size_t)cc;
        std::fill(yl.begin(), yl.end(), 0.0);
        const double* zc = z + c * nloc;
        lop.alpha_volume(c, zc, yl.data());
        for (int f = 0; f < 6; ++f) {
          size_t cn;
          if (grid.neighbor(c, f, cn))
            lop.alpha_skeleton(c, f, cn, zc, z + cn * nloc, yl.data(), dummy.data());
          else
            lop.boundary_integral(c, f, zc, yl.data(), true);
        }
        for (size_t i = 0; i < nloc; ++i) y[c * nloc + i] += yl[i];
      }
    }
  }

  // r += R(x) with two-sided skeleton visits, parallel over cells: the residual engine's
  // terms of residual() (lambda_volume, alpha_volume, alpha_skeleton, boundary with the
  // data terms), each cell writing only its own block. Same operator as residual() up to
  // rounding (pinned against it); it makes full-size checks affordable on the host.
  void residual_twosided(const double* x, double* r, int nthreads) const {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
      std::vector<double> rl(nloc), dummy(nloc);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (long long cc = 0; cc < (long long)grid.ncells; ++cc) {
        const size_t c = (size_t)cc;
        std::fill(rl.begin(), rl.end(), 0.0);
        const double* xc = x + c * nloc;
        lop.lambda_volume(c, rl.data());
        lop.alpha_volume(c, xc, rl.data());
        for (int f = 0; f < 6; ++f) {
          size_t cn;
          if (grid.neighbor(c, f, cn))
            lop.alpha_skeleton(c, f, cn, xc, x + cn * nloc, rl.data(), dummy.data());
          else
            lop.boundary_integral(c, f, xc, rl.data(), false);
        }
        for (size_t i = 0; i < nloc; ++i) r[c * nloc + i] += rl[i];
      }
    }
  }

  // y += J^T z from the same local Jacobian blocks as jacobian_apply_transpose (each face
  // taken from its higher-index cell, assembler.hh:181), but gathered per column cell:
  // cell c collects (J_cc)^T z_c and (J_{n,c})^T z_n of its face neighbours n, so cells run
  // in parallel. Equal to jacobian_apply_transpose up to the summation order.
  void jacobian_apply_transpose_twosided(const double* z, double* y, int nthreads) const {
    const size_t n2 = nloc * nloc;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
      std::vector<double> m(n2), mss(n2), msn(n2), mns(n2), mnn(n2), yl(nloc);
      auto addT = [&](const std::vector<double>& A, const double* x) {
        for (size_t i = 0; i < nloc; ++i)
          for (size_t j = 0; j < nloc; ++j) yl[j] += A[i * nloc + j] * x[i];
      };
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (long long cc = 0; cc < (long long)grid.ncells; ++cc) {
        const size_t c = (size_t)cc;
        std::fill(yl.begin(), yl.end(), 0.0);
        std::fill(m.begin(), m.end(), 0.0);
        lop.jacobian_volume(c, m.data());
        addT(m, z + c * nloc);
        for (int f = 0; f < 6; ++f) {
          size_t cn;
          if (grid.neighbor(c, f, cn)) {
            std::fill(mss.begin(), mss.end(), 0.0); std::fill(msn.begin(), msn.end(), 0.0);
            std::fill(mns.begin(), mns.end(), 0.0); std::fill(mnn.begin(), mnn.end(), 0.0);
            if (c > cn) {  // c inside: dR_c/du_c = mss, dR_cn/du_c = mns
              lop.jacobian_skeleton(c, f, cn, mss.data(), msn.data(), mns.data(), mnn.data());
              addT(mss, z + c * nloc);
              addT(mns, z + cn * nloc);
            } else {       // cn inside: dR_c/du_c = mnn, dR_cn/du_c = msn
              lop.jacobian_skeleton(cn, f ^ 1, c, mss.data(), msn.data(), mns.data(), mnn.data());
              addT(mnn, z + c * nloc);
              addT(msn, z + cn * nloc);
            }
          } else {
            std::fill(m.begin(), m.end(), 0.0);
            lop.jacobian_boundary(c, f, m.data());
            addT(m, z + c * nloc);
          }
        }
        for (size_t i = 0; i < nloc; ++i) y[c * nloc + i] += yl[i];
      }
    }
  }

  // y += J^T z from the local Jacobian blocks (jacobian engine traversal)
  void jacobian_apply_transpose(const double* z, double* y) const {
    const size_t n2 = nloc * nloc;
    std::vector<double> m(n2), mss(n2), msn(n2), mns(n2), mnn(n2);
    auto addT = [&](const std::vector<double>& A, const double* x, double* out) {
      for (size_t i = 0; i < nloc; ++i)
        for (size_t j = 0; j < nloc; ++j) out[j] += A[i * nloc + j] * x[i];
    };
    for (size_t c = 0; c < grid.ncells; ++c) {
      std::fill(m.begin(), m.end(), 0.0);
      lop.jacobian_volume(c, m.data());
      addT(m, z + c * nloc, y + c * nloc);
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          if (c > cn) {
            std::fill(mss.begin(), mss.end(), 0.0); std::fill(msn.begin(), msn.end(), 0.0);
            std::fill(mns.begin(), mns.end(), 0.0); std::fill(mnn.begin(), mnn.end(), 0.0);
            lop.jacobian_skeleton(c, f, cn, mss.data(), msn.data(), mns.data(), mnn.data());
            addT(mss, z + c * nloc, y + c * nloc);
            addT(mns, z + cn * nloc, y + c * nloc);
            addT(msn, z + c * nloc, y + cn * nloc);
            addT(mnn, z + cn * nloc, y + cn * nloc);
          }
        } else {
          std::fill(m.begin(), m.end(), 0.0);
          lop.jacobian_boundary(c, f, m.data());
          addT(m, z + c * nloc, y + c * nloc);
        }
      }
    }
  }

  // Block matrix (GridOperator::jacobian): blocks[c][slot], slot 0 = (c,c),
  // slot 1+f = (c, neighbour across f); cols[c][slot] = column cell or -1.
  void assemble(double* blocks, long long* cols) const {
    const size_t n2 = nloc * nloc;
    std::memset(blocks, 0, sizeof(double) * grid.ncells * 7 * n2);
    for (size_t c = 0; c < grid.ncells; ++c) {
      cols[c * 7] = (long long)c;
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        cols[c * 7 + 1 + f] = grid.neighbor(c, f, cn) ? (long long)cn : -1;
      }
    }
    std::vector<double> mss(n2), msn(n2), mns(n2), mnn(n2);
    for (size_t c = 0; c < grid.ncells; ++c) {
      lop.jacobian_volume(c, blocks + c * 7 * n2);
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          if (c > cn) {
            std::fill(mss.begin(), mss.end(), 0.0); std::fill(msn.begin(), msn.end(), 0.0);
            std::fill(mns.begin(), mns.end(), 0.0); std::fill(mnn.begin(), mnn.end(), 0.0);
            lop.jacobian_skeleton(c, f, cn, mss.data(), msn.data(), mns.data(), mnn.data());
            double* bss = blocks + c * 7 * n2;
            double* bsn = blocks + (c * 7 + 1 + f) * n2;
            double* bns = blocks + (cn * 7 + 1 + (f ^ 1)) * n2;
            double* bnn = blocks + cn * 7 * n2;
            for (size_t i = 0; i < n2; ++i) { bss[i] += mss[i]; bsn[i] += msn[i]; bns[i] += mns[i]; bnn[i] += mnn[i]; }
          }
        } else {
          lop.jacobian_boundary(c, f, blocks + c * 7 * n2);
        }
      }
    }
  }

  // PointDiagonalLocalOperatorWrapper residual (two-sided skeleton, ss-block
  // diagonal only), R:dune/pdelab/localoperator/pointdiagonalwrapper.hh:51-131
  void point_diagonal(double* d) const {
    const size_t n2 = nloc * nloc;
    std::vector<double> m(n2), mss(n2), o1(n2), o2(n2), o3(n2);
    for (size_t c = 0; c < grid.ncells; ++c) {
      std::fill(m.begin(), m.end(), 0.0);
      lop.jacobian_volume(c, m.data());
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          std::fill(mss.begin(), mss.end(), 0.0);
          std::fill(o1.begin(), o1.end(), 0.0); std::fill(o2.begin(), o2.end(), 0.0); std::fill(o3.begin(), o3.end(), 0.0);
          lop.jacobian_skeleton(c, f, cn, mss.data(), o1.data(), o2.data(), o3.data());
          for (size_t i = 0; i < nloc; ++i) m[i * nloc + i] += mss[i * nloc + i];
        } else {
          lop.jacobian_boundary(c, f, m.data());
        }
      }
      for (size_t i = 0; i < nloc; ++i) d[c * nloc + i] += m[i * nloc + i];
    }
  }

  // BlockDiagonalLocalOperatorWrapper jacobian: jacobian_volume, both sides of every
  // interior face keeping the ss block, and the boundary faces,
  // R:dune/pdelab/localoperator/blockdiagonalwrapper.hh:79-160. blocks[c][i][j] row-major.
  void block_diagonal(double* blocks) const {
    const size_t n2 = nloc * nloc;
    std::vector<double> mss(n2), o1(n2), o2(n2), o3(n2);
    for (size_t c = 0; c < grid.ncells; ++c) {
      double* m = blocks + c * n2;
      std::fill(m, m + n2, 0.0);
      lop.jacobian_volume(c, m);
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        if (grid.neighbor(c, f, cn)) {
          std::fill(mss.begin(), mss.end(), 0.0);
          std::fill(o1.begin(), o1.end(), 0.0); std::fill(o2.begin(), o2.end(), 0.0); std::fill(o3.begin(), o3.end(), 0.0);
          lop.jacobian_skeleton(c, f, cn, mss.data(), o1.data(), o2.data(), o3.data());
          for (size_t i = 0; i < n2; ++i) m[i] += mss[i];
        } else {
          lop.jacobian_boundary(c, f, m);
        }
      }
    }
  }

  void mass_apply(const double* z, double* y) const {
    std::vector<double> yl(nloc);
    for (size_t c = 0; c < grid.ncells; ++c) {
      std::fill(yl.begin(), yl.end(), 0.0);
      mass_volume(grid, tab, c, z + c * nloc, yl.data());
      for (size_t i = 0; i < nloc; ++i) y[c * nloc + i] += yl[i];
    }
  }

  // Nodal interpolation of the analytic solution (QkLocalInterpolation,
  // R:dune/pdelab/finiteelement/qkdglagrange.hh:256-281: Lagrange points alpha/k)
  void interpolate(double* x) const {
    const int k = tab.k, n1 = k + 1;
    for (size_t c = 0; c < grid.ncells; ++c) {
      double lo[3], ext[3];
      grid.cell_box(c, lo, ext);
      for (size_t i = 0; i < nloc; ++i) {
        int a[3] = {(int)(i % n1), (int)((i / n1) % n1), (int)(i / (n1 * n1))};
        double xg[3];
        for (int d = 0; d < 3; ++d) xg[d] = lo[d] + ext[d] * ((1.0 * a[d]) / k);
        x[c * nloc + i] = param.exact(xg);
      }
    }
  }

  // integral of (u_h - u)^2 with a quadrature rule of order 10
  double l2_error_sq(const double* x) const {
    Tables t10(tab.k, 10);
    const int nq = t10.m * t10.m * t10.m;
    double err = 0.0;
    for (size_t c = 0; c < grid.ncells; ++c) {
      double lo[3], ext[3];
      grid.cell_box(c, lo, ext);
      const double det = ext[0] * ext[1] * ext[2];
      for (int q = 0; q < nq; ++q) {
        const double* phi = &t10.vphi[(size_t)q * nloc];
        double uh = 0.0;
        for (size_t i = 0; i < nloc; ++i) uh += x[c * nloc + i] * phi[i];
        double xg[3];
        for (int d = 0; d < 3; ++d) xg[d] = lo[d] + ext[d] * t10.vxi[3 * q + d];
        double e = uh - param.exact(xg);
        err += e * e * t10.vw[q] * det;
      }
    }
    return err;
  }

  // g[c] = -lambda^T dR/dlog(kappa_c)(u), by forward-mode differentiation of the
  // local operator (volume of c and all faces of c).
  void grad_logkappa(const double* u, const double* lam, double* g) const {
    LocalOperator<Dual> dl(grid, param, tab);
    std::vector<Dual> rl(nloc), rn(nloc);
    for (size_t c = 0; c < grid.ncells; ++c) {
      dl.seed = c;
      double acc = 0.0;
      std::fill(rl.begin(), rl.end(), Dual());
      dl.alpha_volume(c, u + c * nloc, rl.data());
      for (size_t i = 0; i < nloc; ++i) acc += lam[c * nloc + i] * rl[i].d;
      for (int f = 0; f < 6; ++f) {
        size_t cn;
        std::fill(rl.begin(), rl.end(), Dual());
        if (grid.neighbor(c, f, cn)) {
          std::fill(rn.begin(), rn.end(), Dual());
          size_t cin = std::max(c, cn), cout = std::min(c, cn);
          int fin = (cin == c) ? f : (f ^ 1);
          dl.alpha_skeleton(cin, fin, cout, u + cin * nloc, u + cout * nloc, rl.data(), rn.data());
          for (size_t i = 0; i < nloc; ++i) acc += lam[cin * nloc + i] * rl[i].d + lam[cout * nloc + i] * rn[i].d;
        } else {
          dl.boundary_integral(c, f, u + c * nloc, rl.data(), false);
          for (size_t i = 0; i < nloc; ++i) acc += lam[c * nloc + i] * rl[i].d;
        }
      }
      g[c] = -acc;
    }
  }

  // g[c] = sum_i lambda_{c,i} * integral_c phi_i  (= -lambda^T dR/ds_c)
  void grad_source(const double* lam, double* g) const {
    const int nq = tab.m * tab.m * tab.m;
    for (size_t c = 0; c < grid.ncells; ++c) {
      double lo[3], ext[3];
      grid.cell_box(c, lo, ext);
      const double det = ext[0] * ext[1] * ext[2];
      double acc = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double* phi = &tab.vphi[(size_t)q * nloc];
        for (size_t i = 0; i < nloc; ++i) acc += lam[c * nloc + i] * phi[i] * tab.vw[q] * det;
      }
      g[c] = acc;
    }
  }
};

// ---------------------------------------------------------------------------
// Sequential vector backend (ISTL::BlockVector dot / two_norm / axpy,
// R:dune/pdelab/backend/istl/vector.hh:249-282)
static double dot(size_t n, const double* a, const double* b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static double norm2(size_t n, const double* a) { return std::sqrt(dot(n, a, a)); }

// dune-istl IterativeSolver::Iteration semantics (dune-istl is not part of the reference tree):
// stop when def < reduction * def0 or def < 1e-30, a non-finite defect throws, half steps are
// counted up. Pinned by the CG and BiCGSTAB comparisons with textbook numpy solvers.
struct Iteration {
  double def0 = 0.0, def = 0.0, reduction;
  double i = 0;
  bool converged = false;
  explicit Iteration(double red) : reduction(red) {}
  // returns true when the loop must stop; throws on a non-finite defect
  bool step(double it, double d) {
    if (!std::isfinite(d)) throw std::runtime_error("defect is infinite or NaN");
    if (it == 0) def0 = d;
    def = d;
    i = it;
    converged = (d < def0 * reduction) || (d < 1e-30);
    return converged;
  }
  void finalize(pdelab_linear_solver_result* r, double elapsed) const {
    r->converged = converged ? 1 : 0;
    r->iterations = (unsigned)std::ceil(i);
    r->reduction = def0 > 0 ? def / def0 : 0.0;
    r->conv_rate = std::pow(r->reduction, 1.0 / std::max(1.0, std::ceil(i)));
    r->elapsed = elapsed;
  }
};

using OpFn = std::function<void(const double*, double*)>;  // y = A x (overwrite)

// CGSolver::apply (dune-istl) with the given preconditioner (y = W^{-1} d)
static void cg(size_t n, const OpFn& A, const OpFn& prec, double reduction, int maxit, double* x, double* b,
               pdelab_linear_solver_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  Iteration iter(reduction);
  std::vector<double> p(n), q(n), tmp(n);
  // b = b - A x (applyscaleadd(-1, x, b))
  A(x, tmp.data());
  for (size_t i = 0; i < n; ++i) b[i] -= tmp[i];
  double def = norm2(n, b);
  bool stop = iter.step(0, def);
  if (!stop) {
    prec(b, p.data());
    double rholast = dot(n, p.data(), b);
    for (int it = 1; it <= maxit; ++it) {
      A(p.data(), q.data());
      double alpha = dot(n, p.data(), q.data());
      double lambda = rholast / alpha;
      for (size_t i = 0; i < n; ++i) x[i] += lambda * p[i];
      for (size_t i = 0; i < n; ++i) b[i] -= lambda * q[i];
      def = norm2(n, b);
      if (iter.step(it, def)) break;
      prec(b, q.data());
      double rho = dot(n, q.data(), b);
      double beta = rho / rholast;
      for (size_t i = 0; i < n; ++i) p[i] = beta * p[i] + q[i];
      rholast = rho;
    }
  }
  double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  iter.finalize(res, el);
}

// BiCGSTABSolver::apply (dune-istl) with half-step iteration counting
static void bicgstab(size_t n, const OpFn& A, const OpFn& prec, double reduction, int maxit, double* x, double* b,
                     pdelab_linear_solver_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  const double EPSILON = 1e-80;
  Iteration iter(reduction);
  std::vector<double> p(n), v(n), t(n), y(n), rt(n);
  double* r = b;
  A(x, t.data());
  for (size_t i = 0; i < n; ++i) r[i] -= t[i];
  std::copy(r, r + n, rt.begin());
  double norm = norm2(n, r);
  if (!iter.step(0, norm)) {
    double rho = 1, alpha = 1, omega = 1, rho_new, beta, h;
    for (double it = 0.5; it < maxit; it += .5) {
      rho_new = dot(n, rt.data(), r);
      if (std::fabs(rho) <= EPSILON) throw std::runtime_error("breakdown in BiCGSTAB - rho");
      if (std::fabs(omega) <= EPSILON) throw std::runtime_error("breakdown in BiCGSTAB - omega");
      if (it < 1) {
        std::copy(r, r + n, p.begin());
      } else {
        beta = (norm == 0.0) ? 0.0 : (rho_new / rho) * (alpha / omega);
        for (size_t i = 0; i < n; ++i) p[i] = beta * (p[i] - omega * v[i]) + r[i];
      }
      prec(p.data(), y.data());
      A(y.data(), v.data());
      h = dot(n, rt.data(), v.data());
      if (std::fabs(h) < EPSILON) throw std::runtime_error("abs(h) < EPSILON in BiCGSTAB");
      alpha = (norm == 0.0) ? 0.0 : rho_new / h;
      for (size_t i = 0; i < n; ++i) x[i] += alpha * y[i];
      for (size_t i = 0; i < n; ++i) r[i] -= alpha * v[i];
      norm = norm2(n, r);
      if (iter.step(it, norm)) break;
      it += .5;
      prec(r, y.data());
      A(y.data(), t.data());
      omega = (norm == 0.0) ? 0.0 : dot(n, t.data(), r) / dot(n, t.data(), t.data());
      for (size_t i = 0; i < n; ++i) x[i] += omega * y[i];
      for (size_t i = 0; i < n; ++i) r[i] -= omega * t[i];
      rho = rho_new;
      norm = norm2(n, r);
      if (iter.step(it, norm)) break;
    }
  }
  double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  iter.finalize(res, el);
}

}  // namespace oracle

using oracle::Oracle;

// LU factorisation with partial pivoting of one diagonal block (in place, row-major) and
// its solve: what AssembledBlockJacobiPreconditionerLocalOperator does with Eigen's
// PartialPivLU, R:dune/pdelab/backend/istl/matrixfree/assembledblockjacobipreconditioner.hh:170-227.
static void lu_factor(size_t n, double* A, int* piv) {
  for (size_t k = 0; k < n; ++k) {
    size_t p = k;
    for (size_t i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
    piv[k] = (int)p;
    if (p != k)
      for (size_t j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
    const double d = A[k * n + k];
    for (size_t i = k + 1; i < n; ++i) {
      const double l = A[i * n + k] / d;
      A[i * n + k] = l;
      for (size_t j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
    }
  }
}

static void lu_solve(size_t n, const double* LU, const int* piv, const double* b, double* x) {
  for (size_t i = 0; i < n; ++i) x[i] = b[i];
  for (size_t k = 0; k < n; ++k)
    if ((size_t)piv[k] != k) std::swap(x[k], x[piv[k]]);
  for (size_t i = 0; i < n; ++i) {  // L has ones on the diagonal
    double s = x[i];
    for (size_t j = 0; j < i; ++j) s -= LU[i * n + j] * x[j];
    x[i] = s;
  }
  for (size_t i = n; i-- > 0;) {
    double s = x[i];
    for (size_t j = i + 1; j < n; ++j) s -= LU[i * n + j] * x[j];
    x[i] = s / LU[i * n + i];
  }
}

struct oracle_handle {
  std::unique_ptr<Oracle> o;
  std::vector<double> diag;  // cached point diagonal
  std::vector<double> blu[2];  // LU factors of the diagonal blocks of J and J^T
  std::vector<int> bpiv[2];
  std::vector<double> blocks;
  std::vector<long long> cols;
  int threads = 1;
  double sor_omega = 1.0;
  // iterative block Jacobi: the diagonal blocks of J (row-major) and the block-solver options
  std::vector<double> bdiag;
  double ibj_reduction = 1e-5, ibj_weight = 1.0;
  int ibj_maxit = 100, ibj_precondition = 1;
  unsigned long long ibj_iterations = 0;  // block-solver iterations of the last application
};

static int guard(oracle_handle* h, const std::function<void()>& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    if (h && h->o) h->o->err = e.what();
    return 5;
  }
}

extern "C" {

oracle_handle* oracle_create(const pdelab_config* cfg, const double* kappa) {
  if (!cfg || cfg->degree < 1 || cfg->degree > 4 || cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return nullptr;
  auto* h = new oracle_handle;
  h->o.reset(new Oracle(*cfg, kappa));
  return h;
}
void oracle_destroy(oracle_handle* h) { delete h; }
const char* oracle_last_error(oracle_handle* h) { return h && h->o ? h->o->err.c_str() : ""; }
size_t oracle_size(oracle_handle* h) { return h->o->N; }
void oracle_set_threads(oracle_handle* h, int n) { h->threads = n; }
void oracle_set_sor_omega(oracle_handle* h, double omega) { h->sor_omega = omega; }
void oracle_set_block_solver(oracle_handle* h, double reduction, int maxit, int precondition, double weight) {
  h->ibj_reduction = reduction;
  h->ibj_maxit = maxit;
  h->ibj_precondition = precondition;
  h->ibj_weight = weight;
}
unsigned long long oracle_block_solver_iterations(oracle_handle* h) { return h->ibj_iterations; }
int oracle_iterative_block_jacobi_apply(oracle_handle* h, const double* d, double* v);
int oracle_block_sor_apply(oracle_handle* h, const double* d, double* v, int inplace);

int oracle_set_kappa(oracle_handle* h, const double* kappa) {
  h->o->param.kappa.assign(kappa, kappa + h->o->grid.ncells * (h->o->param.aniso ? 3 : 1));
  h->diag.clear();
  h->blocks.clear();
  for (int t = 0; t < 2; ++t) { h->blu[t].clear(); h->bpiv[t].clear(); }
  return 0;
}
int oracle_set_cell_source(oracle_handle* h, const double* s) {
  if (s) h->o->param.cell_source.assign(s, s + h->o->grid.ncells);
  else h->o->param.cell_source.clear();
  return 0;
}
int oracle_set_tabulated(oracle_handle* h, const double* fq, const double* gq) {
  auto& o = *h->o;
  const size_t m = o.tab.m;
  oracle::BoundaryNumbering bn(o.grid);
  o.param.f_tab.assign(fq, fq + o.grid.ncells * m * m * m);
  o.param.g_tab.assign(gq, gq + bn.offset[6] * m * m);
  return 0;
}
int oracle_quadrature_points(oracle_handle* h) { return h->o->tab.m; }
// Velocity field b(x) at the quadrature points (layouts of Param::b_vol / b_face); NULL
// b_vol restores the constant b of the configuration.
int oracle_set_velocity(oracle_handle* h, const double* bv, const double* bf) {
  auto& o = *h->o;
  if (bv && !bf) return 1;
  o.param.set_velocity(bv, bf, o.grid.n, o.tab.m, o.grid.ncells);
  h->diag.clear();
  h->blocks.clear();
  h->bdiag.clear();
  for (int t = 0; t < 2; ++t) { h->blu[t].clear(); h->bpiv[t].clear(); }
  return 0;
}

int oracle_residual(oracle_handle* h, const double* x, double* r) {
  return guard(h, [&] { h->o->residual(x, r); });
}
int oracle_jacobian_apply(oracle_handle* h, const double* z, double* y) {
  return guard(h, [&] { h->o->jacobian_apply(z, y); });
}
int oracle_jacobian_apply_twosided(oracle_handle* h, const double* z, double* y, int nthreads) {
  return guard(h, [&] { h->o->jacobian_apply_twosided(z, y, nthreads); });
}
int oracle_residual_twosided(oracle_handle* h, const double* x, double* r, int nthreads) {
  return guard(h, [&] { h->o->residual_twosided(x, r, nthreads); });
}
int oracle_jacobian_apply_transpose_twosided(oracle_handle* h, const double* z, double* y, int nthreads) {
  return guard(h, [&] { h->o->jacobian_apply_transpose_twosided(z, y, nthreads); });
}
int oracle_jacobian_apply_transpose(oracle_handle* h, const double* z, double* y) {
  return guard(h, [&] { h->o->jacobian_apply_transpose(z, y); });
}
int oracle_point_diagonal(oracle_handle* h, double* d) {
  return guard(h, [&] { h->o->point_diagonal(d); });
}
int oracle_assemble(oracle_handle* h, double* blocks, long long* cols) {
  return guard(h, [&] { h->o->assemble(blocks, cols); });
}
int oracle_block_diagonal(oracle_handle* h, double* blocks) {
  return guard(h, [&] { h->o->block_diagonal(blocks); });
}
int oracle_mass_apply(oracle_handle* h, const double* z, double* y) {
  return guard(h, [&] { h->o->mass_apply(z, y); });
}
int oracle_interpolate(oracle_handle* h, double* x) {
  return guard(h, [&] { h->o->interpolate(x); });
}
double oracle_l2_error_sq(oracle_handle* h, const double* x) { return h->o->l2_error_sq(x); }
int oracle_grad_logkappa(oracle_handle* h, const double* u, const double* lam, double* g) {
  return guard(h, [&] { h->o->grad_logkappa(u, lam, g); });
}
int oracle_grad_source(oracle_handle* h, const double* lam, double* g) {
  return guard(h, [&] { h->o->grad_source(lam, g); });
}

// Krylov solve of J z = r (ISTL backend apply(z, r, reduction)).
// matrix_based: 0 = OnTheFlyOperator over jacobian_apply, 1 = assembled block matrix
// (MatrixAdapter), 2 = transposed operator (adjoint).
// LU factors of the diagonal blocks of J (bt = 0) or J^T (bt = 1), computed once per handle.
static void block_lu_prepare(oracle_handle* h, int bt) {
  auto& o = *h->o;
  const size_t nloc = o.nloc, n2 = nloc * nloc;
  if (!h->blu[bt].empty()) return;
  std::vector<double>& B = h->blu[bt];
  B.resize(o.grid.ncells * n2);
  h->bpiv[bt].resize(o.N);
  o.block_diagonal(B.data());
  for (size_t c = 0; c < o.grid.ncells; ++c) {
    double* A = &B[c * n2];
    if (bt)
      for (size_t i = 0; i < nloc; ++i)
        for (size_t j = i + 1; j < nloc; ++j) std::swap(A[i * nloc + j], A[j * nloc + i]);
    lu_factor(nloc, A, &h->bpiv[bt][c * nloc]);
  }
}

// Block SOR needs the off-diagonal blocks A_ij (the assembled matrix) and the LU factors of
// the diagonal blocks.
static void sor_prepare(oracle_handle* h) {
  auto& o = *h->o;
  const size_t n2 = o.nloc * o.nloc;
  if (h->blocks.empty()) {
    h->blocks.resize(o.grid.ncells * 7 * n2);
    h->cols.resize(o.grid.ncells * 7);
    o.assemble(h->blocks.data(), h->cols.data());
  }
  block_lu_prepare(h, 0);
}

// One forward block-SOR step, R:dune/pdelab/backend/istl/matrixfree/blocksorpreconditioner.hh:19-27
// (the cell update at :221-250): cells in index order, a_i = d_i - sum_{j<i} A_ij v_j^(k) -
// sum_{j>i} A_ij v_j^(k-1), D_i b_i = a_i, v_i = (1 - omega) v_i + omega b_i, in place on v
// (GridOperatorPreconditioner::apply, gridoperatorpreconditioner.hh:78-84). from_zero: v^(k-1)
// = 0, as on the zeroed v of dune-istl's BiCGSTAB (only the lower neighbours contribute).
static void sor_sweep(oracle_handle* h, const double* d, double* v, bool from_zero = true) {
  auto& o = *h->o;
  const size_t nloc = o.nloc, n2 = nloc * nloc;
  std::vector<double> a(nloc), b(nloc);
  if (from_zero) std::fill(v, v + o.N, 0.0);
  for (size_t c = 0; c < o.grid.ncells; ++c) {
    for (size_t i = 0; i < nloc; ++i) a[i] = d[c * nloc + i];
    for (int sl = 1; sl < 7; ++sl) {
      const long long cc = h->cols[c * 7 + sl];
      if (cc < 0) continue;
      const double* B = &h->blocks[(c * 7 + sl) * n2];
      const double* vn = v + (size_t)cc * nloc;
      for (size_t i = 0; i < nloc; ++i) {
        double s = 0.0;
        for (size_t j = 0; j < nloc; ++j) s += B[i * nloc + j] * vn[j];
        a[i] -= s;
      }
    }
    lu_solve(nloc, &h->blu[0][c * n2], &h->bpiv[0][c * nloc], a.data(), b.data());
    for (size_t i = 0; i < nloc; ++i) v[c * nloc + i] = (1.0 - h->sor_omega) * v[c * nloc + i] + h->sor_omega * b[i];
  }
}

// The iterative block Jacobi preconditioner, R:dune/pdelab/backend/istl/matrixfree/
// iterativeblockjacobipreconditioner.hh:267-450 (IterativeBlockJacobiPreconditionerLocalOperator,
// linear case :355-390): per cell, dune-istl's BiCGSTAB on D_c y = d_c from y = 0 (the
// BlockDiagonalOperator, :84-150), preconditioned by weight * diag(D_c)^{-1}
// (LocalPointJacobiPreconditioner, :28-80; the identity when precondition is off), to the
// residual reduction and iteration limit of BlockSolverOptions (:198-231, defaults 1e-5 and
// 100). v_c = y (GridOperatorPreconditioner::apply on a zeroed v). A breakdown of a block
// solve throws, as the ISTL solver does.
static void ibj_prepare(oracle_handle* h) {
  auto& o = *h->o;
  if (!h->bdiag.empty()) return;
  h->bdiag.resize(o.grid.ncells * o.nloc * o.nloc);
  o.block_diagonal(h->bdiag.data());
}

static void ibj_apply(oracle_handle* h, const double* d, double* v) {
  auto& o = *h->o;
  const size_t nloc = o.nloc, n2 = nloc * nloc;
  std::vector<double> b(nloc), x(nloc);
  h->ibj_iterations = 0;
  for (size_t c = 0; c < o.grid.ncells; ++c) {
    const double* D = &h->bdiag[c * n2];
    oracle::OpFn A = [&](const double* in, double* out) {
      for (size_t i = 0; i < nloc; ++i) {
        double s = 0.0;
        for (size_t j = 0; j < nloc; ++j) s += D[i * nloc + j] * in[j];
        out[i] = s;
      }
    };
    oracle::OpFn P = [&](const double* in, double* out) {
      for (size_t i = 0; i < nloc; ++i)
        out[i] = h->ibj_precondition ? h->ibj_weight * in[i] * (1.0 / D[i * nloc + i]) : in[i];
    };
    std::copy(d + c * nloc, d + (c + 1) * nloc, b.begin());
    std::fill(x.begin(), x.end(), 0.0);
    pdelab_linear_solver_result r;
    oracle::bicgstab(nloc, A, P, h->ibj_reduction, h->ibj_maxit, x.data(), b.data(), &r);
    h->ibj_iterations += r.iterations;
    std::copy(x.begin(), x.end(), v + c * nloc);
  }
}

int oracle_iterative_block_jacobi_apply(oracle_handle* h, const double* d, double* v) {
  return guard(h, [&] {
    ibj_prepare(h);
    ibj_apply(h, d, v);
  });
}

// v = the block-SOR preconditioner applied to d: one sweep from v = 0 (inplace = 0), or one SOR
// step from the v given (inplace = 1)
int oracle_block_sor_apply(oracle_handle* h, const double* d, double* v, int inplace) {
  return guard(h, [&] {
    sor_prepare(h);
    sor_sweep(h, d, v, inplace == 0);
  });
}

int oracle_solve(oracle_handle* h, int method, int precond, double reduction, int maxit, double* z, double* r,
                 pdelab_linear_solver_result* res, int matrix_based) {
  return guard(h, [&] {
    auto& o = *h->o;
    const size_t n = o.N, nloc = o.nloc, n2 = nloc * nloc;
    if (maxit <= 0) maxit = 5000;
    if (precond == PDELAB_PREC_JACOBI && h->diag.empty()) {
      h->diag.assign(n, 0.0);
      o.point_diagonal(h->diag.data());
    }
    const int bt = matrix_based == 2 ? 1 : 0;  // blocks of J^T for the transposed operator
    const bool sor = precond == PDELAB_PREC_BLOCK_SOR;
    if (sor && matrix_based == 2) throw std::runtime_error("block SOR: not for the transposed operator");
    if (sor && method != PDELAB_SOLVER_BICGSTAB)
      throw std::runtime_error("block SOR: BiCGSTAB only (the preconditioner is not symmetric)");
    const bool ibj = precond == PDELAB_PREC_ITERATIVE_BLOCK_JACOBI;
    if (ibj && matrix_based == 2) throw std::runtime_error("iterative block Jacobi: not for the transposed operator");
    if (ibj && method != PDELAB_SOLVER_BICGSTAB)
      throw std::runtime_error("iterative block Jacobi: BiCGSTAB only (the preconditioner is an inner iteration)");
    if (sor) sor_prepare(h);
    else if (ibj) ibj_prepare(h);
    else if (precond == PDELAB_PREC_BLOCK_JACOBI) block_lu_prepare(h, bt);
    oracle::OpFn A;
    if (matrix_based == 1) {
      if (h->blocks.empty()) {
        h->blocks.resize(o.grid.ncells * 7 * n2);
        h->cols.resize(o.grid.ncells * 7);
        o.assemble(h->blocks.data(), h->cols.data());
      }
      A = [&](const double* x, double* y) {
        for (size_t c = 0; c < o.grid.ncells; ++c)
          for (size_t i = 0; i < nloc; ++i) {
            double s = 0.0;
            for (int sl = 0; sl < 7; ++sl) {
              long long cc = h->cols[c * 7 + sl];
              if (cc < 0) continue;
              const double* B = &h->blocks[(c * 7 + sl) * n2 + i * nloc];
              const double* xc = x + (size_t)cc * nloc;
              for (size_t j = 0; j < nloc; ++j) s += B[j] * xc[j];
            }
            y[c * nloc + i] = s;
          }
      };
    } else if (matrix_based == 2) {
      A = [&](const double* x, double* y) {
        std::fill(y, y + n, 0.0);
        o.jacobian_apply_transpose(x, y);
      };
    } else {
      A = [&](const double* x, double* y) {  // OnTheFlyOperator::apply: y = 0; jacobian_apply
        std::fill(y, y + n, 0.0);
        if (h->threads > 1) o.jacobian_apply_twosided(x, y, h->threads);
        else o.jacobian_apply(x, y);
      };
    }
    oracle::OpFn P = [&](const double* d, double* v) {
      if (precond == PDELAB_PREC_JACOBI)
        for (size_t i = 0; i < n; ++i) v[i] = d[i] / h->diag[i];
      else if (precond == PDELAB_PREC_BLOCK_JACOBI)
        for (size_t c = 0; c < o.grid.ncells; ++c)
          lu_solve(nloc, &h->blu[bt][c * n2], &h->bpiv[bt][c * nloc], d + c * nloc, v + c * nloc);
      else if (sor)
        sor_sweep(h, d, v);
      else if (ibj)
        ibj_apply(h, d, v);
      else
        std::copy(d, d + n, v);
    };
    if (method == PDELAB_SOLVER_BICGSTAB) oracle::bicgstab(n, A, P, reduction, maxit, z, r, res);
    else oracle::cg(n, A, P, reduction, maxit, z, r, res);
  });
}

// ---------------------------------------------------------------------------
// Instationary: OneStepGridOperator over the operator J and the DG mass M
// (R:dune/pdelab/gridoperator/onestep.hh, weights of gridoperator/onestep/localassembler.hh:122-126
// in MultiplyOperator0ByDT mode: b dt on J, a on M) and OneStepMethod::apply
// (R:dune/pdelab/instationary/implicitonestep.hh:144-230).

// cell mass matrix (the same for every cell of the uniform grid), row-major
static std::vector<double> cell_mass(oracle::Oracle& o) {
  const size_t nl = o.nloc;
  std::vector<double> m(nl * nl), e(nl), y(nl);
  for (size_t j = 0; j < nl; ++j) {
    std::fill(e.begin(), e.end(), 0.0);
    std::fill(y.begin(), y.end(), 0.0);
    e[j] = 1.0;
    oracle::mass_volume(o.grid, o.tab, 0, e.data(), y.data());
    for (size_t i = 0; i < nl; ++i) m[i * nl + j] = y[i];
  }
  return m;
}

// solves (b_dt J + a_mass M) z = r (OnTheFlyOperator over the one-step grid operator) with
// point or block Jacobi of that operator
static void onestep_solve(oracle_handle* h, int method, int precond, double a_mass, double b_dt, double reduction,
                          int maxit, double* z, double* r, pdelab_linear_solver_result* res) {
  auto& o = *h->o;
  const size_t n = o.N, nloc = o.nloc, n2 = nloc * nloc;
  const std::vector<double> Mc = cell_mass(o);
  std::vector<double> tmp(n);
  oracle::OpFn A = [&](const double* x, double* y) {
    std::fill(y, y + n, 0.0);
    o.jacobian_apply(x, y);
    for (size_t i = 0; i < n; ++i) y[i] *= b_dt;
    std::fill(tmp.begin(), tmp.end(), 0.0);
    o.mass_apply(x, tmp.data());
    for (size_t i = 0; i < n; ++i) y[i] += a_mass * tmp[i];
  };
  std::vector<double> diag;
  std::vector<double> blu;
  std::vector<int> piv;
  if (precond == PDELAB_PREC_JACOBI) {
    diag.assign(n, 0.0);
    o.point_diagonal(diag.data());
    for (size_t i = 0; i < n; ++i) diag[i] = b_dt * diag[i] + a_mass * Mc[(i % nloc) * nloc + (i % nloc)];
  } else if (precond == PDELAB_PREC_BLOCK_JACOBI) {
    blu.resize(o.grid.ncells * n2);
    piv.resize(n);
    o.block_diagonal(blu.data());
    for (size_t c = 0; c < o.grid.ncells; ++c) {
      double* B = &blu[c * n2];
      for (size_t k = 0; k < n2; ++k) B[k] = b_dt * B[k] + a_mass * Mc[k];
      lu_factor(nloc, B, &piv[c * nloc]);
    }
  }
  oracle::OpFn P = [&](const double* d, double* v) {
    if (precond == PDELAB_PREC_JACOBI)
      for (size_t i = 0; i < n; ++i) v[i] = d[i] / diag[i];
    else if (precond == PDELAB_PREC_BLOCK_JACOBI)
      for (size_t c = 0; c < o.grid.ncells; ++c) lu_solve(nloc, &blu[c * n2], &piv[c * nloc], d + c * nloc, v + c * nloc);
    else
      std::copy(d, d + n, v);
  };
  if (maxit <= 0) maxit = 5000;
  if (method == PDELAB_SOLVER_BICGSTAB) oracle::bicgstab(n, A, P, reduction, maxit, z, r, res);
  else oracle::cg(n, A, P, reduction, maxit, z, r, res);
}

int oracle_onestep_jacobian_apply(oracle_handle* h, double a_mass, double b_dt, const double* z, double* y) {
  return guard(h, [&] {
    auto& o = *h->o;
    std::vector<double> t(o.N, 0.0);
    o.jacobian_apply(z, t.data());
    for (size_t i = 0; i < o.N; ++i) y[i] += b_dt * t[i];
    std::fill(t.begin(), t.end(), 0.0);
    o.mass_apply(z, t.data());
    for (size_t i = 0; i < o.N; ++i) y[i] += a_mass * t[i];
  });
}

// coefficients a(r, i) = A[r-1][i], b(r, i) = B[r-1][i],
// R:dune/pdelab/instationary/onestepparameter.hh (ImplicitEuler :193, OneStepTheta :96-101,
// Alexander2 :449-461)
int oracle_onestep_step(oracle_handle* h, int scheme, int method, int precond, double dt, double reduction, int maxit,
                        const double* xold, double* xnew, pdelab_onestep_result* res) {
  return guard(h, [&] {
    auto& o = *h->o;
    const size_t n = o.N;
    int s = 1;
    double A[2][3] = {{-1, 1, 0}, {0, 0, 0}}, B[2][3] = {{0, 1, 0}, {0, 0, 0}};
    if (scheme == PDELAB_TS_CRANK_NICOLSON) {
      B[0][0] = 0.5;
      B[0][1] = 0.5;
    } else if (scheme == PDELAB_TS_ALEXANDER2) {
      const double alpha = 1.0 - 0.5 * std::sqrt(2.0);
      s = 2;
      A[1][0] = -1; A[1][1] = 0; A[1][2] = 1;
      B[0][0] = 0; B[0][1] = alpha; B[0][2] = 0;
      B[1][0] = 0; B[1][1] = 1 - alpha; B[1][2] = alpha;
    } else if (scheme != PDELAB_TS_IMPLICIT_EULER) {
      throw std::runtime_error("unknown one-step scheme");
    }
    std::memset(res, 0, sizeof(*res));
    res->converged = 1;
    res->stages = s;
    std::vector<std::vector<double>> u(s + 1, std::vector<double>(n));
    std::copy(xold, xold + n, u[0].begin());
    std::vector<double> C(n), r(n), z(n), t(n);
    // y += w R(v) + wm M v
    auto add_terms = [&](const std::vector<double>& v, double w, double wm, std::vector<double>& y) {
      if (w != 0.0) {
        std::fill(t.begin(), t.end(), 0.0);
        o.residual(v.data(), t.data());
        for (size_t i = 0; i < n; ++i) y[i] += w * t[i];
      }
      if (wm != 0.0) {
        std::fill(t.begin(), t.end(), 0.0);
        o.mass_apply(v.data(), t.data());
        for (size_t i = 0; i < n; ++i) y[i] += wm * t[i];
      }
    };
    for (int rr = 1; rr <= s; ++rr) {
      std::fill(C.begin(), C.end(), 0.0);
      for (int j = 0; j < rr; ++j) add_terms(u[j], B[rr - 1][j] * dt, A[rr - 1][j], C);
      if (rr == 1) std::copy(xnew, xnew + n, u[1].begin());
      else u[rr] = u[rr - 1];
      const double bdt = B[rr - 1][rr] * dt, am = A[rr - 1][rr];
      r = C;
      add_terms(u[rr], bdt, am, r);
      const double defect = oracle::norm2(n, r.data());
      const double red = std::max(reduction, 1e-99 / defect);
      std::fill(z.begin(), z.end(), 0.0);
      pdelab_linear_solver_result ls;
      onestep_solve(h, method, precond, am, bdt, red, maxit, z.data(), r.data(), &ls);
      res->linear_solver_iterations += (int)ls.iterations;
      if (!ls.converged) res->converged = 0;
      for (size_t i = 0; i < n; ++i) u[rr][i] -= z[i];
    }
    std::copy(u[s].begin(), u[s].end(), xnew);
  });
}

// StationaryLinearProblemSolver::apply, R:dune/pdelab/stationary/linearproblem.hh:188-302
int oracle_linear_problem_solve(oracle_handle* h, int method, int precond, double reduction, double min_defect,
                                int maxit, double* x, pdelab_linear_problem_result* res, int matrix_based) {
  auto& o = *h->o;
  const size_t n = o.N;
  std::vector<double> r(n, 0.0), z(n, 0.0);
  auto t0 = std::chrono::steady_clock::now();
  int st = guard(h, [&] { o.residual(x, r.data()); });
  if (st) return st;
  double tasm = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double defect = oracle::norm2(n, r.data());
  double red = std::max(reduction, min_defect / defect);
  auto t1 = std::chrono::steady_clock::now();
  st = oracle_solve(h, method, precond, red, maxit, z.data(), r.data(), &res->solver, matrix_based);
  if (st) return st;
  res->linear_solver_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  res->assembler_time = tasm;
  res->first_defect = defect;
  res->defect = defect * res->solver.reduction;
  res->linear_solver_iterations = (int)res->solver.iterations;
  for (size_t i = 0; i < n; ++i) x[i] -= z[i];
  return 0;
}

// Times jacobian_apply: repeats until min_seconds of CPU work, returns seconds/apply.
double oracle_time_apply(oracle_handle* h, int nthreads, double min_seconds, int* reps_out) {
  auto& o = *h->o;
  std::vector<double> z(o.N), y(o.N, 0.0);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (size_t i = 0; i < o.N; ++i) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t t = s;
    t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
    t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
    t ^= t >> 31;
    z[i] = 2.0 * ((t >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
  int reps = 0;
  auto t0 = std::chrono::steady_clock::now();
  double el = 0.0;
  do {
    if (nthreads > 1) o.jacobian_apply_twosided(z.data(), y.data(), nthreads);
    else o.jacobian_apply(z.data(), y.data());
    ++reps;
    el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } while (el < min_seconds);
  if (reps_out) *reps_out = reps;
  return el / reps;
}

// 1D tables for the fixture check: Gauss rule of `order` and p_i, p_i' at its nodes
// and at 0 and 1. Layout: x[m], w[m], P[m][k+1], D[m][k+1], e0[k+1], e1[k+1], d0[k+1], d1[k+1]
int oracle_tables_1d(int k, int order, double* out) {
  oracle::Rule1D r = oracle::gauss_legendre(order);
  int m = (int)r.x.size(), n1 = k + 1, o = 0;
  for (int q = 0; q < m; ++q) out[o++] = r.x[q];
  for (int q = 0; q < m; ++q) out[o++] = r.w[q];
  for (int q = 0; q < m; ++q)
    for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_p(k, i, r.x[q]);
  for (int q = 0; q < m; ++q)
    for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_dp(k, i, r.x[q]);
  for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_p(k, i, 0.0);
  for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_p(k, i, 1.0);
  for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_dp(k, i, 0.0);
  for (int i = 0; i < n1; ++i) out[o++] = oracle::lagrange_dp(k, i, 1.0);
  return m;
}

}  // extern "C"
