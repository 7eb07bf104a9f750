// 3D restatement of R:dune/pdelab/test/testconvectiondiffusiondg.cc on the HIP path,
// written against the PDELab-shaped host classes: Poisson problem with the manufactured
// solution u = exp(-|x - 1/2|^2), Q1 DG on 16^3 cells, ConvectionDiffusionDG with the
// default SIPG / weightsOn / alpha 1, StationaryLinearProblemSolver with a matrix-free
// CG + Jacobi backend to 1e-12; pass iff the squared L2 error <= 1e-6 (the reference's
// threshold). Further checks: the nonlinear jacobian_apply throws for this linear
// operator (gridoperator.hh:202-203), OnTheFlyOperator::apply equals jacobian_apply,
// and the mixed-precision backend reaches 1e-10 on a lognormal-kappa problem.
// Returns nonzero on failure, like the reference tests.
#include <cmath>
#include <cstdio>
#include <random>

#include "pdelab_amd.hh"

using namespace pdelab_amd;

struct PoissonProblem : ConvectionDiffusionModelProblem {
  static double u(const std::array<double, 3>& x) {
    double c = 0;
    for (int d = 0; d < 3; ++d) c += (0.5 - x[d]) * (0.5 - x[d]);
    return std::exp(-c);
  }
  double f(const Element& e, const std::array<double, 3>& xl) const {
    auto x = e.geometry().global(xl);
    double c = 0;
    for (int d = 0; d < 3; ++d) c += (0.5 - x[d]) * (0.5 - x[d]);
    return (6.0 - 4.0 * c) * std::exp(-c);  // -lap u in 3D
  }
  double g(const Element& e, const std::array<double, 3>& xl) const { return u(e.geometry().global(xl)); }
};

struct LognormalProblem : ConvectionDiffusionModelProblem {
  std::vector<double> kappa;
  double A(const Element& e, const std::array<double, 3>&) const { return kappa[e.index]; }
  double f(const Element&, const std::array<double, 3>&) const { return 1.0; }
  double g(const Element&, const std::array<double, 3>&) const { return 0.0; }
};

// The same field as a diagonal diffusion tensor (PermTensorType); a full tensor is refused.
struct TensorProblem : LognormalProblem {
  bool full = false;
  std::array<std::array<double, 3>, 3> A(const Element& e, const std::array<double, 3>&) const {
    const double k = kappa[e.index];
    return {{{k, full ? 0.1 : 0.0, 0.0}, {0.0, k, 0.0}, {0.0, 0.0, k}}};
  }
};

// u = 0.3 + x + 2y - z + x^2 + 1.5 y^2 - 2 z^2 with the divergence-free linear velocity
// b = (0.5 - (y - 1/2) + 0.3 z, 0.25 + (x - 1/2) - 0.2 z, 0.2 + 0.5 (x - 1/2) - 0.3 (y - 1/2)),
// kappa 0.7: f = -kappa lap u + b.grad u, g = u. Every integral of the Q2 scheme is exact, so
// the discrete solution is the interpolant of u: the tabulated velocity field end to end.
struct VelocityFieldProblem : ConvectionDiffusionModelProblem {
  static double u(const std::array<double, 3>& x) {
    return 0.3 + x[0] + 2 * x[1] - x[2] + x[0] * x[0] + 1.5 * x[1] * x[1] - 2 * x[2] * x[2];
  }
  static std::array<double, 3> bfield(const std::array<double, 3>& x) {
    return {0.5 - (x[1] - 0.5) + 0.3 * x[2], 0.25 + (x[0] - 0.5) - 0.2 * x[2], 0.2 + 0.5 * (x[0] - 0.5) - 0.3 * (x[1] - 0.5)};
  }
  double A(const Element&, const std::array<double, 3>&) const { return 0.7; }
  std::array<double, 3> b(const Element& e, const std::array<double, 3>& xl) const {
    return bfield(e.geometry().global(xl));
  }
  double f(const Element& e, const std::array<double, 3>& xl) const {
    const auto x = e.geometry().global(xl);
    const auto bb = bfield(x);
    const double gu[3] = {1 + 2 * x[0], 2 + 3 * x[1], -1 - 4 * x[2]};
    return -0.7 * (2 + 3 - 4) + bb[0] * gu[0] + bb[1] * gu[1] + bb[2] * gu[2];
  }
  double g(const Element& e, const std::array<double, 3>& xl) const { return u(e.geometry().global(xl)); }
};

int main() {
  int fails = 0;
  try {
    // --- testconvectiondiffusiondg.cc in 3D
    StructuredGrid grid({1.0, 1.0, 1.0}, {16, 16, 16});
    using FEM = QkDGLocalFiniteElementMap<double, double, 1, 3>;
    FEM fem;
    using GFS = GridFunctionSpace<FEM>;
    GFS gfs(grid, fem);
    PoissonProblem problem;
    using LOP = ConvectionDiffusionDG<PoissonProblem, FEM>;
    LOP lop(problem);
    using GO = GridOperator<GFS, GFS, LOP>;
    GO go(gfs, gfs, lop);
    std::printf("gfs with %zu dofs generated\n", gfs.size());
    auto x = go.makeVector();
    x = 0.0;
    using LS = ISTLBackend_SEQ_MatrixFree_CG_Jac<GO>;
    LS ls(go, 5000, 0);
    StationaryLinearProblemSolver<GO, LS, GO::Domain> solver(go, ls, x, 1e-12);
    solver.apply();
    const auto& res = solver.result();
    const double err = l2_error_squared(gfs, x.download(), PoissonProblem::u, 10);
    std::printf("l2errorsquared: %g (iterations %u, converged %d)\n", err, res.iterations, (int)res.converged);
    if (std::isnan(err) || std::fabs(err) > 1e-6 || !res.converged) ++fails;

    // --- instationary, as testinstationaryfastdgassembler.cc:137-209: OneStepGridOperator over the
    // operator and the DG mass, Alexander2, one step of dt = 0.1 from the Dirichlet-free start
    // u = 0; then two steps of dt = 1e6, which must reach the stationary solution above
    {
      using IGO = OneStepGridOperator<GO, L2>;
      IGO igo(go, L2(2 * 1));
      auto xo = go.makeVector(), xn = go.makeVector();
      xo = 0.0;
      xn = 0.0;
      StationaryLinearProblemSolver<IGO, LS, GO::Domain> pdesolver(igo, ls, xn, 1e-10);
      Alexander2Parameter<double> method;
      OneStepMethod<double, IGO, decltype(pdesolver), GO::Domain, GO::Domain> osm(method, igo, pdesolver);
      double time = osm.apply(0.0, 0.1, xo, xn);
      std::printf("Alexander2 step to t = %g: stages %d, iterations %d, converged %d\n", time,
                  osm.result().stages, osm.result().linear_solver_iterations, osm.result().converged);
      if (!osm.result().converged || osm.result().stages != 2) ++fails;
      for (int i = 0; i < 2; ++i) {
        xo = xn;
        osm.apply(time, 1e6, xo, xn);
      }
      auto a = xn.download(), b = x.download();
      double num = 0, den = 0;
      for (size_t i = 0; i < a.size(); ++i) { num += (a[i] - b[i]) * (a[i] - b[i]); den += b[i] * b[i]; }
      std::printf("Alexander2 with dt = 1e6 vs the stationary solution: relative difference %g\n",
                  std::sqrt(num / den));
      if (!(std::sqrt(num / den) < 1e-8)) ++fails;
    }

    // --- linear / nonlinear jacobian_apply dispatch
    bool threw = false;
    try {
      auto y = go.makeVector();
      go.jacobian_apply(x, x, y);
    } catch (const Exception&) {
      threw = true;
    }
    std::printf("nonlinear jacobian_apply on linear operator throws: %d\n", (int)threw);
    if (!threw) ++fails;

    // --- OnTheFlyOperator::apply == jacobian_apply
    OnTheFlyOperator<GO::Domain, GO::Range, GO> op(go);
    auto y1 = go.makeVector();
    auto y2 = go.makeVector();
    op.apply(x, y1);
    go.jacobian_apply(x, y2);
    auto h1 = y1.download(), h2 = y2.download();
    double diff = 0;
    for (size_t i = 0; i < h1.size(); ++i) diff = std::max(diff, std::fabs(h1[i] - h2[i]));
    std::printf("OnTheFlyOperator vs jacobian_apply max diff: %g\n", diff);
    if (diff != 0.0) ++fails;

    // --- mixed precision on a lognormal kappa field, Q3
    StructuredGrid grid3({1.0, 1.0, 1.0}, {12, 12, 12});
    using FEM3 = QkDGLocalFiniteElementMap<double, double, 3, 3>;
    FEM3 fem3;
    GridFunctionSpace<FEM3> gfs3(grid3, fem3);
    LognormalProblem lp;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (size_t c = 0; c < grid3.size(); ++c) lp.kappa.push_back(std::exp(nd(rng)));
    using LOP3 = ConvectionDiffusionDG<LognormalProblem, FEM3>;
    LOP3 lop3(lp, ConvectionDiffusionDGMethod::SIPG, ConvectionDiffusionDGWeights::weightsOn, 3.0);
    using GO3 = GridOperator<GridFunctionSpace<FEM3>, GridFunctionSpace<FEM3>, LOP3>;
    GO3 go3(gfs3, gfs3, lop3);
    auto xm = go3.makeVector();
    MixedPrecisionSolver<GO3, Krylov::CG> mixed(go3, 0.1);
    auto mr = mixed.apply(xm, 1e-10);
    std::printf("mixed precision: converged %d outer %d inner %d reduction %g\n", mr.converged, mr.outer_iterations,
                mr.inner_iterations, mr.reduction);
    if (!mr.converged || mr.reduction > 1e-10) ++fails;
    // FP64 reference solve on the same operator
    auto x64 = go3.makeVector();
    ISTLBackend_SEQ_MatrixFree_CG_Jac<GO3> ls3(go3, 20000, 0);
    StationaryLinearProblemSolver<GO3, decltype(ls3), GO3::Domain> s3(go3, ls3, x64, 1e-12);
    s3.apply();
    auto a = xm.download(), b = x64.download();
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) { num += (a[i] - b[i]) * (a[i] - b[i]); den += b[i] * b[i]; }
    std::printf("mixed vs FP64 relative difference: %g\n", std::sqrt(num / den));
    if (std::sqrt(num / den) > 1e-8) ++fails;
    // the same solve with per-cell block Jacobi (Blocking::fixed): fewer iterations, same solution
    auto xb = go3.makeVector();
    ISTLBackend_SEQ_MatrixFree_CG_BlockJac<GO3> lsb(go3, 20000, 0);
    StationaryLinearProblemSolver<GO3, decltype(lsb), GO3::Domain> sb(go3, lsb, xb, 1e-12);
    sb.apply();
    auto c = xb.download();
    num = 0;
    for (size_t i = 0; i < c.size(); ++i) num += (c[i] - b[i]) * (c[i] - b[i]);
    std::printf("block Jacobi: %d iterations (point Jacobi %d), relative difference %g\n",
                (int)lsb.result().iterations, (int)ls3.result().iterations, std::sqrt(num / den));
    if (std::sqrt(num / den) > 1e-8 || lsb.result().iterations >= ls3.result().iterations) ++fails;
    // the preconditioner object on its own: v += D^{-1} d, with d = J xb from OnTheFlyOperator
    AssembledBlockJacobiPreconditioner<GO3> bjp(go3);
    auto d = go3.makeVector(), v = go3.makeVector(), d2 = go3.makeVector();
    go3.jacobian_apply(xb, d);
    OnTheFlyOperator<GO3::Domain, GO3::Range, GO3> op3(go3);
    d2 = 1.0;
    op3.apply(xb, d2);
    {
      auto p = d.download(), q = d2.download();
      double e = 0, m = 0;
      for (size_t i = 0; i < p.size(); ++i) { e = std::max(e, std::fabs(p[i] - q[i])); m = std::max(m, std::fabs(p[i])); }
      std::printf("OnTheFlyOperator::apply vs jacobian_apply: max difference %g (max %g)\n", e, m);
      if (e > 1e-14 * m) ++fails;
    }
    bjp.apply(v, d);
    auto vh = v.download();
    double vn = 0;
    for (double e : vh) vn += e * e;
    if (!(vn > 0) || !std::isfinite(vn)) ++fails;

    // --- BiCGSTAB with the matrix-free block SOR and the iterative block Jacobi (the two
    // preconditioners of matrix_free_linear.cc): the same solution
    {
      auto xs = go3.makeVector();
      ISTLBackend_SEQ_MatrixFree_BCGS_BlockSOR<GO3> lss(go3, 20000, 0);
      StationaryLinearProblemSolver<GO3, decltype(lss), GO3::Domain> ss(go3, lss, xs, 1e-12);
      ss.apply();
      auto xi = go3.makeVector();
      ISTLBackend_SEQ_MatrixFree_BCGS_IterativeBlockJac<GO3> lsi(go3, 20000, 0);
      StationaryLinearProblemSolver<GO3, decltype(lsi), GO3::Domain> si(go3, lsi, xi, 1e-12);
      si.apply();
      auto hs = xs.download(), hi = xi.download();
      double es = 0, ei = 0;
      for (size_t i = 0; i < hs.size(); ++i) {
        es += (hs[i] - b[i]) * (hs[i] - b[i]);
        ei += (hi[i] - b[i]) * (hi[i] - b[i]);
      }
      std::printf("BiCGSTAB block SOR: %d iterations, rel. difference %g; iterative block Jacobi: %d, %g\n",
                  (int)lss.result().iterations, std::sqrt(es / den), (int)lsi.result().iterations, std::sqrt(ei / den));
      if (std::sqrt(es / den) > 1e-8 || std::sqrt(ei / den) > 1e-8 || !lss.result().converged || !lsi.result().converged)
        ++fails;
      // the preconditioner objects: SOR on v = 0 and the accumulating iterative block Jacobi
      BlockSORPreconditioner<GO3> sor(go3, 1.0);
      IterativeBlockJacobiPreconditioner<GO3> ibj(go3, BlockSolverOptions(1e-5, 100));
      auto vs = go3.makeVector(), vi = go3.makeVector();
      vs = 0.0;
      vi = 0.0;
      sor.apply(vs, d);
      ibj.apply(vi, d);
      auto a1 = vs.download(), a2 = vi.download();
      double n1 = 0, n2 = 0;
      for (size_t i = 0; i < a1.size(); ++i) { n1 += a1[i] * a1[i]; n2 += a2[i] * a2[i]; }
      std::printf("preconditioner objects: |W_sor^-1 d| %g, |W_ibj^-1 d| %g, block iterations %llu\n", std::sqrt(n1),
                  std::sqrt(n2), ibj.iterations());
      if (!(n1 > 0) || !(n2 > 0) || !std::isfinite(n1 + n2) || ibj.iterations() == 0) ++fails;
    }

    // --- the lognormal field given as the diagonal tensor diag(k, k, k): the same solve, bit for bit
    {
      TensorProblem tp;
      tp.kappa = lp.kappa;
      using LOPT = ConvectionDiffusionDG<TensorProblem, FEM3>;
      LOPT lopt(tp, ConvectionDiffusionDGMethod::SIPG, ConvectionDiffusionDGWeights::weightsOn, 3.0);
      using GOT = GridOperator<GridFunctionSpace<FEM3>, GridFunctionSpace<FEM3>, LOPT>;
      GOT got(gfs3, gfs3, lopt);
      auto xt = got.makeVector();
      ISTLBackend_SEQ_MatrixFree_CG_Jac<GOT> lst(got, 20000, 0);
      StationaryLinearProblemSolver<GOT, decltype(lst), GOT::Domain> st(got, lst, xt, 1e-12);
      st.apply();
      auto t = xt.download();
      double dmax = 0;
      for (size_t i = 0; i < t.size(); ++i) dmax = std::max(dmax, std::fabs(t[i] - b[i]));
      std::printf("diagonal tensor vs scalar kappa: max difference %g, iterations %d and %d\n", dmax,
                  (int)lst.result().iterations, (int)ls3.result().iterations);
      if (dmax != 0.0 || lst.result().iterations != ls3.result().iterations) ++fails;
      tp.full = true;
      LOPT lopf(tp, ConvectionDiffusionDGMethod::SIPG, ConvectionDiffusionDGWeights::weightsOn, 3.0);
      bool refused = false;
      try {
        GOT gof(gfs3, gfs3, lopf);
      } catch (const NotImplemented&) {
        refused = true;
      }
      std::printf("full tensor refused: %d\n", (int)refused);
      if (!refused) ++fails;
    }

    // --- a velocity field b(x) given by the parameter class, tabulated at the quadrature points
    {
      StructuredGrid gridv({1.0, 1.0, 1.0}, {6, 5, 4});
      using FEM2 = QkDGLocalFiniteElementMap<double, double, 2, 3>;
      FEM2 fem2;
      GridFunctionSpace<FEM2> gfsv(gridv, fem2);
      VelocityFieldProblem vp;
      using LOPV = ConvectionDiffusionDG<VelocityFieldProblem, FEM2>;
      LOPV lopv(vp, ConvectionDiffusionDGMethod::SIPG, ConvectionDiffusionDGWeights::weightsOn, 4.0);
      using GOV = GridOperator<GridFunctionSpace<FEM2>, GridFunctionSpace<FEM2>, LOPV>;
      GOV gov(gfsv, gfsv, lopv);
      auto xv = gov.makeVector();
      xv = 0.0;
      ISTLBackend_SEQ_MatrixFree_BCGS_Jac<GOV> lsv(gov, 5000, 0);
      StationaryLinearProblemSolver<GOV, decltype(lsv), GOV::Domain> sv(gov, lsv, xv, 1e-13);
      sv.apply();
      const double ev = l2_error_squared(gfsv, xv.download(), VelocityFieldProblem::u, 10);
      std::printf("velocity field b(x), Q2 BiCGSTAB: l2errorsquared %g (iterations %u, converged %d)\n", ev,
                  sv.result().iterations, (int)sv.result().converged);
      if (!(ev < 1e-18) || !sv.result().converged) ++fails;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::printf(fails ? "FAILED (%d)\n" : "PASSED\n", fails);
  return fails ? 1 : 0;
}
