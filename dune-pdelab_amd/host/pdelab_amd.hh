// pdelab_amd.hh -- C++ host classes shaped like PDELab's GridOperator / LocalOperator /
// ISTL solver backends for the SIPG DG hot path, on top of the C-ABI of
// include/pdelab_gpu.h (no HIP headers needed; link with libpdelab_gpu.so).
//
// Reference shapes mirrored here (same names, argument meaning and error behaviour):
//   ConvectionDiffusionModelProblem   R:dune/pdelab/localoperator/convectiondiffusionparameter.hh:134-209
//   ConvectionDiffusionDG             R:dune/pdelab/localoperator/convectiondiffusiondg.hh:54-102
//   QkDGLocalFiniteElementMap         R:dune/pdelab/finiteelementmap/qkdg.hh:36-76
//   GridFunctionSpace / Backend::Vector  (DG leaf ordering, cell-blocked)
//   GridOperator                      R:dune/pdelab/gridoperator/gridoperator.hh:30-240
//   OnTheFlyOperator                  R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:44-100
//   ISTLBackend_SEQ_MatrixFree_Base   R:dune/pdelab/backend/istl/matrixfree/backends.hh:62-143
//   SequentialNorm, LinearSolverResult R:dune/pdelab/backend/solver.hh:16-90
//   StationaryLinearProblemSolver     R:dune/pdelab/stationary/linearproblem.hh:59-338
//
// The user's parameter class keeps the reference interface (A, b, c, f, g, bctype taking
// an element and a local coordinate). The GridOperator evaluates it once on the host:
// kappa at cell centres (permeabilityIsConstantPerCell), b and c at cell centres (they
// must be constant), f at the volume quadrature points and g at the boundary face
// quadrature points; the device then works from those tables.
#ifndef PDELAB_AMD_HOST_HH
#define PDELAB_AMD_HOST_HH

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/pdelab_gpu.h"

namespace pdelab_amd {

// ---- exceptions (the DUNE_THROW sites of the path) --------------------------
struct Exception : std::runtime_error {
  explicit Exception(const std::string& m) : std::runtime_error(m) {}
};
struct InvalidStateException : Exception {
  using Exception::Exception;
};
struct NotImplemented : Exception {
  using Exception::Exception;
};
struct SolverAbort : Exception {
  using Exception::Exception;
};
struct HipError : Exception {
  using Exception::Exception;
};

inline void check(int status, const pdelab_gpu_gridoperator* h) {
  if (status == PDELAB_OK) return;
  std::string msg = pdelab_gpu_last_error(h);
  switch (status) {
    case PDELAB_ERR_LINEARITY_MISMATCH:
    case PDELAB_ERR_OUTFLOW_ON_INFLOW: throw Exception(msg);
    case PDELAB_ERR_SOLVER_ABORT: throw SolverAbort(msg);
    case PDELAB_ERR_UNSUPPORTED: throw NotImplemented(msg);
    case PDELAB_ERR_HIP:
    case PDELAB_ERR_NO_DEVICE: throw HipError(msg);
    default: throw Exception(msg);
  }
}

// ---- grid ---------------------------------------------------------------------
// YaspGrid<3> on [0,L0]x[0,L1]x[0,L2] with n0*n1*n2 cells, lexicographic (x fastest).
struct StructuredGrid {
  std::array<int, 3> cells;
  std::array<double, 3> upper;
  StructuredGrid(std::array<double, 3> upper_, std::array<int, 3> cells_) : cells(cells_), upper(upper_) {}
  size_t size() const { return (size_t)cells[0] * cells[1] * cells[2]; }
  double h(int d) const { return upper[d] / cells[d]; }
};

struct Geometry {
  std::array<double, 3> lower, extent;
  std::array<double, 3> global(const std::array<double, 3>& xi) const {
    return {lower[0] + extent[0] * xi[0], lower[1] + extent[1] * xi[1], lower[2] + extent[2] * xi[2]};
  }
  std::array<double, 3> center() const { return global({0.5, 0.5, 0.5}); }
  double volume() const { return extent[0] * extent[1] * extent[2]; }
};

// Codim-0 entity handle given to the parameter class.
struct Element {
  size_t index;
  Geometry geo;
  const Geometry& geometry() const { return geo; }
};

// Boundary intersection handle: side 0..5 = -x,+x,-y,+y,-z,+z (YaspGrid face order).
struct Intersection {
  int side;
  Element inside;
};

// ---- finite element map / function space / vectors ---------------------------
template <class D, class R, int k, int dim>
struct QkDGLocalFiniteElementMap {
  static_assert(dim == 3, "this build covers 3D hexahedral grids");
  static_assert(k >= 1 && k <= 3, "Q1..Q3");
  static constexpr int degree = k;
  static constexpr int size = (k + 1) * (k + 1) * (k + 1);
};

template <class FEM>
class GridFunctionSpace {
 public:
  using FiniteElementMapType = FEM;
  GridFunctionSpace(const StructuredGrid& g, const FEM& fem) : grid_(g), fem_(fem) {}
  size_t size() const { return grid_.size() * FEM::size; }
  size_t maxLocalSize() const { return FEM::size; }
  const StructuredGrid& gridView() const { return grid_; }
  const FEM& finiteElementMap() const { return fem_; }

 private:
  StructuredGrid grid_;
  FEM fem_;
};

// Device-resident coefficient vector in the cell-blocked DG layout
// (Backend::Vector<GFS, double>). Allocation goes through the operator handle.
class GridOperatorBase;

template <class T>
class DeviceVector {
 public:
  DeviceVector(pdelab_gpu_gridoperator* h, size_t n) : h_(h), n_(n) {
    void* p = nullptr;
    check(pdelab_gpu_malloc(h_, &p, n_ * sizeof(T)), h_);
    p_ = static_cast<T*>(p);
    *this = T(0);
  }
  DeviceVector(const DeviceVector& o) : DeviceVector(o.h_, o.n_) {
    check(pdelab_gpu_memcpy_dtod(h_, p_, o.p_, n_ * sizeof(T)), h_);
  }
  DeviceVector& operator=(const DeviceVector& o) {
    if (this != &o) check(pdelab_gpu_memcpy_dtod(h_, p_, o.p_, n_ * sizeof(T)), h_);
    return *this;
  }
  ~DeviceVector() {
    if (p_) pdelab_gpu_free(h_, p_);
  }
  // v = 0.0 (only zero-fill is supported on the device without an extra kernel)
  DeviceVector& operator=(T value) {
    if (value != T(0)) {
      std::vector<T> host(n_, value);
      upload(host);
    } else {
      check(pdelab_gpu_memset_zero(h_, p_, n_ * sizeof(T)), h_);
    }
    return *this;
  }
  void upload(const std::vector<T>& host) {
    if (host.size() != n_) throw Exception("DeviceVector::upload: size mismatch");
    check(pdelab_gpu_memcpy_htod(h_, p_, host.data(), n_ * sizeof(T)), h_);
  }
  std::vector<T> download() const {
    std::vector<T> host(n_);
    check(pdelab_gpu_memcpy_dtoh(h_, host.data(), p_, n_ * sizeof(T)), h_);
    return host;
  }
  size_t N() const { return n_; }
  T* data() { return p_; }
  const T* data() const { return p_; }
  pdelab_gpu_gridoperator* handle() const { return h_; }

 private:
  pdelab_gpu_gridoperator* h_;
  size_t n_;
  T* p_ = nullptr;
};

// ---- parameter class ------------------------------------------------------------
struct ConvectionDiffusionBoundaryConditions {
  enum Type { Dirichlet = 1, Neumann = -1, Outflow = -2, None = -3 };
};

// ConvectionDiffusionModelProblem with the reference defaults (A = I, b = 0, c = 0,
// f = 0, Dirichlet everywhere, g = |x|); derive and override like in PDELab.
struct ConvectionDiffusionModelProblem {
  static constexpr bool permeabilityIsConstantPerCell() { return true; }
  double A(const Element&, const std::array<double, 3>&) const { return 1.0; }  // kappa of A = kappa*I
  std::array<double, 3> b(const Element&, const std::array<double, 3>&) const { return {0.0, 0.0, 0.0}; }
  double c(const Element&, const std::array<double, 3>&) const { return 0.0; }
  double f(const Element&, const std::array<double, 3>&) const { return 0.0; }
  ConvectionDiffusionBoundaryConditions::Type bctype(const Intersection&, const std::array<double, 2>&) const {
    return ConvectionDiffusionBoundaryConditions::Dirichlet;
  }
  double g(const Element& e, const std::array<double, 3>& x) const {
    auto xg = e.geometry().global(x);
    return std::sqrt(xg[0] * xg[0] + xg[1] * xg[1] + xg[2] * xg[2]);
  }
  // Neumann flux (b u - A grad u).n = j and outflow data -(A grad u).n = o
  // (convectiondiffusionparameter.hh:121-124, 196-208), at face-local coordinates
  double j(const Intersection&, const std::array<double, 2>&) const { return 0.0; }
  double o(const Intersection&, const std::array<double, 2>&) const { return 0.0; }
};

struct ConvectionDiffusionDGMethod {
  enum Type { NIPG = PDELAB_NIPG, SIPG = PDELAB_SIPG, IIPG = PDELAB_IIPG };
};
struct ConvectionDiffusionDGWeights {
  enum Type { weightsOn = PDELAB_WEIGHTS_ON, weightsOff = PDELAB_WEIGHTS_OFF };
};

// ConvectionDiffusionDG(param, method = SIPG, weights = weightsOn, alpha = 1, intorderadd = 0)
template <class T, class FEM>
class ConvectionDiffusionDG {
 public:
  enum { doPatternVolume = true, doPatternSkeleton = true };
  enum { doAlphaVolume = true, doAlphaSkeleton = true, doAlphaBoundary = true, doLambdaVolume = true };
  static constexpr bool isLinear = true;

  explicit ConvectionDiffusionDG(T& param, ConvectionDiffusionDGMethod::Type method = ConvectionDiffusionDGMethod::SIPG,
                                 ConvectionDiffusionDGWeights::Type weights = ConvectionDiffusionDGWeights::weightsOn,
                                 double alpha = 1.0, int intorderadd = 0)
      : param_(param), method_(method), weights_(weights), alpha_(alpha), intorderadd_(intorderadd) {}

  T& param() const { return param_; }
  int method() const { return method_; }
  int weights() const { return weights_; }
  double alpha() const { return alpha_; }
  int intorderadd() const { return intorderadd_; }

 private:
  T& param_;
  ConvectionDiffusionDGMethod::Type method_;
  ConvectionDiffusionDGWeights::Type weights_;
  double alpha_;
  int intorderadd_;
};

namespace detail {
// Gauss-Legendre points and weights on [0,1] (m points)
inline void gauss01(int m, std::vector<double>& x, std::vector<double>& w) {
  x.assign(m, 0.0);
  w.assign(m, 0.0);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < m; ++i) {
    long double t = std::cos(pi * (i + 0.75L) / (m + 0.5L)), p0 = 1, p1 = t, dp = 1;
    for (int it = 0; it < 100; ++it) {
      p0 = 1; p1 = t;
      for (int j = 1; j < m; ++j) { long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1); p0 = p1; p1 = p2; }
      if (m == 1) { p1 = t; p0 = 1; }
      dp = m * (t * p1 - p0) / (t * t - 1.0L);
      long double dt = p1 / dp;
      t -= dt;
      if (std::fabs((double)dt) < 1e-30) break;
    }
    p0 = 1; p1 = t;
    for (int j = 1; j < m; ++j) { long double p2 = ((2 * j + 1) * t * p1 - j * p0) / (j + 1); p0 = p1; p1 = p2; }
    if (m == 1) { p1 = t; p0 = 1; }
    dp = m * (t * p1 - p0) / (t * t - 1.0L);
    x[m - 1 - i] = (double)((1.0L + t) / 2.0L);
    w[m - 1 - i] = (double)(1.0L / ((1.0L - t * t) * dp * dp));
  }
}
inline Element element(const StructuredGrid& g, size_t c) {
  int i = (int)(c % g.cells[0]), j = (int)((c / g.cells[0]) % g.cells[1]), l = (int)(c / ((size_t)g.cells[0] * g.cells[1]));
  Element e;
  e.index = c;
  const int ijk[3] = {i, j, l};
  for (int d = 0; d < 3; ++d) {
    e.geo.lower[d] = ijk[d] * g.h(d);
    e.geo.extent[d] = (ijk[d] + 1) * g.h(d) - e.geo.lower[d];
  }
  return e;
}
}  // namespace detail

// ---- GridOperator -----------------------------------------------------------------
// GridOperator<GFSU, GFSV, LOP> (matrix backend, field types and constraints of the
// reference template list are not needed on this path: DG imposes Dirichlet data
// weakly and the operator is never assembled).
template <class GFSU, class GFSV, class LOP>
class GridOperator {
 public:
  struct Traits {
    using Domain = DeviceVector<double>;
    using Range = DeviceVector<double>;
    using TrialGridFunctionSpace = GFSU;
    using TestGridFunctionSpace = GFSV;
  };
  using Domain = typename Traits::Domain;
  using Range = typename Traits::Range;

  // mirrors localAssembler().localOperator() / isLinear()
  struct LocalAssembler {
    const LOP& lop;
    const LOP& localOperator() const { return lop; }
    static constexpr bool isLinear() { return LOP::isLinear; }
  };

  GridOperator(const GFSU& gfsu, const GFSV& gfsv, const LOP& lop) : gfsu_(gfsu), gfsv_(gfsv), lop_(lop), la_{lop} {
    const StructuredGrid& grid = gfsu.gridView();
    const size_t nc = grid.size();
    auto& par = lop.param();
    pdelab_config cfg;
    pdelab_gpu_default_config(&cfg);
    cfg.nx = grid.cells[0]; cfg.ny = grid.cells[1]; cfg.nz = grid.cells[2];
    cfg.lx = grid.upper[0]; cfg.ly = grid.upper[1]; cfg.lz = grid.upper[2];
    cfg.degree = GFSU::FiniteElementMapType::degree;
    cfg.method = lop.method();
    cfg.weights = lop.weights();
    cfg.alpha = lop.alpha();
    cfg.intorderadd = lop.intorderadd();
    // A evaluated at cell centres (permeabilityIsConstantPerCell): a scalar kappa (A = kappa I)
    // or a 3x3 tensor (PermTensorType, convectiondiffusionparameter.hh:74-84), which must be
    // diagonal on this path (it then keeps the Kronecker structure)
    const std::array<double, 3> centre = {0.5, 0.5, 0.5};
    Element e0 = detail::element(grid, 0);
    using AType = decltype(par.A(e0, centre));
    constexpr bool scalarA = std::is_arithmetic<AType>::value;
    std::vector<double> kappa(nc * (scalarA ? 1 : 3));
    auto b0 = par.b(e0, centre);
    double c0 = par.c(e0, centre);
    for (size_t c = 0; c < nc; ++c) {
      Element e = detail::element(grid, c);
      if constexpr (scalarA) {
        kappa[c] = par.A(e, centre);
      } else {
        const auto A = par.A(e, centre);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            if (i != j && A[i][j] != 0.0)
              throw NotImplemented("GridOperator: the diffusion tensor must be diagonal on this path");
        for (int d = 0; d < 3; ++d) kappa[3 * c + d] = A[d][d];
      }
      if (par.c(e, centre) != c0) throw NotImplemented("GridOperator: c must be constant over the grid on this path");
    }
    cfg.kappa_per_cell = scalarA ? 1 : 2;
    // b at every quadrature point (convectiondiffusiondg.hh:178, 426, 797): a constant b is
    // passed in the configuration, a varying one as the tabulated velocity field
    tabulate_velocity(grid, par, 2 * cfg.degree + cfg.intorderadd);
    const bool bconst = velocity_constant(b0);
    for (int d = 0; d < 3; ++d) cfg.b[d] = bconst ? b0[d] : 0.0;
    cfg.c = c0;
    for (int s = 0; s < 6; ++s) {
      Intersection is{s, detail::element(grid, 0)};
      cfg.bctype[s] = par.bctype(is, {0.5, 0.5});
      cfg_bctype_[s] = cfg.bctype[s];
    }
    cfg.problem = PDELAB_PROBLEM_TABULATED;
    check(pdelab_gpu_create(&cfg, kappa.data(), &h_), nullptr);
    if (!bconst) check(pdelab_gpu_set_velocity_field(h_, bvol_.data(), bface_.data()), h_);
    std::vector<double>().swap(bvol_);
    std::vector<double>().swap(bface_);
    tabulate(grid, par);
  }
  ~GridOperator() {
    if (h_) pdelab_gpu_destroy(h_);
  }
  GridOperator(const GridOperator&) = delete;
  GridOperator& operator=(const GridOperator&) = delete;

  const GFSU& trialGridFunctionSpace() const { return gfsu_; }
  const GFSV& testGridFunctionSpace() const { return gfsv_; }
  const LocalAssembler& localAssembler() const { return la_; }
  pdelab_gpu_gridoperator* handle() const { return h_; }

  Domain makeVector() const { return Domain(h_, gfsu_.size()); }

  // r += R(x)   (gridoperator.hh:176-181)
  void residual(const Domain& x, Range& r) const { check(pdelab_gpu_residual(h_, x.data(), r.data()), h_); }
  // y += J z for a linear operator; throws otherwise (gridoperator.hh:192-197)
  void jacobian_apply(const Domain& z, Range& y) const {
    if (!LOP::isLinear) throw Exception("Your trying to use a non linear jacobian apply for a linear problem.");
    check(pdelab_gpu_jacobian_apply(h_, z.data(), y.data()), h_);
  }
  // nonlinear form; throws for a linear operator (gridoperator.hh:200-205)
  void jacobian_apply(const Domain& x, const Domain& z, Range& y) const {
    check(pdelab_gpu_jacobian_apply_nonlinear(h_, x.data(), z.data(), y.data()), h_);
  }

 private:
  // b(e, x) at the volume and face Gauss points of the rule of the given order, in the
  // layouts of pdelab_gpu_set_velocity_field: an interior face is evaluated from the cell
  // the reference's assembler visits it from, the one with the higher index (its lower face,
  // default/assembler.hh:181), a boundary face from its only cell.
  template <class P>
  void tabulate_velocity(const StructuredGrid& grid, const P& par, int order) {
    const int m = order / 2 + 1;
    std::vector<double> qx, qw;
    detail::gauss01(m, qx, qw);
    const size_t nc = grid.size();
    bvol_.resize(nc * m * m * m * 3);
    for (size_t c = 0; c < nc; ++c) {
      Element e = detail::element(grid, c);
      for (int q2 = 0; q2 < m; ++q2)
        for (int q1 = 0; q1 < m; ++q1)
          for (int q0 = 0; q0 < m; ++q0) {
            const auto b = par.b(e, {qx[q0], qx[q1], qx[q2]});
            const size_t o = 3 * (c * m * m * m + q0 + m * (q1 + m * q2));
            for (int d = 0; d < 3; ++d) bvol_[o + d] = b[d];
          }
    }
    bface_.clear();
    const int* n = grid.cells.data();
    for (int d = 0; d < 3; ++d) {
      const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      int ext[3] = {n[0], n[1], n[2]};
      ext[d] += 1;
      face_count_[d] = (size_t)ext[0] * ext[1] * ext[2] * m * m;
      for (int p2 = 0; p2 < ext[2]; ++p2)
        for (int p1 = 0; p1 < ext[1]; ++p1)
          for (int p0 = 0; p0 < ext[0]; ++p0) {
            int ijk[3] = {p0, p1, p2};
            const bool top = ijk[d] == n[d];  // the upper boundary face: its cell lies below
            if (top) ijk[d] -= 1;
            const size_t c = (size_t)ijk[0] + (size_t)n[0] * ((size_t)ijk[1] + (size_t)n[1] * ijk[2]);
            Element e = detail::element(grid, c);
            for (int q2 = 0; q2 < m; ++q2)
              for (int q1 = 0; q1 < m; ++q1) {
                std::array<double, 3> xi;
                xi[d] = top ? 1.0 : 0.0;
                xi[t1] = qx[q1];
                xi[t2] = qx[q2];
                bface_.push_back(par.b(e, xi)[d]);
              }
          }
    }
  }
  // whether every tabulated value equals b0 (then the constant-coefficient path applies)
  template <class B>
  bool velocity_constant(const B& b0) const {
    for (size_t i = 0; i < bvol_.size(); ++i)
      if (bvol_[i] != b0[i % 3]) return false;
    size_t o = 0;
    for (int d = 0; d < 3; ++d)
      for (size_t i = 0; i < face_count_[d]; ++i, ++o)
        if (bface_[o] != b0[d]) return false;
    return true;
  }

  template <class P>
  void tabulate(const StructuredGrid& grid, const P& par) {
    const int m = pdelab_gpu_quadrature_points(h_);
    std::vector<double> qx, qw;
    detail::gauss01(m, qx, qw);
    const size_t nc = grid.size();
    std::vector<double> fq(nc * m * m * m);
    for (size_t c = 0; c < nc; ++c) {
      Element e = detail::element(grid, c);
      for (int q2 = 0; q2 < m; ++q2)
        for (int q1 = 0; q1 < m; ++q1)
          for (int q0 = 0; q0 < m; ++q0)
            fq[c * m * m * m + q0 + m * (q1 + m * q2)] = par.f(e, {qx[q0], qx[q1], qx[q2]});
    }
    std::vector<double> gq;
    const int* n = grid.cells.data();
    for (int s = 0; s < 6; ++s) {
      const int d = s / 2, t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      for (int a2 = 0; a2 < n[t2]; ++a2)
        for (int a1 = 0; a1 < n[t1]; ++a1) {
          int ijk[3];
          ijk[d] = (s % 2) ? n[d] - 1 : 0;
          ijk[t1] = a1;
          ijk[t2] = a2;
          size_t c = (size_t)ijk[0] + (size_t)n[0] * ((size_t)ijk[1] + (size_t)n[1] * ijk[2]);
          Element e = detail::element(grid, c);
          const Intersection is{s, e};
          // the boundary value of the side's type: g (Dirichlet), j (Neumann), o (outflow)
          const int bct = cfg_bctype_[s];
          for (int q2 = 0; q2 < m; ++q2)
            for (int q1 = 0; q1 < m; ++q1) {
              std::array<double, 3> xi;
              xi[d] = (s % 2) ? 1.0 : 0.0;
              xi[t1] = qx[q1];
              xi[t2] = qx[q2];
              const std::array<double, 2> xf = {qx[q1], qx[q2]};
              if (bct == ConvectionDiffusionBoundaryConditions::Neumann) gq.push_back(par.j(is, xf));
              else if (bct == ConvectionDiffusionBoundaryConditions::Outflow) gq.push_back(par.o(is, xf));
              else gq.push_back(par.g(e, xi));
            }
        }
    }
    check(pdelab_gpu_set_tabulated_data(h_, fq.data(), gq.data()), h_);
  }

  const GFSU& gfsu_;
  const GFSV& gfsv_;
  const LOP& lop_;
  int cfg_bctype_[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> bvol_, bface_;  // the tabulated velocity field, until the handle has it
  size_t face_count_[3] = {0, 0, 0};  // values of bface_ per face direction
  LocalAssembler la_;
  pdelab_gpu_gridoperator* h_ = nullptr;
};

// ---- ISTL-backend shapes ---------------------------------------------------------------
struct SequentialNorm {
  template <class V>
  double norm(const V& v) const {
    double out = 0.0;
    check(pdelab_gpu_norm(v.handle(), v.data(), &out), v.handle());
    return out;
  }
};

template <class RF>
struct LinearSolverResult {
  bool converged = false;
  unsigned int iterations = 0;
  double elapsed = 0.0;
  RF reduction = 0.0;
  RF conv_rate = 0.0;
  void clear() { *this = LinearSolverResult(); }
};

class LinearResultStorage {
 public:
  const LinearSolverResult<double>& result() const { return res; }

 protected:
  LinearSolverResult<double> res;
  void store(const pdelab_linear_solver_result& r) {
    res.converged = r.converged != 0;
    res.iterations = r.iterations;
    res.elapsed = r.elapsed;
    res.reduction = r.reduction;
    res.conv_rate = r.conv_rate;
  }
};

// OnTheFlyOperator<X, Y, GO>: apply = (y = 0; go.jacobian_apply(x, y))
template <class X, class Y, class GO>
class OnTheFlyOperator {
 public:
  using domain_type = X;
  using range_type = Y;
  static constexpr bool isLinear = GO::LocalAssembler::isLinear();
  explicit OnTheFlyOperator(const GO& go) : go_(go) {}
  void setLinearizationPoint(const X& u) { u_ = &u; }
  void apply(const X& x, Y& y) const {
    if (isLinear) {
      // y = 0; jacobian_apply(x, y) in one pass
      check(pdelab_gpu_operator_apply(go_.handle(), x.data(), y.data()), go_.handle());
    } else {
      y = 0.0;
      if (!u_) throw InvalidStateException("You seem to apply a nonlinear operator without setting the linearization point first!");
      go_.jacobian_apply(*u_, x, y);
    }
  }
  void applyscaleadd(double alpha, const X& x, Y& y) const {
    Y temp(y);
    temp = 0.0;
    if (isLinear) go_.jacobian_apply(x, temp);
    else {
      if (!u_) throw InvalidStateException("You seem to apply a nonlinear operator without setting the linearization point first!");
      go_.jacobian_apply(*u_, x, temp);
    }
    check(pdelab_gpu_axpy(y.handle(), alpha, temp.data(), y.data()), y.handle());
  }

 private:
  const GO& go_;
  const X* u_ = nullptr;
};

enum class Krylov { CG = PDELAB_SOLVER_CG, BiCGSTAB = PDELAB_SOLVER_BICGSTAB };

// Preconditioner of the device Krylov backends: point Jacobi (Blocking::none), per-cell
// block Jacobi with exact block inverses (Blocking::fixed SeqJac; the matrix-free
// AssembledBlockJacobiPreconditionerLocalOperator), one block-SOR sweep
// (BlockSORPreconditionerLocalOperator) or block Jacobi with iteratively solved blocks
// (IterativeBlockJacobiPreconditionerLocalOperator); the last two with BiCGSTAB only.
enum class Preconditioner {
  Jacobi = PDELAB_PREC_JACOBI,
  BlockJacobi = PDELAB_PREC_BLOCK_JACOBI,
  BlockSOR = PDELAB_PREC_BLOCK_SOR,
  IterativeBlockJacobi = PDELAB_PREC_ITERATIVE_BLOCK_JACOBI
};

// BlockSolverOptions, R:dune/pdelab/backend/istl/matrixfree/iterativeblockjacobipreconditioner.hh:198-231
struct BlockSolverOptions {
  BlockSolverOptions(const double resreduction = 1e-5, const size_t maxiter = 100, const bool precondition = true,
                     const double weight = 1.0, const int verbose = 0)
      : _resreduction(resreduction), _maxiter(maxiter), _precondition(precondition), _weight(weight), _verbose(verbose) {}
  double _resreduction;
  size_t _maxiter;
  bool _precondition;
  double _weight;
  int _verbose;
};

// GridOperatorPreconditioner over a grid operator of
// AssembledBlockJacobiPreconditionerLocalOperator<BlockDiagonalLocalOperatorWrapper<LOP>>
// (R:dune/pdelab/backend/istl/matrixfree/gridoperatorpreconditioner.hh:18-91): apply(v, d)
// adds D^{-1} d, D the block diagonal of the grid operator's Jacobian. The blocks are
// inverted on first use (pre() is a no-op) and again after the next kappa change.
template <class GO>
class AssembledBlockJacobiPreconditioner {
  using V = typename GO::Traits::Domain;
  using W = typename GO::Traits::Range;

 public:
  static constexpr bool isLinear = true;
  explicit AssembledBlockJacobiPreconditioner(const GO& go) : go_(go) {}
  void pre(V&, W&) {}
  void apply(V& v, const W& d) { check(pdelab_gpu_block_jacobi_apply(go_.handle(), d.data(), v.data()), go_.handle()); }
  void post(V&) {}

 private:
  const GO& go_;
};

// GridOperatorPreconditioner over IterativeBlockJacobiPreconditionerLocalOperator
// (iterativeblockjacobipreconditioner.hh:267-450): apply(v, d) adds y with D_c y_c = d_c solved per
// cell by BiCGSTAB to the block-solver options. The options are the grid operator's: they also
// apply to its solves with Preconditioner::IterativeBlockJacobi. iterations() is the block
// solvers' iteration count of the last apply, summed over the cells (SolverStatistics).
template <class GO>
class IterativeBlockJacobiPreconditioner {
  using V = typename GO::Traits::Domain;
  using W = typename GO::Traits::Range;

 public:
  static constexpr bool isLinear = true;
  explicit IterativeBlockJacobiPreconditioner(const GO& go, BlockSolverOptions options = BlockSolverOptions())
      : go_(go) {
    check(pdelab_gpu_set_block_solver_options(go_.handle(), options._resreduction, (int)options._maxiter,
                                              options._precondition ? 1 : 0, options._weight),
          go_.handle());
  }
  void pre(V&, W&) {}
  void apply(V& v, const W& d) {
    check(pdelab_gpu_iterative_block_jacobi_apply(go_.handle(), d.data(), v.data(), &iterations_), go_.handle());
  }
  void post(V&) {}
  unsigned long long iterations() const { return iterations_; }

 private:
  const GO& go_;
  unsigned long long iterations_ = 0;
};

// GridOperatorPreconditioner over BlockSORPreconditionerLocalOperator
// (blocksorpreconditioner.hh:38-296) with exact block solves: apply(v, d) is one SOR step on v in
// place, v_c = (1 - omega) v_c + omega D_c^{-1} (d_c - sum_{j != c} A_cj v_j) with the cells in index
// order (the preconditioner of BiCGSTAB on its zeroed v). omega is the grid operator's: it also
// applies to its solves with Preconditioner::BlockSOR.
template <class GO>
class BlockSORPreconditioner {
  using V = typename GO::Traits::Domain;
  using W = typename GO::Traits::Range;

 public:
  static constexpr bool isLinear = true;
  explicit BlockSORPreconditioner(const GO& go, const double omega = 1.0) : go_(go) {
    check(pdelab_gpu_set_sor_relaxation(go_.handle(), omega), go_.handle());
  }
  void pre(V&, W&) {}
  void apply(V& v, const W& d) { check(pdelab_gpu_block_sor_apply(go_.handle(), d.data(), v.data()), go_.handle()); }
  void post(V&) {}

 private:
  const GO& go_;
};

// Matrix-free Krylov backend with device preconditioning: the role of
// ISTLBackend_SEQ_MatrixFree_Base<GO, PrecGO, Solver> with a point-diagonal or
// assembled-block-Jacobi preconditioner grid operator, and of ISTLBackend_SEQ_{CG,BCGS}_Jac
// on the same matrix (Blocking::none / Blocking::fixed).
template <class GO, Krylov method, Preconditioner prec = Preconditioner::Jacobi>
class ISTLBackend_SEQ_MatrixFree_Jac : public SequentialNorm, public LinearResultStorage {
  using V = typename GO::Traits::Domain;
  using W = typename GO::Traits::Range;

 public:
  static constexpr int krylov_method = (int)method;
  static constexpr int preconditioner = (int)prec;
  explicit ISTLBackend_SEQ_MatrixFree_Jac(const GO& go, unsigned maxiter = 5000, int verbose = 1)
      : go_(go), maxiter_(maxiter), verbose_(verbose) {}
  unsigned maxiter() const { return maxiter_; }
  // solves J z = r to the given relative reduction; r becomes the defect
  void apply(V& z, W& r, double reduction) {
    pdelab_linear_solver_result st;
    check(pdelab_gpu_solve(go_.handle(), (int)method, (int)prec, reduction, (int)maxiter_, z.data(), r.data(), &st),
          go_.handle());
    store(st);
  }
  void setLinearizationPoint(const V&) {}
  static constexpr bool isMatrixFree = true;

 private:
  const GO& go_;
  unsigned maxiter_;
  int verbose_;
};

template <class GO>
using ISTLBackend_SEQ_MatrixFree_CG_Jac = ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::CG>;
template <class GO>
using ISTLBackend_SEQ_MatrixFree_BCGS_Jac = ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::BiCGSTAB>;
template <class GO>
using ISTLBackend_SEQ_MatrixFree_CG_BlockJac = ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::CG, Preconditioner::BlockJacobi>;
template <class GO>
using ISTLBackend_SEQ_MatrixFree_BCGS_BlockJac =
    ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::BiCGSTAB, Preconditioner::BlockJacobi>;
// BiCGSTAB with the matrix-free block SOR and iterative block Jacobi preconditioners
// (matrix_free_linear.cc:185-250, 295-330: ISTLBackend_SEQ_MatrixFree_Base<GO, PrecGO, BiCGSTABSolver>)
template <class GO>
using ISTLBackend_SEQ_MatrixFree_BCGS_BlockSOR =
    ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::BiCGSTAB, Preconditioner::BlockSOR>;
template <class GO>
using ISTLBackend_SEQ_MatrixFree_BCGS_IterativeBlockJac =
    ISTLBackend_SEQ_MatrixFree_Jac<GO, Krylov::BiCGSTAB, Preconditioner::IterativeBlockJacobi>;

// Mixed precision backend: FP32 inner Krylov inside FP64 refinement with reliable updates
// (inner_reduction = FP32 residual reduction between FP64 residual replacements, 0.1 for CG,
// 1e-2 for BiCGSTAB); used in place of the linear problem solver (the refinement owns the
// residual evaluations).
template <class GO, Krylov method, Preconditioner prec = Preconditioner::Jacobi>
class MixedPrecisionSolver {
 public:
  MixedPrecisionSolver(const GO& go, double inner_reduction = method == Krylov::CG ? 0.1 : 1e-2,
                       unsigned inner_maxit = 20000)
      : go_(go), inner_(inner_reduction), inner_maxit_(inner_maxit) {}
  pdelab_mixed_result apply(typename GO::Traits::Domain& x, double reduction, int outer_maxit = 50) {
    pdelab_mixed_result r;
    check(pdelab_gpu_solve_mixed(go_.handle(), (int)method, (int)prec, inner_, (int)inner_maxit_, reduction,
                                 outer_maxit, x.data(), &r),
          go_.handle());
    return r;
  }

 private:
  const GO& go_;
  double inner_;
  unsigned inner_maxit_;
};

// StationaryLinearProblemSolver<GO, LS, V> (residual form: r = R(x); solve J z = r; x -= z)
template <class RF>
struct StationaryLinearProblemSolverResult : LinearSolverResult<RF> {
  RF first_defect = 0.0;
  RF defect = 0.0;
  double assembler_time = 0.0;
  double linear_solver_time = 0.0;
  int linear_solver_iterations = 0;
};

template <class GO, class LS, class V>
class StationaryLinearProblemSolver {
 public:
  StationaryLinearProblemSolver(const GO& go, LS& ls, V& x, double reduction, double min_defect = 1e-99,
                                int verbose = 1)
      : go_(go), ls_(ls), x_(&x), reduction_(reduction), min_defect_(min_defect), verbose_(verbose) {}
  void apply(V& x) {
    x_ = &x;
    apply();
  }
  void apply() {
    V r = go_.makeVector();
    auto t0 = std::chrono::steady_clock::now();
    go_.residual(*x_, r);  // residual is additive
    res_.assembler_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double defect = ls_.norm(r);
    V z = go_.makeVector();
    const double red = std::max(reduction_, min_defect_ / defect);
    auto t1 = std::chrono::steady_clock::now();
    ls_.apply(z, r, red);
    res_.linear_solver_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    const auto& lr = ls_.result();
    res_.converged = lr.converged;
    res_.iterations = lr.iterations;
    res_.elapsed = lr.elapsed;
    res_.reduction = lr.reduction;
    res_.conv_rate = lr.conv_rate;
    res_.first_defect = defect;
    res_.defect = defect * lr.reduction;
    res_.linear_solver_iterations = (int)lr.iterations;
    check(pdelab_gpu_axpy(x_->handle(), -1.0, z.data(), x_->data()), x_->handle());  // x -= z
    check(pdelab_gpu_synchronize(x_->handle()), x_->handle());
  }
  const StationaryLinearProblemSolverResult<double>& result() const { return res_; }
  double reduction() const { return reduction_; }
  const LS& linearSolver() const { return ls_; }
  using LinearSolver = LS;

 private:
  const GO& go_;
  LS& ls_;
  V* x_;
  double reduction_, min_defect_;
  int verbose_;
  StationaryLinearProblemSolverResult<double> res_;
};

// ---- instationary ------------------------------------------------------------------------
// Implicit one-step schemes (R:dune/pdelab/instationary/onestepparameter.hh).
template <class R>
struct ImplicitEulerParameter {
  int scheme() const { return PDELAB_TS_IMPLICIT_EULER; }
  unsigned s() const { return 1; }
};
template <class R>
struct Alexander2Parameter {
  int scheme() const { return PDELAB_TS_ALEXANDER2; }
  unsigned s() const { return 2; }
};
// theta = 1 (implicit Euler) and theta = 1/2 (Crank-Nicolson) are implemented
template <class R>
class OneStepThetaParameter {
 public:
  explicit OneStepThetaParameter(R theta) : theta_(theta) {
    if (theta != R(1) && theta != R(0.5)) throw NotImplemented("OneStepThetaParameter: theta must be 1 or 1/2");
  }
  int scheme() const { return theta_ == R(1) ? PDELAB_TS_IMPLICIT_EULER : PDELAB_TS_CRANK_NICOLSON; }
  unsigned s() const { return 1; }

 private:
  R theta_;
};

// The DG mass local operator (ScalarL2 / L2FastDG with the operator's quadrature).
struct L2 {
  explicit L2(int intorder = 2) : intorder_(intorder) {}
  int intorder_;
};

// OneStepGridOperator<GO0, GO1> (R:dune/pdelab/gridoperator/onestep.hh): the spatial grid
// operator and the DG mass. On the device the mass is the handle's own; GO1 only names it.
template <class GO0, class GO1>
class OneStepGridOperator {
 public:
  using Traits = typename GO0::Traits;
  OneStepGridOperator(const GO0& go0, const GO1&) : go0_(go0) {}
  pdelab_gpu_gridoperator* handle() const { return go0_.handle(); }
  const GO0& spatialGridOperator() const { return go0_; }
  typename Traits::Domain makeVector() const { return go0_.makeVector(); }
  // y += b_dt J z + a_mass M z (the stage Jacobian apply)
  void jacobian_apply(double a_mass, double b_dt, const typename Traits::Domain& z, typename Traits::Range& y) const {
    check(pdelab_gpu_onestep_jacobian_apply(handle(), a_mass, b_dt, z.data(), y.data()), handle());
  }

 private:
  const GO0& go0_;
};

// OneStepMethod<R, IGO, PDESOLVER, TrlV, TstV>::apply (R:dune/pdelab/instationary/implicitonestep.hh:144-230)
// for time-independent problems: each stage one StationaryLinearProblemSolver step with the
// pde solver's reduction and its backend's Krylov method and preconditioner.
template <class R, class IGO, class PDESOLVER, class TrlV, class TstV = TrlV>
class OneStepMethod {
 public:
  template <class Method>
  OneStepMethod(const Method& method, const IGO& igo, PDESOLVER& pdesolver)
      : scheme_(method.scheme()), igo_(igo), pdesolver_(pdesolver) {}
  R apply(R time, R dt, TrlV& xold, TrlV& xnew) {
    using LS = typename PDESOLVER::LinearSolver;
    pdelab_onestep_result r;
    check(pdelab_gpu_onestep_step(igo_.handle(), scheme_, LS::krylov_method, LS::preconditioner, (double)dt,
                                  pdesolver_.reduction(), (int)pdesolver_.linearSolver().maxiter(), xold.data(),
                                  xnew.data(), &r),
          igo_.handle());
    last_ = r;
    return time + dt;
  }
  const pdelab_onestep_result& result() const { return last_; }

 private:
  int scheme_;
  const IGO& igo_;
  PDESOLVER& pdesolver_;
  pdelab_onestep_result last_{};
};

// ---- error integration ------------------------------------------------------------------
// integral over the grid of (u_h - u)^2 for a host copy of the coefficients, with a Gauss
// rule of the given order: integrateGridFunction(DifferenceSquaredAdapter(u, u_h), order)
// (R:dune/pdelab/test/testconvectiondiffusiondg.cc:160-166).
template <class GFS, class F>
double l2_error_squared(const GFS& gfs, const std::vector<double>& x, F exact, int order = 10) {
  const StructuredGrid& grid = gfs.gridView();
  const int k = GFS::FiniteElementMapType::degree, n1 = k + 1, nloc = n1 * n1 * n1;
  const int m = order / 2 + 1;
  std::vector<double> qx, qw;
  detail::gauss01(m, qx, qw);
  // 1D Lagrange values at the points (nodes j/k)
  std::vector<double> P(m * n1);
  for (int q = 0; q < m; ++q)
    for (int i = 0; i < n1; ++i) {
      double r = 1.0;
      for (int j = 0; j <= k; ++j)
        if (j != i) r *= (k * qx[q] - j) / (double)(i - j);
      P[q * n1 + i] = r;
    }
  double err = 0.0;
  for (size_t c = 0; c < grid.size(); ++c) {
    Element e = detail::element(grid, c);
    const double det = e.geo.volume();
    for (int q2 = 0; q2 < m; ++q2)
      for (int q1 = 0; q1 < m; ++q1)
        for (int q0 = 0; q0 < m; ++q0) {
          double uh = 0.0;
          for (int i = 0; i < nloc; ++i)
            uh += x[c * nloc + i] * P[q0 * n1 + i % n1] * P[q1 * n1 + (i / n1) % n1] * P[q2 * n1 + i / (n1 * n1)];
          const double d = uh - exact(e.geo.global({qx[q0], qx[q1], qx[q2]}));
          err += d * d * qw[q0] * qw[q1] * qw[q2] * det;
        }
  }
  return err;
}

}  // namespace pdelab_amd

#endif
