// handle.hh -- the state behind one pdelab_gpu_gridoperator handle (internal to the
// library) and the per-degree apply launcher shared by its translation units.
#ifndef PDELAB_HANDLE_HH
#define PDELAB_HANDLE_HH

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pdelab_gpu.h"
#include "block_jacobi.hh"
#include "dg_kernels.hh"
#include "dg_load.hh"
#include "dg_tables.hh"
#include "dg_velocity.hh"
#include "output.hh"

using namespace pdelab_amd;

// Block Jacobi data of one operator (J or J^T): fast-diagonalisation entries or dense
// transposed inverse blocks, per cell or per boundary type.
struct BlockJacobiData {
  bool fdm = false, per_cell = false;
  bool valid64 = false, valid32 = false;
  long long entries = 0, esize = 0;
  double* d64 = nullptr;
  float* d32 = nullptr;
  // row-major copies of per-type dense blocks (bj_typed_kernel)
  double* rm64 = nullptr;
  float* rm32 = nullptr;
};

struct pdelab_gpu_gridoperator {
  pdelab_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Tables1D tab;
  DGArgs<double> a64;
  DGArgs<float> a32;
  LoadArgs la;
  GradArgs ga;
  long long ncells = 0, nloc = 0, N = 0;
  int n1 = 2;
  bool conv = false, vark = false;
  double* d_kappa = nullptr;
  float* d_kappa32 = nullptr;
  double* d_fc64 = nullptr;  // per-cell face coefficients (FC values per cell) of kappa
  float* d_fc32 = nullptr;
  double* d_R0 = nullptr;
  bool R0_valid = false;
  double* d_src = nullptr;
  double* d_ftab = nullptr;
  double* d_gtab = nullptr;
  double* d_diag = nullptr;
  double* d_dinv = nullptr;
  float* d_dinv32 = nullptr;
  bool diag_valid = false;
  double* w64[8] = {nullptr};
  float* w32[8] = {nullptr};
  double* d_part64 = nullptr;
  float* d_part32 = nullptr;
  double* d_scal64 = nullptr;
  float* d_scal32 = nullptr;
  double* h_scal64 = nullptr;
  float* h_scal32 = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double* d_applydot = nullptr;  // per-workgroup partials of the fused apply dot
  int* d_state = nullptr;        // device Krylov loop: {stopped, aborted, iteration, -}
  double* d_dstate = nullptr;    // {defect, def0, reduction, -}
  int* h_state = nullptr;        // pinned mirrors
  double* h_dstate = nullptr;
  double* d_bscal = nullptr;     // device BiCGSTAB loop: {rho, rho_new, alpha, omega, beta}
  // constant-coefficient line matrices: [0] J, [1] J^T
  LineMats<double>* d_lm64 = nullptr;
  LineMats<float>* d_lm32 = nullptr;
  bool sym[2] = {false, false};  // J / J^T have the structure of line_op_sym
  // apply kernel: 0 the row kernel (x-row chunks), 1 the general kernel (chunks of consecutive cells)
  int apply_variant = 0;
  // [0] J, [1] J^T, [2] J with dense inverse blocks (block SOR), [3] the diagonal blocks of J
  // themselves (iterative block Jacobi)
  BlockJacobiData bj[4];
  IBJOptions ibj_opt{1e-5, 100, 1, 1.0};
  void* d_ibj_stat = nullptr;  // int error code at 0, unsigned long long iteration sum at 8
  // reaction shift of the one-step stage operators: the kernels apply J + shift M
  double shift = 0.0;
  double prec_shift = 0.0;  // shift for which the cached diagonal and block Jacobi data hold
  double* os[4] = {nullptr, nullptr, nullptr, nullptr};  // one-step work vectors
  // flexible GMRES of the mixed-precision solve: the Arnoldi basis V (kFgmresRestart + 1
  // vectors) and the preconditioned directions Z (kFgmresRestart vectors), FP64
  static constexpr int kFgmresRestart = 10;
  double* fg[2 * kFgmresRestart + 1] = {nullptr};
  void* d_bjpart = nullptr;  // per-workgroup partials of the block Jacobi application
  // block SOR: cells ordered by hyperplane x + y + z, hyperplane s at
  // d_sor_order[sor_first[s] .. sor_first[s + 1]); relaxation factor omega
  int* d_sor_order = nullptr;
  std::vector<int> sor_first;
  double sor_omega = 1.0;
  // velocity field b(x) tabulated at the quadrature points (replaces the constant b of cfg):
  // device layouts of dg_velocity.hh, FP64 and FP32
  bool varb = false;
  double* d_bv64 = nullptr;
  double* d_bf64 = nullptr;
  float* d_bv32 = nullptr;
  float* d_bf32 = nullptr;
  VelArgs<double> va64;
  VelArgs<float> va32;
  VelArgs<double>* d_va64 = nullptr;  // device copy for the cell-block set-up
  double* d_sortmp64 = nullptr;       // block SOR with a field: the right-hand side less the couplings
  float* d_sortmp32 = nullptr;
};

using H = pdelab_gpu_gridoperator;

#define HIPCHK(h, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      (h)->err = std::string(#expr " failed: ") + hipGetErrorString(e_);             \
      return PDELAB_ERR_HIP;                                                         \
    }                                                                                \
  } while (0)

#define CHK(...)                     \
  do {                               \
    int s_ = (__VA_ARGS__);          \
    if (s_ != PDELAB_OK) return s_;  \
  } while (0)

inline long long apply_cpb(const H* h) {
  return h->n1 == 2 ? ApplyShape<2>::CPB : (h->n1 == 3 ? ApplyShape<3>::CPB : ApplyShape<4>::CPB);
}

// Whether an apply launch takes the row kernel (chunks of x-rows) rather than the general
// kernel (chunks of consecutive cells). Rows that split into whole chunks always do. The
// partial last chunk of a row is moved back to overlap the previous one (nx > CPB), and
// its overlapping lanes do work that is not stored, so such grids take the row kernel only
// for the constant-coefficient symmetric operator, where it is fastest (0.78 of the
// general kernel's time at Q3, 0.90 at Q2), with at most 1/8 (Q3) or 1/16 (Q2) of the
// row's lanes wasted. Decided per launch: set_kappa may change the operator after creation.
inline bool row_kernel(const H* h) {
  if (h->apply_variant != 0) return false;
  const long long cpb = apply_cpb(h), nx = h->a64.n[0];
  const long long span = (nx + cpb - 1) / cpb * cpb;
  if (span == nx) return true;
  if (nx < cpb || h->n1 == 2 || h->vark || h->conv || !h->sym[0] || !h->sym[1]) return false;
  return (span - nx) * (h->n1 == 4 ? 8 : 16) <= span;
}

// Chunks of one apply launch: the row kernel takes ceil(nx / CPB) chunks of every x-row,
// the general kernel ceil(ncells / CPB) chunks of consecutive cells.
inline long long apply_blocks_of(const H* h, bool row) {
  const long long cpb = apply_cpb(h);
  if (!row) return (h->ncells + cpb - 1) / cpb;
  return (h->a64.n[0] + cpb - 1) / cpb * h->a64.n[1] * h->a64.n[2];
}
inline long long apply_blocks(const H* h) { return apply_blocks_of(h, row_kernel(h)); }

// z-chunks walked by one workgroup of the row apply, for constant-coefficient symmetric
// operators (measured faster there; slower with cell-wise kappa): 2 at Q3; 8 at Q2, or 4
// where nz is a multiple of 4 but not of 8; else 1. The launch and the count of fused-dot
// partials both use it.
inline int apply_walk(const H* h) {
  if (!row_kernel(h) || h->vark || h->conv || !h->sym[0] || !h->sym[1]) return 1;
  const long long nz = h->a64.n[2];
  if (h->n1 == 4) return nz % 2 == 0 ? 2 : 1;
  if (h->n1 == 3) return nz % 8 == 0 ? 8 : (nz % 4 == 0 ? 4 : 1);
  return 1;
}

inline int fail(H* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (int)std::max(1LL, std::min(b, 4096LL * 8));
}

// per-cell face coefficients and operator arguments of the precision T
template <typename T>
inline const T* fc_for(H* h);
template <>
inline const double* fc_for<double>(H* h) { return h->d_fc64; }
template <>
inline const float* fc_for<float>(H* h) { return h->d_fc32; }
template <typename T>
inline const DGArgs<T>& args_for(H* h);
template <>
inline const DGArgs<double>& args_for<double>(H* h) { return h->a64; }
template <>
inline const DGArgs<float>& args_for<float>(H* h) { return h->a32; }

template <typename T>
inline const LineMats<T>* line_mats(H* h, bool trans);
template <>
inline const LineMats<double>* line_mats<double>(H* h, bool trans) { return h->d_lm64 + (trans ? 1 : 0); }
template <>
inline const LineMats<float>* line_mats<float>(H* h, bool trans) { return h->d_lm32 + (trans ? 1 : 0); }

// y (+)= J z (or J^T z) on the handle's stream for Q_{N1-1}: mode 0 y = J z, 1 y += J z,
// 2 y = J z with the fused z.(Jz) partials in dotp. Defined in apply_q<N1-1>.hip.
template <int N1, typename T>
int launch_apply_n(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool trans, int mode, T* dotp,
                   const int* stop);

// Block Jacobi (block_jacobi.hip): data set-up for the given precision (64 or 32),
// invalidation when kappa changes, the launches (BJ_PREC / BJ_ADD / BJ_CGR, see
// block_jacobi.hh; the partials go to d_bjpart, their count to *nparts) and the
// diagonal blocks themselves (ncells * nloc^2, row-major).
int bj_prepare(H* h, bool trans, int precision);
void bj_invalidate(H* h);
void bj_free(H* h);
template <typename T, typename A>
int bj_launch(H* h, bool trans, int mode, BJVec<T, A> v, int* nparts);
int bj_block_diagonal(H* h, bool trans, double* blocks);
// Block SOR (block_sor.hh): data set-up (dense blocks of J) and one sweep, from v = 0
// (v = W^{-1} d, the solvers) or in place from the v given (inplace); the kernels store
// nothing when *stop is set.
int sor_prepare(H* h, int precision);
template <typename T>
int sor_launch(H* h, const T* d, T* v, const int* stop, bool inplace);
// Iterative block Jacobi (bj_iter_kernel): the diagonal blocks of J, and one application
// v = W^{-1} d (v += W^{-1} d with add); the block solvers' error code goes to *err (kept at
// the first nonzero), their iteration sum is added to *iters when non-null.
int ibj_prepare(H* h, int precision);
template <typename T>
int ibj_launch(H* h, const T* d, T* v, int* err, unsigned long long* iters, const int* stop, bool add);
// Assembled Jacobian in the 7-point block pattern (jacobian.hh) and its apply y = A z.
int jac_assemble(H* h, bool trans, double* blocks, long long* cols);
int jac_block_apply(H* h, const double* blocks, const long long* cols, const double* z, double* y);
// Velocity field (velocity.hip): set or clear the tabulated b(x) (host arrays in the layout of
// pdelab_gpu_set_velocity_field), add its convection terms y += C z or y += C^T z (nothing
// stored when *stop is set) and their diagonal, free the data.
int vel_set(H* h, const double* b_vol, const double* b_face);
template <typename T>
int vel_launch(H* h, const T* z, T* y, const int* stop, bool trans);
int vel_diag(H* h, double* diag);
void vel_free(H* h);
// Block SOR with a velocity field: *out = d less the convection couplings of the cells of
// hyperplane d_sor_order[first .. first + count) (lower neighbours; all with inplace).
template <typename T>
int vel_sor_rhs(H* h, const T* d, const T* v, int first, int count, bool inplace, const T** out);

#endif
