/*
 * pdelab_gpu.h -- C-ABI of the gfx950 matrix-free SIPG DG hot path.
 *
 * The entry points replace, one for one, the numeric calls that PDELab makes on
 * the path  StationaryLinearProblemSolver -> ISTL solver backend -> OnTheFlyOperator
 * -> GridOperator::jacobian_apply -> ConvectionDiffusionDG::alpha_* :
 *
 *   pdelab_gpu_residual              GridOperator::residual(x, r)
 *                                    R:dune/pdelab/gridoperator/gridoperator.hh:176-181
 *   pdelab_gpu_jacobian_apply        GridOperator::jacobian_apply(z, y)
 *                                    R:dune/pdelab/gridoperator/gridoperator.hh:192-197
 *   pdelab_gpu_operator_apply        OnTheFlyOperator::apply(x, y): y = 0, then
 *                                    jacobian_apply, fused into one pass
 *                                    R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:66-76
 *   pdelab_gpu_jacobian_apply_nonlinear
 *                                    GridOperator::jacobian_apply(x, z, y) (throws for a
 *                                    linear operator) gridoperator.hh:200-205
 *   pdelab_gpu_point_diagonal        residual of PointDiagonalLocalOperatorWrapper
 *                                    R:dune/pdelab/localoperator/pointdiagonalwrapper.hh:51-131
 *   pdelab_gpu_jacobian              GridOperator::jacobian(x, A) into the BCRS matrix of
 *                                    the 7-point block pattern
 *                                    R:dune/pdelab/gridoperator/gridoperator.hh:184-189
 *   pdelab_gpu_block_matrix_apply    MatrixAdapter::apply over that matrix (y = A z), the
 *                                    operator of the matrix-based ISTL backends
 *                                    R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:237-241
 *   pdelab_gpu_block_diagonal        jacobian of BlockDiagonalLocalOperatorWrapper
 *                                    R:dune/pdelab/localoperator/blockdiagonalwrapper.hh:79-160
 *   pdelab_gpu_block_jacobi_apply    GridOperatorPreconditioner::apply over
 *                                    AssembledBlockJacobiPreconditionerLocalOperator
 *                                    R:dune/pdelab/backend/istl/matrixfree/gridoperatorpreconditioner.hh:78-84,
 *                                    assembledblockjacobipreconditioner.hh:193-227
 *   pdelab_gpu_solve                 ISTLBackend_SEQ_{CG,BCGS}_Jac::apply /
 *                                    ISTLBackend_SEQ_MatrixFree_Base::apply(z, r, reduction)
 *                                    R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:231-252
 *                                    R:dune/pdelab/backend/istl/matrixfree/backends.hh:116-126
 *   pdelab_gpu_linear_problem_solve  StationaryLinearProblemSolver::apply
 *                                    R:dune/pdelab/stationary/linearproblem.hh:188-302
 *   pdelab_gpu_dot / _norm / _axpy   ISTL::BlockVector dot / two_norm / axpy
 *                                    R:dune/pdelab/backend/istl/vector.hh:249-282
 *   pdelab_gpu_onestep_jacobian_apply, pdelab_gpu_onestep_step
 *                                    OneStepGridOperator / OneStepMethod::apply
 *                                    R:dune/pdelab/gridoperator/onestep.hh,
 *                                    R:dune/pdelab/instationary/implicitonestep.hh:144-230
 *   pdelab_gpu_l2_error_squared      integrateGridFunction(DifferenceSquaredAdapter(u, u_h))
 *                                    R:dune/pdelab/common/functionutilities.hh:51-83
 *   pdelab_gpu_sample                DiscreteGridFunction values at the SubsamplingVTKWriter
 *                                    vertices, R:dune/pdelab/common/vtk.hh:655-690
 *   pdelab_gpu_mass_apply            ScalarL2::jacobian_apply_volume
 *                                    R:dune/pdelab/localoperator/l2.hh:45-90
 *
 * Extensions with no single reference counterpart (they compose the above):
 *   pdelab_gpu_jacobian_apply_f32, pdelab_gpu_solve_mixed (FP32 Krylov inside FP64
 *   iterative refinement), pdelab_gpu_jacobian_apply_transpose (adjoint operator),
 *   pdelab_gpu_grad_logkappa / pdelab_gpu_grad_source (adjoint parameter gradients).
 *
 * Conventions
 *   - Vectors are device pointers in PDELab's cell-blocked DG layout
 *     index = cell * (k+1)^3 + ix + (k+1)*iy + (k+1)^2*iz, cell = i + nx*(j + ny*l)
 *     (R:dune/pdelab/ordering/leaforderingbase.hh:97-141 with
 *      R:dune/pdelab/finiteelement/qkdglagrange.hh:118-128).
 *   - residual / jacobian_apply ACCUMULATE into their output (r += R(x), y += J z), as
 *     the reference engines do.
 *   - Every call is ordered on the handle's own HIP stream; calls that return host
 *     scalars (dot, norm, solve) synchronise that stream before returning.
 *   - No C++ exception crosses this boundary: each call returns a pdelab_status;
 *     pdelab_gpu_last_error() returns the message of the last failure on the handle.
 *   - Non-convergence of a Krylov solve is NOT an error: it is reported in
 *     pdelab_linear_solver_result.converged (R:dune/pdelab/backend/solver.hh:29-51).
 *   - Calls on one handle must be serialised by the caller; separate handles may be
 *     used from separate host threads.
 */
#ifndef PDELAB_GPU_H
#define PDELAB_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDELAB_GPU_ABI_VERSION 2
#define PDELAB_MAX_GAUSSIANS 16

/* Status codes. */
typedef enum pdelab_status {
  PDELAB_OK = 0,
  PDELAB_ERR_INVALID_ARG = 1,        /* bad config / null pointer / size */
  PDELAB_ERR_LINEARITY_MISMATCH = 2, /* gridoperator.hh:194-195, 202-203 */
  PDELAB_ERR_HIP = 3,                /* HIP runtime error */
  PDELAB_ERR_UNSUPPORTED = 4,        /* configuration outside the implemented path */
  PDELAB_ERR_SOLVER_ABORT = 5,       /* NaN/inf defect or BiCGSTAB breakdown (dune-istl SolverAbort) */
  PDELAB_ERR_OUTFLOW_ON_INFLOW = 6,  /* convectiondiffusiondg.hh:802-806 */
  PDELAB_ERR_NO_DEVICE = 7
} pdelab_status;

/* ConvectionDiffusionDGMethod::Type, R:dune/pdelab/localoperator/convectiondiffusiondg.hh:29-32 */
enum { PDELAB_NIPG = 0, PDELAB_SIPG = 1, PDELAB_IIPG = 2 };
/* ConvectionDiffusionDGWeights::Type, convectiondiffusiondg.hh:34-37 */
enum { PDELAB_WEIGHTS_ON = 0, PDELAB_WEIGHTS_OFF = 1 };
/* ConvectionDiffusionBoundaryConditions::Type,
   R:dune/pdelab/localoperator/convectiondiffusionparameter.hh:111-115 */
enum { PDELAB_BC_DIRICHLET = 1, PDELAB_BC_NEUMANN = -1, PDELAB_BC_OUTFLOW = -2, PDELAB_BC_NONE = -3 };

/* Source term f, Dirichlet data g, Neumann flux j and outflow data o of the parameter class
   (ConvectionDiffusionModelProblem::f / ::g / ::j / ::o, convectiondiffusionparameter.hh:174-208;
   (b u - A grad u).n = j on Neumann sides, -(A grad u).n = o on outflow sides, :121-124).
   The built-in problems are evaluated on the device at the quadrature points. The problems
   with an exact solution u (EXP, QUADRATIC) take its fluxes, j = (b.n) u - kappa0 grad(u).n
   and o = -kappa0 grad(u).n; the others the constants j0 and o0 of the config. */
enum {
  PDELAB_PROBLEM_ZERO = 0,      /* f = 0, g = 0 */
  PDELAB_PROBLEM_EXP = 1,       /* u = exp(-|x-(1/2,1/2,1/2)|^2); f = -kappa0*lap(u) + b.grad(u) + c u; g = u */
  PDELAB_PROBLEM_CONSTANT = 2,  /* f = f0, g = g0 */
  PDELAB_PROBLEM_GAUSSIANS = 3, /* f = sum_i a_i exp(-|x-x_i|^2/(2 s_i^2)); g = g0 */
  PDELAB_PROBLEM_QUADRATIC = 4, /* u = q0 + q.x + sum_d p_d x_d^2; f = -kappa0*lap(u) + b.grad(u) + c u; g = u */
  PDELAB_PROBLEM_TABULATED = 5  /* f, g given at the quadrature points, see pdelab_gpu_set_tabulated_data */
};

/* Krylov method and preconditioner of the ISTL backends. */
enum { PDELAB_SOLVER_CG = 0, PDELAB_SOLVER_BICGSTAB = 1 };
/* PDELAB_PREC_JACOBI: point Jacobi (Blocking::none); PDELAB_PREC_BLOCK_JACOBI: per-cell
   block Jacobi with exact block inverses (Blocking::fixed SeqJac,
   R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:240-243, and the matrix-free
   AssembledBlockJacobiPreconditionerLocalOperator,
   R:dune/pdelab/backend/istl/matrixfree/assembledblockjacobipreconditioner.hh:96-242). */
enum {
  PDELAB_PREC_NONE = 0,
  PDELAB_PREC_JACOBI = 1,
  PDELAB_PREC_BLOCK_JACOBI = 2,
  PDELAB_PREC_BLOCK_SOR = 3,
  PDELAB_PREC_ITERATIVE_BLOCK_JACOBI = 4
};
/* PDELAB_PREC_BLOCK_SOR: one forward block-SOR sweep over the cells in their index order from
   v = 0, v_i = omega D_i^{-1} (d_i - sum_{j<i} A_ij v_j) with the exact inverse D_i^{-1} of the
   cell's diagonal block (matrix-free BlockSORPreconditionerLocalOperator,
   R:dune/pdelab/backend/istl/matrixfree/blocksorpreconditioner.hh:38-296, applied by
   GridOperatorPreconditioner::apply, gridoperatorpreconditioner.hh:78-84, on the zeroed v of
   dune-istl's BiCGSTAB; the block solve of AssembledBlockJacobiPreconditionerLocalOperator).
   Not symmetric: BiCGSTAB only, J only (not J^T). omega is set by
   pdelab_gpu_set_sor_relaxation (default 1).
   PDELAB_PREC_ITERATIVE_BLOCK_JACOBI: block Jacobi whose blocks are solved iteratively, per cell
   dune-istl's BiCGSTAB on D_c y = d_c from y = 0, preconditioned by weight * diag(D_c)^{-1}, to
   the residual reduction and iteration limit of BlockSolverOptions (1e-5, 100 by default;
   IterativeBlockJacobiPreconditionerLocalOperator,
   R:dune/pdelab/backend/istl/matrixfree/iterativeblockjacobipreconditioner.hh:198-450).
   An inner iteration: BiCGSTAB only, J only. Options: pdelab_gpu_set_block_solver_options. */

typedef struct pdelab_config {
  int abi_version;            /* must be PDELAB_GPU_ABI_VERSION */
  /* YaspGrid box [0,lx]x[0,ly]x[0,lz] with nx*ny*nz hexahedra */
  int nx, ny, nz;
  double lx, ly, lz;
  /* QkDGLocalFiniteElementMap<.., k, 3> */
  int degree;                 /* k, 1..3 */
  /* ConvectionDiffusionDG(param, method, weights, alpha, intorderadd) */
  int method;
  int weights;
  double alpha;
  int intorderadd;            /* >= 0; affects only the load vector (f, g) quadrature */
  /* parameter class: A = kappa * I (constant per cell), b constant, c constant */
  int kappa_per_cell;         /* 0: A = kappa0 I everywhere; 1: A = kappa_c I, one value per cell given at
                                 create; 2: A = diag(a_x, a_y, a_z) per cell, three values per cell
                                 (the diffusion tensor of ConvectionDiffusionParameterInterface::A,
                                 R:dune/pdelab/localoperator/convectiondiffusionparameter.hh:74-84,
                                 constant per cell and diagonal: it keeps the Kronecker structure) */
  double kappa0;
  double b[3];
  double c;
  int bctype[6];              /* per boundary side -x,+x,-y,+y,-z,+z; 0 means Dirichlet */
  /* source / Dirichlet data */
  int problem;
  double f0, g0;              /* CONSTANT / GAUSSIANS g0 */
  double quad[7];             /* QUADRATIC: q0, q_x, q_y, q_z, p_x, p_y, p_z */
  int n_gaussians;
  double gaussians[PDELAB_MAX_GAUSSIANS][5]; /* x, y, z, sigma, amplitude */
  double j0, o0;              /* ZERO / CONSTANT / GAUSSIANS: Neumann flux j and outflow data o */
} pdelab_config;

/* LinearSolverResult<double>, R:dune/pdelab/backend/solver.hh:29-51 */
typedef struct pdelab_linear_solver_result {
  int converged;
  unsigned int iterations;
  double elapsed;             /* seconds */
  double reduction;           /* final defect / first defect */
  double conv_rate;           /* reduction^(1/iterations) */
} pdelab_linear_solver_result;

/* StationaryLinearProblemSolverResult<double>, R:dune/pdelab/stationary/linearproblem.hh:27-44 */
typedef struct pdelab_linear_problem_result {
  pdelab_linear_solver_result solver;
  double first_defect;
  double defect;
  double assembler_time;
  double linear_solver_time;
  int linear_solver_iterations;
} pdelab_linear_problem_result;

/* Result of the mixed-precision solve (FP32 inner Krylov, FP64 refinement). */
typedef struct pdelab_mixed_result {
  int converged;
  int outer_iterations;
  int inner_iterations;       /* summed over all refinement steps */
  double first_defect;        /* ||R(x0)||_2 in FP64 */
  double defect;              /* final ||R(x)||_2 in FP64 */
  double reduction;           /* defect / first_defect */
  double elapsed;             /* seconds */
} pdelab_mixed_result;

/* Implicit one-step schemes (R:dune/pdelab/instationary/onestepparameter.hh): implicit Euler
   (ImplicitEulerParameter), Crank-Nicolson (OneStepThetaParameter, theta = 1/2), Alexander's
   two-stage order-2 DIRK (Alexander2Parameter). */
enum { PDELAB_TS_IMPLICIT_EULER = 0, PDELAB_TS_CRANK_NICOLSON = 1, PDELAB_TS_ALEXANDER2 = 2 };

/* Result of one time step (OneStepMethodPartialResult). */
typedef struct pdelab_onestep_result {
  int converged;                /* every stage solve converged */
  int stages;
  int linear_solver_iterations; /* summed over the stages */
  double linear_solver_time;    /* seconds */
  double elapsed;               /* seconds, whole step */
} pdelab_onestep_result;

typedef struct pdelab_gpu_gridoperator pdelab_gpu_gridoperator;

/* ---- lifecycle ------------------------------------------------------------ */

/* Fills *cfg with the reference defaults: SIPG, weightsOn, alpha 1, intorderadd 0
   (convectiondiffusiondg.hh:85-90), model problem A=I, b=0, c=0, Dirichlet everywhere
   (convectiondiffusionparameter.hh:134-209), unit cube, 4^3 cells, k=1. */
void pdelab_gpu_default_config(pdelab_config* cfg);

/* Creates a grid operator on the current HIP device. kappa_host (host memory: nx*ny*nz doubles
   for kappa_per_cell == 1, 3*nx*ny*nz doubles a_x, a_y, a_z cell by cell for kappa_per_cell == 2;
   every value finite and > 0) is read when cfg->kappa_per_cell != 0, else ignored. */
int pdelab_gpu_create(const pdelab_config* cfg, const double* kappa_host,
                      pdelab_gpu_gridoperator** out);
int pdelab_gpu_destroy(pdelab_gpu_gridoperator* h);
/* Message of the last failed call on h (h may be NULL: last failed create). */
const char* pdelab_gpu_last_error(const pdelab_gpu_gridoperator* h);
/* Number of DoFs N = nx*ny*nz*(k+1)^3 and of cells. */
size_t pdelab_gpu_size(const pdelab_gpu_gridoperator* h);
size_t pdelab_gpu_num_cells(const pdelab_gpu_gridoperator* h);
/* isLinear of the wrapped local operator (always 1 for ConvectionDiffusionDG,
   convectiondiffusiondg.hh inherits LocalOperatorDefaultFlags::isLinear = true). */
int pdelab_gpu_is_linear(const pdelab_gpu_gridoperator* h);
/* Replaces the per-cell kappa (host array laid out as at create: nx*ny*nz doubles, or three per
   cell for kappa_per_cell == 2). */
int pdelab_gpu_set_kappa(pdelab_gpu_gridoperator* h, const double* kappa_host);
/* Same from a device array of log(kappa) (kappa = exp(log_kappa), computed on the device);
   A = kappa I only (PDELAB_ERR_UNSUPPORTED for kappa_per_cell == 2, as for the kappa gradient). */
int pdelab_gpu_set_log_kappa_device(pdelab_gpu_gridoperator* h, const double* log_kappa_dev);
/* Per-cell additive source s_c (f -> f + s_c on cell c); host array or NULL to clear. */
int pdelab_gpu_set_cell_source(pdelab_gpu_gridoperator* h, const double* source_host);
/* PDELAB_PROBLEM_TABULATED data, host arrays:
   f_q: per cell, volume quadrature points, x fastest: nx*ny*nz*m^3,
   g_q: per boundary face (side-major: -x,+x,-y,+y,-z,+z; within a side faces
        lexicographic in the two tangential cell indices), m^2 points each: the Dirichlet
        value g on Dirichlet sides, the Neumann flux j on Neumann sides, the outflow data
        o on outflow sides,
   with m = (2k + intorderadd)/2 + 1 Gauss points per direction. */
int pdelab_gpu_set_tabulated_data(pdelab_gpu_gridoperator* h, const double* f_q, const double* g_q);
int pdelab_gpu_quadrature_points(const pdelab_gpu_gridoperator* h); /* m */
/* Velocity field b(x) of the convection term, tabulated at the quadrature points where
   ConvectionDiffusionDG evaluates param.b (convectiondiffusiondg.hh:178 volume, :426 skeleton,
   :797 boundary; ConvectionDiffusionModelProblem::b, convectiondiffusionparameter.hh:159-165).
   Host arrays, with m = pdelab_gpu_quadrature_points:
   b_vol:  per cell (lexicographic, x fastest), the m^3 volume Gauss points (x fastest),
           (b_x, b_y, b_z) per point: nx*ny*nz*m^3*3 doubles;
   b_face: the component b_d along +e_d at the m^2 Gauss points of every face normal to d
           (lower tangential direction fastest): the faces normal to x, then y, then z, each
           direction's faces lexicographic in their position (x fastest; the position along
           d runs 0..n_d): ((nx+1)*ny*nz + nx*(ny+1)*nz + nx*ny*(nz+1))*m^2 doubles. On an
           interior face the reference evaluates b from the cell it visits the face from,
           the one with the higher index (default/assembler.hh:181).
   The field replaces the configuration's constant b, which must be zero, in the apply, the
   residual and its load vector (the Dirichlet inflow term, and the manufactured f and j of
   PDELAB_PROBLEM_EXP / _QUADRATIC), the point diagonal and the FP32 apply. b_vol = NULL
   restores the constant b. Outflow sides must see no inflow (b . n >= -1e-30 at every point,
   else PDELAB_ERR_OUTFLOW_ON_INFLOW). Requires intorderadd 0..3. The transposed operator
   (and with it the adjoint solves), the block diagonal, the block preconditioners (dense
   blocks per cell) and the assembled Jacobian take the field too. With a field set, CG
   returns PDELAB_ERR_UNSUPPORTED (the operator is not symmetric). */
int pdelab_gpu_set_velocity_field(pdelab_gpu_gridoperator* h, const double* b_vol, const double* b_face);

/* ---- device memory helpers (ordered on the handle's stream) --------------- */
int pdelab_gpu_malloc(pdelab_gpu_gridoperator* h, void** dptr, size_t bytes);
int pdelab_gpu_free(pdelab_gpu_gridoperator* h, void* dptr);
int pdelab_gpu_memcpy_htod(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes);
int pdelab_gpu_memcpy_dtoh(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes);
int pdelab_gpu_memcpy_dtod(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes);
int pdelab_gpu_memset_zero(pdelab_gpu_gridoperator* h, void* dptr, size_t bytes);
int pdelab_gpu_synchronize(pdelab_gpu_gridoperator* h);

/* ---- GridOperator --------------------------------------------------------- */
/* r += R(x) = alpha_volume + alpha_skeleton + alpha_boundary + lambda_volume terms */
int pdelab_gpu_residual(pdelab_gpu_gridoperator* h, const double* x, double* r);
/* y += J z (linear operator: the Jacobian apply is the residual without the load terms) */
int pdelab_gpu_jacobian_apply(pdelab_gpu_gridoperator* h, const double* z, double* y);
/* y = J z (overwrites y): the operator apply of the Krylov solvers, OnTheFlyOperator::apply. */
int pdelab_gpu_operator_apply(pdelab_gpu_gridoperator* h, const double* z, double* y);
/* FP32 operator apply, y += J z in single precision (arithmetic in FP32). */
int pdelab_gpu_jacobian_apply_f32(pdelab_gpu_gridoperator* h, const float* z, float* y);
/* Nonlinear form jacobian_apply(x, z, y): always PDELAB_ERR_LINEARITY_MISMATCH here,
   mirroring gridoperator.hh:202-203. */
int pdelab_gpu_jacobian_apply_nonlinear(pdelab_gpu_gridoperator* h, const double* x,
                                        const double* z, double* y);
/* y += J^T z (adjoint operator; convection becomes downwind). */
int pdelab_gpu_jacobian_apply_transpose(pdelab_gpu_gridoperator* h, const double* z, double* y);
/* d += diag(J) */
int pdelab_gpu_point_diagonal(pdelab_gpu_gridoperator* h, double* d);
/* blocks[c][i][j] = dR_i/dx_j of cell c within cell c: the diagonal blocks of J (of J^T
   if transpose), ncells * (k+1)^6 doubles, row-major per cell, device memory
   (overwritten). */
int pdelab_gpu_block_diagonal(pdelab_gpu_gridoperator* h, int transpose, double* blocks);
/* The Jacobian (of J^T if transpose) as a block matrix with the 7-point block pattern,
   device arrays: blocks[(7c + s) nloc^2 + i nloc + j] = dR_i(cell c) / dx_j(cell cols[7c + s]),
   nloc = (k+1)^3; slot 0 is cell c itself, slot 1 + f its neighbour across face f = 2d + side
   (x-, x+, y-, y+, z-, z+), cols = -1 and a zero block where that face is on the boundary.
   blocks holds 7 ncells nloc^2 doubles, cols 7 ncells entries; both are overwritten. */
int pdelab_gpu_jacobian(pdelab_gpu_gridoperator* h, int transpose, double* blocks, long long* cols);
/* y = A z for a block matrix in the layout of pdelab_gpu_jacobian (device arrays). */
int pdelab_gpu_block_matrix_apply(pdelab_gpu_gridoperator* h, const double* blocks, const long long* cols,
                                  const double* z, double* y);
/* y += D^{-1} z with D the block diagonal of J: the block Jacobi preconditioner as the
   Krylov solvers apply it with PDELAB_PREC_BLOCK_JACOBI. The blocks are inverted
   once (and again after the next kappa change). */
int pdelab_gpu_block_jacobi_apply(pdelab_gpu_gridoperator* h, const double* z, double* y);
/* FP32 form of pdelab_gpu_block_jacobi_apply (inverse blocks rounded to FP32). */
int pdelab_gpu_block_jacobi_apply_f32(pdelab_gpu_gridoperator* h, const float* z, float* y);
/* One block-SOR step on v in place, GridOperatorPreconditioner::apply(v, d) over
   BlockSORPreconditionerLocalOperator (R:dune/pdelab/backend/istl/matrixfree/
   gridoperatorpreconditioner.hh:78-84, blocksorpreconditioner.hh:221-250): cells in index
   order, v_c = (1 - omega) v_c + omega D_c^{-1} (d_c - sum_{j<c} A_cj v_j(new) -
   sum_{j>c} A_cj v_j(old)). On v = 0 this is the preconditioner W^{-1} d that BiCGSTAB applies
   with PDELAB_PREC_BLOCK_SOR (dune-istl zeroes v before apply). v must not alias d. One launch
   per hyperplane x + y + z of the grid; the dense inverse blocks of J are built once (and again
   after the next kappa change). */
int pdelab_gpu_block_sor_apply(pdelab_gpu_gridoperator* h, const double* d, double* v);
/* v += W^{-1} d for the iterative block Jacobi preconditioner W (PDELAB_PREC_ITERATIVE_BLOCK_JACOBI):
   per cell BiCGSTAB on D_c y_c = d_c from y_c = 0, added to v_c (GridOperatorPreconditioner::apply
   over IterativeBlockJacobiPreconditionerLocalOperator, R:dune/pdelab/backend/istl/matrixfree/
   iterativeblockjacobipreconditioner.hh:355-390, which adds y_tmp to y). v must not alias d.
   *iterations (when non-null) receives the block solvers' iteration count summed over the cells
   (SolverStatistics, solverstatistics.hh). A breakdown of a block solve returns
   PDELAB_ERR_SOLVER_ABORT, as the ISTL solver throws. */
int pdelab_gpu_iterative_block_jacobi_apply(pdelab_gpu_gridoperator* h, const double* d, double* v,
                                            unsigned long long* iterations);
/* BlockSolverOptions of the iterative block Jacobi preconditioner
   (iterativeblockjacobipreconditioner.hh:198-231): residual reduction (default 1e-5), iteration
   limit (100), point-Jacobi preconditioning of the block solves (1) and its weight (1).
   PDELAB_ERR_INVALID_ARG unless reduction > 0, maxiter >= 1 and weight finite and nonzero. */
int pdelab_gpu_set_block_solver_options(pdelab_gpu_gridoperator* h, double reduction, int maxiter, int precondition,
                                        double weight);
/* Relaxation factor omega of the block-SOR preconditioner (BlockSORPreconditionerLocalOperator's
   omega, default 1); finite and in (0, 2), else PDELAB_ERR_INVALID_ARG. */
int pdelab_gpu_set_sor_relaxation(pdelab_gpu_gridoperator* h, double omega);
/* y += M z, DG mass matrix (ScalarL2 with scaling 1). */
int pdelab_gpu_mass_apply(pdelab_gpu_gridoperator* h, const double* z, double* y);

/* ---- vector backend (FP64) ------------------------------------------------ */
int pdelab_gpu_dot(pdelab_gpu_gridoperator* h, const double* x, const double* y, double* out);
int pdelab_gpu_norm(pdelab_gpu_gridoperator* h, const double* x, double* out);
int pdelab_gpu_axpy(pdelab_gpu_gridoperator* h, double a, const double* x, double* y);

/* ---- solver backends ------------------------------------------------------ */
/* ISTL backend apply(z, r, reduction): solves J z = r starting from the given z with
   Krylov method `method` and preconditioner `precond`; r is overwritten by the final
   defect, as dune-istl does with its right-hand side. maxit <= 0 means 5000. */
int pdelab_gpu_solve(pdelab_gpu_gridoperator* h, int method, int precond, double reduction,
                     int maxit, double* z, double* r, pdelab_linear_solver_result* res);
/* Same on J^T (adjoint solve). */
int pdelab_gpu_solve_transpose(pdelab_gpu_gridoperator* h, int method, int precond,
                               double reduction, int maxit, double* z, double* r,
                               pdelab_linear_solver_result* res);
/* StationaryLinearProblemSolver::apply on x (device): r = R(x); solve J z = r to
   max(reduction, min_defect/||r||); x -= z. */
int pdelab_gpu_linear_problem_solve(pdelab_gpu_gridoperator* h, int method, int precond,
                                    double reduction, double min_defect, int maxit, double* x,
                                    pdelab_linear_problem_result* res);
/* Mixed precision: FP32 Krylov inside FP64 refinement until ||R(x)|| <= outer_reduction *
   ||R(x0)||: one FP32 CG or BiCGSTAB run with reliable updates; whenever its recursive
   residual has dropped by inner_reduction, the FP32 correction is subtracted from x in FP64
   and the residual replaced by the FP64 R(x), with the recurrence kept (CG's search direction;
   BiCGSTAB's direction, shadow residual and rho). 0.1 is a good interval for CG, 0.01 for
   BiCGSTAB; much smaller ones let the FP32 recursion drift and stall. outer_iterations counts
   the FP64 residual replacements. */
/* Variants of pdelab_gpu_solve_mixed, OR'd into its method argument:
   PDELAB_MIXED_RESTARTED  restarted refinement instead of reliable updates:
                           x <- x - J32^{-1} R(x), each FP32 solve to inner_reduction (or, when
                           weaker, to a quarter of the reduction still missing), a new Krylov
                           space per refinement step;
   PDELAB_MIXED_FLEXIBLE   flexible GMRES(10) in FP64 as the outer iteration, preconditioned
                           by the FP32 solve (CG or BiCGSTAB to inner_reduction from zero on a
                           unit-norm right-hand side); outer_iterations then counts FGMRES steps
                           and outer_maxit bounds them. */
enum { PDELAB_MIXED_RESTARTED = 16, PDELAB_MIXED_FLEXIBLE = 32 };
int pdelab_gpu_solve_mixed(pdelab_gpu_gridoperator* h, int method, int precond,
                           double inner_reduction, int inner_maxit, double outer_reduction,
                           int outer_maxit, double* x, pdelab_mixed_result* res);

/* ---- instationary (OneStepGridOperator over the operator and the DG mass) ---- */
/* y += b_dt J z + a_mass M z: the Jacobian apply of a stage of OneStepGridOperator<GO, GO_mass>
   (R:dune/pdelab/gridoperator/onestep.hh; weights b_rr dt on the spatial operator and a_rr on
   the mass, the MultiplyOperator0ByDT mode of gridoperator/onestep/localassembler.hh:122-126),
   with M the ScalarL2 mass of pdelab_gpu_mass_apply. */
int pdelab_gpu_onestep_jacobian_apply(pdelab_gpu_gridoperator* h, double a_mass, double b_dt, const double* z,
                                      double* y);
/* One step xold -> xnew of OneStepMethod::apply (R:dune/pdelab/instationary/implicitonestep.hh:144-230)
   for the time-independent problem of the handle: each stage r solves
     sum_{j<r} (b_rj dt R(u_j) + a_rj M u_j) + b_rr dt R(u_r) + a_rr M u_r = 0
   with one StationaryLinearProblemSolver step (Krylov method, preconditioner, reduction).
   Stage 1 starts from xnew as given, later stages from the previous stage. Device vectors. */
int pdelab_gpu_onestep_step(pdelab_gpu_gridoperator* h, int scheme, int method, int precond, double dt,
                            double reduction, int maxit, const double* xold, double* xnew,
                            pdelab_onestep_result* res);

/* ---- output and error integration ----------------------------------------- */
/* *err = integral over the grid of (u_h - u)^2, u the exact solution of the built-in problem
   (PDELAB_PROBLEM_EXP or _QUADRATIC), with the tensor Gauss rule of the given order (at most
   23): integrateGridFunction(DifferenceSquaredAdapter(u, u_h), order),
   R:dune/pdelab/common/functionutilities.hh:51-83, as testconvectiondiffusiondg.cc:160-166 uses it. */
int pdelab_gpu_l2_error_squared(pdelab_gpu_gridoperator* h, const double* x, int order, double* err);
/* pdelab_gpu_l2_error_squared without any allocation. No counterpart in the reference: it
   is for callers that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc). The per-cell partials live in a
   library-owned device array of PDELAB_NOALLOC_MAX_CELLS + 1 doubles (8 MiB + 8 B), one
   per GPU, held until the process ends; grids with more cells return
   PDELAB_ERR_INVALID_ARG. Every call on one GPU shares that array, so calls that overlap on
   one GPU (handles used from separate host threads) race. Same result as
   pdelab_gpu_l2_error_squared, bit for bit. */
#define PDELAB_NOALLOC_MAX_CELLS 1048576
int pdelab_gpu_l2_error_squared_noalloc(pdelab_gpu_gridoperator* h, const double* x, int order, double* err);
/* values[c (L+1)^3 + i + (L+1) j + (L+1)^2 l] = u_h(i/L, j/L, l/L) in cell c, L = intervals >= 1
   (at most 11): the DiscreteGridFunction values (gridfunctionspaceutilities.hh:54) at the
   vertices of the refined cells that SubsamplingVTKWriter with refinementIntervals(L) writes
   (R:dune/pdelab/common/vtk.hh:655-690). values is a device array of ncells (L+1)^3 doubles. */
int pdelab_gpu_sample(pdelab_gpu_gridoperator* h, const double* x, int intervals, double* values);

/* ---- adjoint gradients ---------------------------------------------------- */
/* g[c] = -lambda^T (dR/dlog kappa_c)(u) for every cell (device arrays, g has nx*ny*nz): the
   derivative of the residual R(u) = J u + R(0), i.e. of the operator and, with Dirichlet data
   g != 0, of the data terms that scale with kappa (R:dune/pdelab/localoperator/
   convectiondiffusiondg.hh:717-734, 839-877); weightsOn and weightsOff. */
int pdelab_gpu_grad_logkappa(pdelab_gpu_gridoperator* h, const double* u, const double* lambda,
                             double* g);
/* g[c] = lambda^T (d b / d s_c): the load-vector derivative for a per-cell source s_c. */
int pdelab_gpu_grad_source(pdelab_gpu_gridoperator* h, const double* lambda, double* g);

/* ---- measurement ---------------------------------------------------------- */
/* Runs `reps` back-to-back jacobian applies (precision 64 or 32) after `warmup`
   untimed ones and returns the average duration per launch in milliseconds,
   measured with HIP events on the handle's stream. */
int pdelab_gpu_time_apply(pdelab_gpu_gridoperator* h, int precision, const void* z, void* y,
                          int warmup, int reps, double* avg_ms);
/* Stream reference of the same box: `reps` back-to-back y += 0.5 x over the handle's
   vector length (FP64: 16 B read and 8 B written per entry) after `warmup` untimed ones;
   average duration per launch in milliseconds (HIP events, the handle's stream). No
   reference counterpart: it is the measured bandwidth ceiling the apply is compared with. */
int pdelab_gpu_time_stream(pdelab_gpu_gridoperator* h, const double* x, double* y, int warmup, int reps,
                           double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* PDELAB_GPU_H */
