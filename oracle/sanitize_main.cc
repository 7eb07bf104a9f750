// sanitize_main.cc -- TEST INFRASTRUCTURE: drives every entry point of the oracle on small
// grids, built with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
// `sanitize`, run by tests/test_sanitize.py). Any out-of-bounds access, use after free, leak
// or undefined operation aborts the run with a report; exit status 0 means none was found.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/pdelab_gpu.h"

struct oracle_handle;
extern "C" {
oracle_handle* oracle_create(const pdelab_config* cfg, const double* kappa);
void oracle_destroy(oracle_handle* h);
const char* oracle_last_error(oracle_handle* h);
size_t oracle_size(oracle_handle* h);
void oracle_set_threads(oracle_handle* h, int n);
int oracle_set_kappa(oracle_handle* h, const double* kappa);
int oracle_set_cell_source(oracle_handle* h, const double* s);
int oracle_set_tabulated(oracle_handle* h, const double* fq, const double* gq);
int oracle_quadrature_points(oracle_handle* h);
int oracle_set_velocity(oracle_handle* h, const double* bv, const double* bf);
int oracle_residual(oracle_handle* h, const double* x, double* r);
int oracle_residual_twosided(oracle_handle* h, const double* x, double* r, int nthreads);
int oracle_jacobian_apply(oracle_handle* h, const double* z, double* y);
int oracle_jacobian_apply_twosided(oracle_handle* h, const double* z, double* y, int nthreads);
int oracle_jacobian_apply_transpose(oracle_handle* h, const double* z, double* y);
int oracle_jacobian_apply_transpose_twosided(oracle_handle* h, const double* z, double* y, int nthreads);
int oracle_point_diagonal(oracle_handle* h, double* d);
int oracle_assemble(oracle_handle* h, double* blocks, long long* cols);
int oracle_block_diagonal(oracle_handle* h, double* blocks);
int oracle_mass_apply(oracle_handle* h, const double* z, double* y);
int oracle_interpolate(oracle_handle* h, double* x);
double oracle_l2_error_sq(oracle_handle* h, const double* x);
int oracle_grad_logkappa(oracle_handle* h, const double* u, const double* lam, double* g);
int oracle_grad_source(oracle_handle* h, const double* lam, double* g);
int oracle_solve(oracle_handle* h, int method, int precond, double reduction, int maxit, double* z, double* r,
                 pdelab_linear_solver_result* res, int matrix_based);
int oracle_linear_problem_solve(oracle_handle* h, int method, int precond, double reduction, double min_defect,
                                int maxit, double* x, pdelab_linear_problem_result* res, int matrix_based);
int oracle_onestep_jacobian_apply(oracle_handle* h, double a_mass, double b_dt, const double* z, double* y);
int oracle_onestep_step(oracle_handle* h, int scheme, int method, int precond, double dt, double reduction, int maxit,
                        const double* xold, double* xnew, pdelab_onestep_result* res);
double oracle_time_apply(oracle_handle* h, int nthreads, double min_seconds, int* reps_out);
int oracle_tables_1d(int k, int order, double* out);
}

static int failures = 0;
static void check(int st, oracle_handle* h, const char* what) {
  if (st != 0) {
    std::printf("FAIL %s: %s\n", what, h ? oracle_last_error(h) : "");
    ++failures;
  }
}

static pdelab_config base(int nx, int ny, int nz, int k) {
  pdelab_config c;
  std::memset(&c, 0, sizeof(c));
  c.abi_version = PDELAB_GPU_ABI_VERSION;
  c.nx = nx; c.ny = ny; c.nz = nz;
  c.lx = c.ly = c.lz = 1.0;
  c.degree = k;
  c.method = PDELAB_SIPG;
  c.weights = PDELAB_WEIGHTS_ON;
  c.alpha = 3.0;
  c.kappa0 = 1.0;
  c.problem = PDELAB_PROBLEM_EXP;
  return c;
}

static void exercise(const pdelab_config& c, bool cellwise) {
  const size_t nc = (size_t)c.nx * c.ny * c.nz;
  std::vector<double> kappa(nc);
  for (size_t i = 0; i < nc; ++i) kappa[i] = std::exp(std::sin(1.7 * i));
  oracle_handle* h = oracle_create(&c, cellwise ? kappa.data() : nullptr);
  if (!h) { std::printf("FAIL create\n"); ++failures; return; }
  const size_t n = oracle_size(h), nloc = n / nc;
  std::vector<double> x(n), y(n), z(n), r(n), d(n), g(nc);
  for (size_t i = 0; i < n; ++i) x[i] = std::cos(0.37 * i);
  check(oracle_residual(h, x.data(), r.data()), h, "residual");
  check(oracle_residual_twosided(h, x.data(), r.data(), 2), h, "residual_twosided");
  check(oracle_jacobian_apply(h, x.data(), y.data()), h, "jacobian_apply");
  check(oracle_jacobian_apply_twosided(h, x.data(), y.data(), 2), h, "jacobian_apply_twosided");
  check(oracle_jacobian_apply_transpose(h, x.data(), y.data()), h, "jacobian_apply_transpose");
  check(oracle_jacobian_apply_transpose_twosided(h, x.data(), y.data(), 2), h, "transpose_twosided");
  check(oracle_point_diagonal(h, d.data()), h, "point_diagonal");
  std::vector<double> blocks(nc * 7 * nloc * nloc), bd(nc * nloc * nloc);
  std::vector<long long> cols(nc * 7);
  check(oracle_assemble(h, blocks.data(), cols.data()), h, "assemble");
  check(oracle_block_diagonal(h, bd.data()), h, "block_diagonal");
  check(oracle_mass_apply(h, x.data(), y.data()), h, "mass_apply");
  check(oracle_interpolate(h, z.data()), h, "interpolate");
  (void)oracle_l2_error_sq(h, z.data());
  check(oracle_grad_logkappa(h, x.data(), z.data(), g.data()), h, "grad_logkappa");
  check(oracle_grad_source(h, x.data(), g.data()), h, "grad_source");
  check(oracle_onestep_jacobian_apply(h, 1.0, 0.1, x.data(), y.data()), h, "onestep_jacobian_apply");
  const bool spd = c.method == PDELAB_SIPG && c.b[0] == 0 && c.b[1] == 0 && c.b[2] == 0;
  for (int method : {PDELAB_SOLVER_CG, PDELAB_SOLVER_BICGSTAB}) {
    if (method == PDELAB_SOLVER_CG && !spd) continue;
    for (int prec : {PDELAB_PREC_NONE, PDELAB_PREC_JACOBI, PDELAB_PREC_BLOCK_JACOBI})
      for (int mb : {0, 1, 2}) {
        std::vector<double> zz(n, 0.0), rr = r;
        pdelab_linear_solver_result res;
        check(oracle_solve(h, method, prec, 1e-8, 60, zz.data(), rr.data(), &res, mb), h, "solve");
      }
    std::vector<double> xs(n, 0.0);
    pdelab_linear_problem_result lres;
    check(oracle_linear_problem_solve(h, method, PDELAB_PREC_JACOBI, 1e-8, 1e-99, 60, xs.data(), &lres, 0), h,
          "linear_problem_solve");
    for (int scheme : {PDELAB_TS_IMPLICIT_EULER, PDELAB_TS_CRANK_NICOLSON, PDELAB_TS_ALEXANDER2}) {
      std::vector<double> xn = x;
      pdelab_onestep_result ores;
      check(oracle_onestep_step(h, scheme, method, PDELAB_PREC_JACOBI, 0.01, 1e-8, 60, x.data(), xn.data(), &ores),
            h, "onestep_step");
    }
  }
  check(oracle_set_kappa(h, kappa.data()), h, "set_kappa");
  check(oracle_set_cell_source(h, g.data()), h, "set_cell_source");
  check(oracle_residual(h, x.data(), r.data()), h, "residual after set_kappa / set_cell_source");
  check(oracle_set_cell_source(h, nullptr), h, "clear cell source");
  int reps = 0;
  (void)oracle_time_apply(h, 2, 0.0, &reps);
  oracle_destroy(h);
}

int main() {
  std::vector<double> t(4096);
  for (int k = 1; k <= 3; ++k)
    for (int order = 1; order <= 12; ++order) check(oracle_tables_1d(k, order, t.data()) <= 0, nullptr, "tables_1d");
  // Q1, Q2 on 3x2x2 and 2x3x2 cells, Q3 on 2x1x2 and 1x2x2 (every face orientation, boundary
  // and interior, in each case)
  for (int k = 1; k <= 3; ++k) {
    const int m = k == 3 ? 1 : 0;
    pdelab_config c = base(3 - m, 2 - m, 2, k);
    exercise(c, false);
    if (k == 3) continue;  // the operator variants below at Q1 and Q2 only (time under the sanitizers)
    c.method = PDELAB_NIPG;
    c.b[0] = 1.0; c.b[1] = -0.5; c.b[2] = 0.25;
    c.c = 0.3;
    c.kappa_per_cell = 1;
    exercise(c, true);
    c = base(2 - m, 3 - m, 2, k);
    c.method = PDELAB_IIPG;
    c.weights = PDELAB_WEIGHTS_OFF;
    c.b[0] = 1.0;
    c.bctype[1] = PDELAB_BC_OUTFLOW;
    c.bctype[2] = PDELAB_BC_NEUMANN;
    c.problem = PDELAB_PROBLEM_GAUSSIANS;
    c.n_gaussians = 1;
    c.gaussians[0][0] = 0.5; c.gaussians[0][1] = 0.5; c.gaussians[0][2] = 0.5; c.gaussians[0][3] = 0.2;
    c.gaussians[0][4] = 1.0;
    exercise(c, false);
  }
  // tabulated data of the quadrature points
  {
    pdelab_config c = base(2, 2, 2, 2);
    c.problem = PDELAB_PROBLEM_TABULATED;
    c.intorderadd = 1;
    oracle_handle* h = oracle_create(&c, nullptr);
    const int m = oracle_quadrature_points(h);
    std::vector<double> fq(8 * m * m * m, 1.0), gq(6 * 4 * m * m, 0.5), x(oracle_size(h), 0.0), r(oracle_size(h));
    check(oracle_set_tabulated(h, fq.data(), gq.data()), h, "set_tabulated");
    check(oracle_residual(h, x.data(), r.data()), h, "residual tabulated");
    oracle_destroy(h);
  }
  // a velocity field at the quadrature points (3 x 2 x 2 cells, Q2, outflow on +x)
  {
    pdelab_config c = base(3, 2, 2, 2);
    c.bctype[1] = PDELAB_BC_OUTFLOW;
    c.bctype[2] = PDELAB_BC_NEUMANN;
    oracle_handle* h = oracle_create(&c, nullptr);
    const size_t m = oracle_quadrature_points(h), nf = 4 * 2 * 2 + 3 * 3 * 2 + 3 * 2 * 3;
    std::vector<double> bv(12 * m * m * m * 3), bf(nf * m * m, 0.75), x(oracle_size(h)), y(oracle_size(h)),
        d(oracle_size(h));
    for (size_t i = 0; i < bv.size(); ++i) bv[i] = 0.5 + 0.25 * std::sin(0.1 * (double)i);
    for (size_t i = 0; i < x.size(); ++i) x[i] = std::cos(0.3 * (double)i);
    check(oracle_set_velocity(h, bv.data(), bf.data()), h, "set_velocity");
    check(oracle_residual(h, x.data(), y.data()), h, "residual with a velocity field");
    check(oracle_jacobian_apply(h, x.data(), y.data()), h, "apply with a velocity field");
    check(oracle_jacobian_apply_twosided(h, x.data(), y.data(), 2), h, "two-sided apply with a velocity field");
    check(oracle_point_diagonal(h, d.data()), h, "point diagonal with a velocity field");
    check(oracle_set_velocity(h, nullptr, nullptr), h, "clear the velocity field");
    oracle_destroy(h);
  }
  if (failures) {
    std::printf("%d oracle calls failed\n", failures);
    return 1;
  }
  std::printf("sanitized oracle run: no errors\n");
  return 0;
}
