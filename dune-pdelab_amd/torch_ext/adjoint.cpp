// adjoint.cpp -- torch::autograd::Function around the DG solve of the HIP library.
//
// Forward: log kappa (one value per cell, float64 on the GPU) -> u solving J(kappa) u = -R(0)
// (StationaryLinearProblemSolver residual form on x0 = 0, CG with point Jacobi by default,
// or block Jacobi).
// Backward: the forward solution is recomputed (no Krylov history is stored), the
// adjoint system J^T lambda = dL/du is solved, and the per-cell kernel returns
// dL/dlog kappa_c = -lambda^T (dJ/dlog kappa_c) u.
#include <torch/extension.h>

#include "pdelab_gpu.h"

namespace {

pdelab_gpu_gridoperator* handle_of(int64_t h) { return reinterpret_cast<pdelab_gpu_gridoperator*>(h); }

void check(int st, pdelab_gpu_gridoperator* h) {
  TORCH_CHECK(st == PDELAB_OK, "pdelab_gpu error ", st, ": ", pdelab_gpu_last_error(h));
}

torch::Tensor solve_forward(pdelab_gpu_gridoperator* h, const torch::Tensor& logk, double reduction, int64_t maxit,
                            int64_t precond) {
  TORCH_CHECK(logk.is_cuda() && logk.scalar_type() == torch::kFloat64 && logk.is_contiguous(),
              "log kappa must be a contiguous float64 GPU tensor");
  TORCH_CHECK((size_t)logk.numel() == pdelab_gpu_num_cells(h), "log kappa must have one entry per cell");
  // The library runs on its own stream, which does not wait for torch's: torch's work on
  // the inputs is finished here, and every buffer the library reads is allocated without a
  // torch kernel (empty) and cleared on the library's stream.
  torch::cuda::synchronize();
  check(pdelab_gpu_set_log_kappa_device(h, logk.data_ptr<double>()), h);
  auto u = torch::empty({(int64_t)pdelab_gpu_size(h)}, logk.options());
  check(pdelab_gpu_memset_zero(h, u.data_ptr<double>(), sizeof(double) * (size_t)u.numel()), h);
  pdelab_linear_problem_result res;
  check(pdelab_gpu_linear_problem_solve(h, PDELAB_SOLVER_CG, (int)precond, reduction, 1e-99, (int)maxit,
                                        u.data_ptr<double>(), &res),
        h);
  TORCH_CHECK(res.solver.converged, "forward solve did not converge");
  return u;
}

class DGSolve : public torch::autograd::Function<DGSolve> {
 public:
  static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor logk, int64_t handle,
                               double reduction, int64_t maxit, int64_t precond) {
    auto* h = handle_of(handle);
    auto u = solve_forward(h, logk, reduction, maxit, precond);
    ctx->save_for_backward({logk});
    ctx->saved_data["handle"] = handle;
    ctx->saved_data["reduction"] = reduction;
    ctx->saved_data["maxit"] = maxit;
    ctx->saved_data["precond"] = precond;
    return u;
  }

  static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::variable_list grads) {
    auto logk = ctx->get_saved_variables()[0];
    auto* h = handle_of(ctx->saved_data["handle"].toInt());
    const double reduction = ctx->saved_data["reduction"].toDouble();
    const int64_t maxit = ctx->saved_data["maxit"].toInt();
    const int64_t precond = ctx->saved_data["precond"].toInt();
    // recompute the forward solution
    auto u = solve_forward(h, logk, reduction, maxit, precond);
    // adjoint solve J^T lambda = dL/du
    auto rhs = grads[0].contiguous().to(torch::kFloat64).clone();
    auto lam = torch::empty_like(rhs);
    auto g = torch::empty_like(logk);
    torch::cuda::synchronize();  // rhs is complete before the library's stream reads it
    check(pdelab_gpu_memset_zero(h, lam.data_ptr<double>(), sizeof(double) * (size_t)lam.numel()), h);
    check(pdelab_gpu_memset_zero(h, g.data_ptr<double>(), sizeof(double) * (size_t)g.numel()), h);
    pdelab_linear_solver_result res;
    check(pdelab_gpu_solve_transpose(h, PDELAB_SOLVER_CG, (int)precond, reduction, (int)maxit,
                                     lam.data_ptr<double>(), rhs.data_ptr<double>(), &res),
          h);
    TORCH_CHECK(res.converged, "adjoint solve did not converge");
    check(pdelab_gpu_grad_logkappa(h, u.data_ptr<double>(), lam.data_ptr<double>(), g.data_ptr<double>()), h);
    check(pdelab_gpu_synchronize(h), h);  // g is complete before torch's stream uses it
    return {g, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor dg_solve(torch::Tensor logk, int64_t handle, double reduction, int64_t maxit, int64_t precond) {
  TORCH_CHECK(precond == PDELAB_PREC_NONE || precond == PDELAB_PREC_JACOBI || precond == PDELAB_PREC_BLOCK_JACOBI,
              "precond must be PDELAB_PREC_NONE, PDELAB_PREC_JACOBI or PDELAB_PREC_BLOCK_JACOBI");
  return DGSolve::apply(logk, handle, reduction, maxit, precond);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("dg_solve", &dg_solve, "u = J(exp(log_kappa))^{-1} (-R(0)) with an adjoint backward",
        py::arg("log_kappa"), py::arg("handle"), py::arg("reduction") = 1e-12, py::arg("maxit") = 20000,
        py::arg("precond") = (int64_t)PDELAB_PREC_JACOBI);
}
