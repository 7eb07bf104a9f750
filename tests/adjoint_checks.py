"""Adjoint gradient checks through the libtorch autograd Function
(dune-pdelab_amd/torch_ext), run as a child process by tests/test_gpu_adjoint.py so that
torch (and its HIP runtime) is the first to load in that process.

dL/dlog kappa for L = 1/2 (u - u_obs)^T M (u - u_obs), u = J(kappa)^{-1}(-R(0)), against
(a) the oracle's gradient (its own solves and the dual-number derivative of the
restated local operator) and (b) central finite differences of L with GPU solves.
"""
import os
import sys

import torch  # first: fixes the HIP runtime of this process

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import numpy as np  # noqa: E402
from conftest import import_package  # noqa: E402

pkg = import_package()
from dune_pdelab_amd import adjoint  # noqa: E402


def mass_quadratic(go):
    class MassQ(torch.autograd.Function):
        @staticmethod
        def forward(ctx, d):
            dd = d.detach().contiguous()
            md = torch.zeros_like(dd)
            torch.cuda.synchronize()
            assert pkg.lib().pdelab_gpu_mass_apply(go.h, dd.data_ptr(), md.data_ptr()) == 0
            go.synchronize()
            ctx.save_for_backward(md)
            return 0.5 * torch.dot(dd, md)

        @staticmethod
        def backward(ctx, g):
            (md,) = ctx.saved_tensors
            return g * md

    return MassQ.apply


def setup(n, k, seed):
    cfg = pkg.make_config(n=n, degree=k, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0, g0=0.0)
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 0.5, n ** 3)
    go = pkg.GridOperator(cfg, np.exp(theta))
    return cfg, go, theta, rng


def check_oracle():
    import oracle_ffi as of
    cfg, go, theta, rng = setup(6, 2, 21)
    uobs = rng.normal(0.0, 0.01, go.N)
    o = of.Oracle(cfg, np.exp(theta))
    x, _ = o.linear_problem_solve(reduction=1e-14, maxit=20000)
    lam, _, _ = o.solve(o.mass_apply(x - uobs), reduction=1e-14, maxit=20000, matrix_based=2)
    g_ref = o.grad_logkappa(x, lam)
    for precond, name in ((1, "point Jacobi"), (2, "block Jacobi")):
        t = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
        u = adjoint.dg_solve(t, go, reduction=1e-13, precond=precond)
        L = mass_quadratic(go)(u - torch.tensor(uobs, device="cuda"))
        L.backward()
        g = t.grad.detach().cpu().numpy()
        eu = np.linalg.norm(u.detach().cpu().numpy() - x) / np.linalg.norm(x)
        eg = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
        print(f"oracle check 6^3 Q2, {name}: u rel err {eu:.3e}, gradient rel err {eg:.3e}", flush=True)
        assert eu <= 1e-10 and eg <= 1e-8


def check_fd(n=32, ncheck=20):
    cfg, go, theta, rng = setup(n, 2, 31)
    theta_b = rng.normal(0.0, 0.5, n ** 3)
    with torch.no_grad():
        ub = adjoint.dg_solve(torch.tensor(theta_b, dtype=torch.float64, device="cuda"), go, reduction=1e-13)
    sigma = 0.01 * float(torch.sqrt(torch.mean(ub * ub)))
    uobs = ub + sigma * torch.tensor(rng.normal(size=go.N), device="cuda")
    mq = mass_quadratic(go)

    def J(th):
        with torch.no_grad():
            u = adjoint.dg_solve(torch.tensor(th, dtype=torch.float64, device="cuda"), go, reduction=1e-14)
            return float(mq(u - uobs))

    t = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    L = mq(adjoint.dg_solve(t, go, reduction=1e-14) - uobs)
    L.backward()
    g = t.grad.detach().cpu().numpy()
    cells = rng.choice(n ** 3, size=ncheck, replace=False)
    eps = 1e-4
    worst = 0.0
    for c in cells:
        tp = theta.copy(); tp[c] += eps
        tm = theta.copy(); tm[c] -= eps
        fd = (J(tp) - J(tm)) / (2 * eps)
        worst = max(worst, abs(fd - g[c]) / max(abs(fd), 1e-3 * np.abs(g).max()))
    print(f"FD check {n}^3 Q2 ({n ** 3} log-kappa entries): worst relative error over {ncheck} cells {worst:.3e}",
          flush=True)
    assert worst <= 1e-5


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("oracle", "all"):
        check_oracle()
    if which in ("fd", "all"):
        check_fd()
    print("ADJOINT CHECKS PASSED", flush=True)
