"""Benchmark of the matrix-free SIPG DG hot path on one MI355X.

Workload (one "step" = one operator apply y = J z over the whole vector, as the Krylov
solvers call it: OnTheFlyOperator::apply, i.e. y = 0 and GridOperator::jacobian_apply):
SIPG Poisson on the unit cube, 64^3 hexahedra, Q3 DG (64 DoF/cell, 16,777,216 DoF),
alpha 3, FP64, applied to a seeded uniform[-1,1) vector resident in HBM.

The JSON line also carries:
  roofline      algorithmic bytes per apply (16 B/DoF: read z, write y) over the
                HIP-event-measured average launch duration of the timed region, against
                8 TB/s HBM; burst_* give the same for 10 launches after a second of idle and
                two warm-up launches (the GPU clock drops under sustained FP64 load);
  roofline_q2, apply_variants
                the same HIP-event measurement for the 96^3 Q2 operators (constant kappa,
                cell-wise lognormal kappa, convection with a constant b and with a velocity
                field b(x) tabulated at the quadrature points) and the 64^3 Q3
                cell-wise-kappa operator, FP64 and FP32;
  cpu_baseline  the oracle (CPU restatement of the reference quadrature-loop
                algorithm, its -O3 build with FP contraction) timed on this host's cores on
                the same apply, and on one core;
                a projected single-core time of the full CG/Jacobi solve (GPU iteration
                count times the single-core apply, labelled projected); and the 16^3 Q2
                reference run (exp problem, CG/Jacobi to 1e-10) solved on the host cores,
                next to the same solve on the GPU (solve.c1_fp64_cg);
  solve         time to a 1e-10 relative residual: plain FP64 CG on the manufactured
                exp problem, and for the lognormal-kappa problem plain FP64 CG vs FP32
                CG with FP64 residual replacement (reliable updates), each with point
                Jacobi and with per-cell block Jacobi, and the stated configuration (restarted
                refinement, each FP32 CG solve to 1e-4); the oracle's FP64 residual of the
                64^3 solutions and of the 96^3 convection solutions (cpu_baseline leg);
  convection    96^3 Q2 convection-diffusion (kappa 1e-2, b = (1, .5, .25), eight seeded
                Gaussian sources, g = 0): FP64 BiCGSTAB/Jacobi vs FP32-inner BiCGSTAB in
                FP64 refinement to 1e-8, and the adjoint (J^T, downwind) gradient of
                1/2 u^T M u with respect to a per-cell source;
  adjoint_logkappa  32^3 Q2, log-kappa ~ N(0, 0.5^2) per cell: forward solve and backward
                (recomputed forward, adjoint solve, per-cell gradient kernel) of the gradient
                of 1/2 ||u - u_obs||_M^2, as the libtorch autograd Function runs them;
  matrix_based  32^3 Q2: the assembled Jacobian (7-point block pattern) and its block-matrix
                apply against the matrix-free apply;
  instationary  three Alexander2 steps of the heat equation at 64^3 Q3 (OneStepMethod).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-cpu]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "dune-pdelab_amd")
HBM_PEAK_GBS = 8000.0


def import_package():
    spec = importlib.util.spec_from_file_location("dune_pdelab_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["dune_pdelab_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def splitmix_uniform(n, seed):
    """uniform[-1,1) from splitmix64 (the counter generator the CPU oracle uses)."""
    s = (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(seed))
    t = s.copy()
    t ^= t >> np.uint64(30)
    t *= np.uint64(0xBF58476D1CE4E5B9)
    t ^= t >> np.uint64(27)
    t *= np.uint64(0x94D049BB133111EB)
    t ^= t >> np.uint64(31)
    return 2.0 * ((t >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def lognormal_kappa(ncells, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, 1.0, ncells))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kron_flops_per_dof(k):
    """FP64 flops per DoF of the implemented apply (constant kappa, no convection):
    per line of n1 = k+1 values, A u (n1 mul + n1(n1-1) FMA), two neighbour couplings
    (trace derivative: 1 mul + (n1-1) FMA, then 2 n1 FMA each), the S_d scaling (n1
    mul), one mass sweep (n1 mul + n1(n1-1) FMA), three directions, plus 2 n1 adds
    gathering the partial sums; an FMA counts 2 flops."""
    n1 = k + 1
    line = n1 + 2 * n1 * (n1 - 1) + 2 * (1 + 2 * (n1 - 1) + 4 * n1) + n1
    mass = n1 + 2 * n1 * (n1 - 1)
    return (3 * (line + mass) + 2 * n1) / n1


def cpu_baseline(cfg, min_seconds, gpu_cg_iterations=None, checks=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi

    cores = len(os.sched_getaffinity(0))
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    # timed with the -O3 build that allows FP contraction (the checker build keeps it off)
    build = os.path.basename(oracle_ffi.fast_lib_path())
    o = oracle_ffi.Oracle(cfg, timing=True)
    sec, reps = o.time_apply(nthreads=cores, min_seconds=min_seconds)
    sec1, reps1 = o.time_apply(nthreads=1, min_seconds=min_seconds / 2)
    n = oracle_ffi.cfgmod.ndofs(cfg)
    out = {"value": n / sec / 1e6, "unit": "MDoF/s", "cores": cores, "kind": "port",
           "sample": f"{reps} full jacobian_apply of the 64^3 Q3 FP64 workload (oracle quadrature-loop "
                     f"restatement, OpenMP over cells with two-sided faces), {sec * reps:.1f} s total",
           "build": f"{build}: g++ -O3 -ffp-contract=fast, -march=x86-64-v4 (AVX-512) or x86-64-v3, "
                    f"the widest this host supports",
           "single_core": {"value": n / sec1 / 1e6, "unit": "MDoF/s", "cores": 1,
                           "sample": f"{reps1} full jacobian_apply on one core, {sec1 * reps1:.1f} s total"}}
    if gpu_cg_iterations:
        # a single-core CG/Jacobi solve of this workload is infeasible in a run: projected from
        # the GPU solve's iteration count and the single-core apply (applies only, a lower bound)
        out["projected_single_core_solve"] = {
            "seconds": gpu_cg_iterations * sec1, "iterations": gpu_cg_iterations,
            "note": "projected: iterations of the GPU FP64 CG/Jacobi solve times the single-core apply; "
                    "vector operations excluded"}
    # the 16^3 Q2 reference run (exp problem, CG/Jacobi to 1e-10) on the host cores
    c1 = oracle_ffi.cfgmod.make_config(n=16, degree=2, alpha=3.0, problem=oracle_ffi.cfgmod.PROBLEM_EXP)
    o1 = oracle_ffi.Oracle(c1, timing=True)
    o1.set_threads(cores)
    t0 = time.perf_counter()
    x1, r1 = o1.linear_problem_solve(reduction=1e-10, maxit=5000)
    t1 = time.perf_counter()
    out["c1_solve"] = {"workload": "16^3 Q2 SIPG Poisson, exp manufactured solution, CG/Jacobi to 1e-10 "
                                   "(oracle, matrix-free, host cores)",
                       "seconds": t1 - t0, "cores": cores, "iterations": r1["solver"]["iterations"],
                       "converged": bool(r1["solver"]["converged"]),
                       "l2_error": float(np.sqrt(o1.l2_error_sq(x1)))}
    # error of the GPU solutions against the FP64 oracle: the oracle's residual of each returned
    # solution relative to its residual at x = 0 (the solves' first defect)
    for name, c in (checks or {}).items():
        oc = oracle_ffi.Oracle(c["cfg"], c.get("kappa"))
        t0 = time.perf_counter()
        r0 = float(np.linalg.norm(oc.residual_twosided(np.zeros(oc.N), cores)))
        res = {k: float(np.linalg.norm(oc.residual_twosided(v, cores))) / r0 for k, v in c["solutions"].items()}
        out[f"oracle_residual_{name}"] = {
            "note": f"||R_oracle(x)|| / ||R_oracle(0)|| of the GPU solutions of the {c['workload']} "
                    f"(target {c['target']:g}); R_oracle is the oracle's FP64 residual (two-sided faces, host cores)",
            "relative_residual": res, "gpu_first_defect_vs_oracle": abs(c["first_defect"] - r0) / r0,
            "seconds": time.perf_counter() - t0}
    return out


def swirl(x, y, z):
    """A divergence-free velocity field of the magnitude of the constant b = (1, 0.5, 0.25)."""
    return (1.0 - 0.5 * (y - 0.5), 0.5 + 0.5 * (x - 0.5), 0.25 + 0.0 * z)


APPLY_VARIANTS = [
    # (label, grid cells per direction, degree, config keywords, cell-wise lognormal kappa,
    #  velocity field tabulated at the quadrature points or None)
    ("q2_96_constant_kappa", 96, 2, dict(problem=1), False, None),
    ("q2_96_lognormal_kappa", 96, 2, dict(problem=2, f0=1.0), True, None),
    ("q2_96_convection", 96, 2, dict(kappa0=1e-2, b=(1.0, 0.5, 0.25)), False, None),
    ("q2_96_velocity_field", 96, 2, dict(kappa0=1e-2), False, swirl),
    ("q3_64_lognormal_kappa", 64, 3, dict(problem=2, f0=1.0), True, None),
]


def apply_variants_leg(pkg, reps=20):
    """FP64 and FP32 applies y = J z of the other operators on the path: the Q2 operators of the
    16^3 reference run, the adjoint and the convection-diffusion workloads at 96^3 (constant
    kappa, cell-wise lognormal kappa, convection), and the 64^3 Q3 cell-wise-kappa operator of the
    mixed-precision solve, and the 96^3 Q2 convection-diffusion operator with a velocity field
    b(x) tabulated at the quadrature points (two kernels: the apply with b = 0, then the
    convection terms). HIP-event average over `reps` launches after 3 warm-up launches.
    Algorithmic bytes: 16 B/DoF FP64 (8 B/DoF FP32), plus 8 B (4 B) per cell for cell-wise
    kappa, plus the field (3 m^3 + 6 m^2 values per cell) for the velocity field."""
    out = {}
    for label, n, k, kw, cellwise, field in APPLY_VARIANTS:
        cfg = pkg.make_config(n=n, degree=k, alpha=3.0, kappa_per_cell=cellwise, **kw)
        go = pkg.GridOperator(cfg, lognormal_kappa(n ** 3, 4242) if cellwise else None)
        N, nc = go.N, go.ncells
        fvals = 0
        if field is not None:
            bv, bf = pkg.tabulate_velocity(cfg, field)
            go.set_velocity_field(bv, bf)
            m = pkg.quadrature_points(cfg)
            fvals = 3 * m ** 3 + 6 * m ** 2
            del bv, bf
        z = go.vector(splitmix_uniform(N, 777))
        y = go.vector()
        ms64 = go.time_apply(z, y, warmup=3, reps=reps)
        z32 = go.vector(splitmix_uniform(N, 777).astype(np.float32), dtype=np.float32)
        y32 = go.vector(dtype=np.float32)
        ms32 = go.time_apply(z32, y32, warmup=3, reps=reps)
        b64 = 16.0 * N + (8.0 * nc if cellwise else 0.0) + 8.0 * fvals * nc
        b32 = 8.0 * N + (4.0 * nc if cellwise else 0.0) + 4.0 * fvals * nc
        out[label] = {"dofs": N, "fp64_us": ms64 * 1e3, "fp64_GBs": b64 / (ms64 * 1e-3) / 1e9,
                      "fp64_frac": b64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "fp64_MDoFs": N / (ms64 * 1e-3) / 1e6,
                      "fp32_us": ms32 * 1e3, "fp32_GBs": b32 / (ms32 * 1e-3) / 1e9,
                      "fp32_frac": b32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "algorithmic_bytes_fp64": b64, "algorithmic_bytes_fp32": b32}
        log(f"bench: apply {label}: {ms64 * 1e3:.1f} us fp64 ({out[label]['fp64_frac']:.3f} of HBM peak), "
            f"{ms32 * 1e3:.1f} us fp32")
        go.close()
    return out


def convection_leg(pkg, checks):
    """96^3 Q2 convection-diffusion solves and the source gradient (see the module doc)."""
    rng = np.random.default_rng(96)
    gauss = [(*rng.uniform(0.15, 0.85, 3), rng.uniform(0.05, 0.15), rng.uniform(0.5, 1.5)) for _ in range(8)]
    cfg = pkg.make_config(n=96, degree=2, alpha=3.0, kappa0=1e-2, b=(1.0, 0.5, 0.25), problem=pkg.PROBLEM_GAUSSIANS,
                          gaussians=gauss, g0=0.0)
    go = pkg.GridOperator(cfg)
    out = {"workload": "96^3 Q2 SIPG convection-diffusion, kappa 1e-2, b (1, 0.5, 0.25), 8 seeded Gaussian "
                       "sources, Dirichlet g = 0, 23,887,872 DoF"}
    x64 = go.vector()
    log("bench: 96^3 Q2 FP64 BiCGSTAB/Jacobi")
    r64 = go.linear_problem_solve(x64, method=pkg.SOLVER_BICGSTAB, reduction=1e-8, maxit=20000)
    t64 = r64["assembler_time"] + r64["linear_solver_time"]
    out["fp64_bicgstab"] = {"seconds": t64, "iterations": r64["solver"]["iterations"],
                            "converged": bool(r64["solver"]["converged"]), "reduction": r64["solver"]["reduction"]}
    xm = go.vector()
    log("bench: 96^3 Q2 mixed-precision BiCGSTAB (reliable updates)")
    rm = go.solve_mixed(xm, method=pkg.SOLVER_BICGSTAB, inner_reduction=1e-2, outer_reduction=1e-8,
                        inner_maxit=20000)
    a64, am = x64.download(), xm.download()
    checks["convection"] = {"cfg": cfg, "first_defect": rm["first_defect"], "target": 1e-8,
                            "workload": "96^3 Q2 convection-diffusion problem",
                            "solutions": {"fp64_bicgstab": a64, "mixed_bicgstab": am}}
    out["mixed_bicgstab"] = {"seconds": rm["elapsed"], "outer": rm["outer_iterations"],
                             "inner_iterations": rm["inner_iterations"], "converged": bool(rm["converged"]),
                             "reduction": rm["reduction"],
                             "rel_diff_vs_fp64_gpu": float(np.linalg.norm(am - a64) / np.linalg.norm(a64)),
                             "speedup_vs_fp64": t64 / rm["elapsed"] if rm["elapsed"] > 0 else None,
                             "note": "one FP32 BiCGSTAB run, FP64 residual replaced at each 1e-2 reduction"}
    xr = go.vector()
    log("bench: 96^3 Q2 mixed-precision BiCGSTAB (restarted refinement)")
    rr = go.solve_mixed(xr, method=pkg.SOLVER_BICGSTAB | pkg.MIXED_RESTARTED, inner_reduction=1e-4,
                        outer_reduction=1e-8, inner_maxit=20000)
    out["mixed_bicgstab_restarted"] = {"seconds": rr["elapsed"], "outer": rr["outer_iterations"],
                                       "inner_iterations": rr["inner_iterations"], "converged": bool(rr["converged"]),
                                       "reduction": rr["reduction"],
                                       "speedup_vs_fp64": t64 / rr["elapsed"] if rr["elapsed"] > 0 else None,
                                       "note": "each FP32 BiCGSTAB solve to 1e-4 from zero"}
    del xr
    # dJ/ds for J = 1/2 u^T M u: J^T lambda = M u, dJ/ds_c = lambda^T (int_c phi)
    log("bench: 96^3 Q2 adjoint source gradient")
    t0 = time.perf_counter()
    rhs = go.vector()
    go.mass_apply(x64, rhs)
    lam = go.vector()
    ra = go.solve(lam, rhs, method=pkg.SOLVER_BICGSTAB, reduction=1e-8, maxit=20000, transpose=True)
    g = go.cell_vector()
    go.grad_source(lam, g)
    go.synchronize()
    t1 = time.perf_counter()
    gh = g.download()
    out["adjoint_source_gradient"] = {"seconds": t1 - t0, "iterations": ra["iterations"],
                                      "converged": bool(ra["converged"]), "grad_norm": float(np.linalg.norm(gh))}
    for v in (x64, xm, rhs, lam, g):
        v.free()
    go.close()
    return out


def adjoint_leg(pkg):
    """32^3 Q2 SIPG, log-kappa per cell N(0, 0.5^2): the gradient of
    J = 1/2 (u - u_obs)^T M (u - u_obs) with respect to log-kappa, u_obs the solution at a
    second seeded field plus 1% Gaussian noise. Times the native calls of the autograd
    Function (dune-pdelab_amd/torch_ext/adjoint.cpp): the forward solve, then the backward
    (forward recomputed, adjoint solve J^T lambda = M (u - u_obs), per-cell gradient kernel),
    CG to 1e-12 with block and with point Jacobi (the Function's default)."""
    n, k = 32, 2
    rng = np.random.default_rng(32)
    theta, theta_b = rng.normal(0.0, 0.5, n ** 3), rng.normal(0.0, 0.5, n ** 3)
    cfg = pkg.make_config(n=n, degree=k, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0,
                          g0=0.0)
    go = pkg.GridOperator(cfg, np.exp(theta_b))
    ub = go.vector()
    go.linear_problem_solve(ub, precond=pkg.PREC_BLOCK_JACOBI, reduction=1e-12, maxit=20000)
    uh = ub.download()
    uobs = go.vector(uh + 0.01 * np.sqrt(np.mean(uh * uh)) * rng.normal(size=go.N))
    go.set_kappa(np.exp(theta))
    out = {"workload": "32^3 Q2 SIPG, log-kappa ~ N(0, 0.25) per cell (32,768 entries), J = 1/2 ||u - u_obs||_M^2, "
                       "u_obs from a second field plus 1% noise, CG to 1e-12"}
    for name, prec in (("block_jacobi", pkg.PREC_BLOCK_JACOBI), ("point_jacobi", pkg.PREC_JACOBI)):
        u, d, rhs, lam, g = go.vector(), go.vector(), go.vector(), go.vector(), go.cell_vector()
        go.synchronize()
        t0 = time.perf_counter()
        rf = go.linear_problem_solve(u, precond=prec, reduction=1e-12, maxit=20000)
        go.synchronize()
        t1 = time.perf_counter()
        u.free()
        u = go.vector()
        rb = go.linear_problem_solve(u, precond=prec, reduction=1e-12, maxit=20000)  # recomputed
        d.copy_from(u)
        go.axpy(-1.0, uobs, d)
        go.mass_apply(d, rhs)
        ra = go.solve(lam, rhs, precond=prec, reduction=1e-12, maxit=20000, transpose=True)
        go.grad_logkappa(u, lam, g)
        go.synchronize()
        t2 = time.perf_counter()
        out[name] = {"forward_seconds": t1 - t0, "backward_seconds": t2 - t1,
                     "forward_iterations": rf["solver"]["iterations"], "adjoint_iterations": ra["iterations"],
                     "converged": bool(rf["solver"]["converged"] and rb["solver"]["converged"] and ra["converged"]),
                     "grad_norm": float(np.linalg.norm(g.download()))}
        for v in (u, d, rhs, lam, g):
            v.free()
    for v in (ub, uobs):
        v.free()
    go.close()
    return out


def matrix_based_leg(pkg):
    """32^3 Q2 SIPG Poisson: the assembled Jacobian (7-point block pattern, GridOperator::jacobian)
    and its block-matrix apply y = A z against the matrix-free operator apply."""
    cfg = pkg.make_config(n=32, degree=2, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    N, nloc = go.N, 27
    z = go.vector(splitmix_uniform(N, 7))
    y = go.vector()
    blocks, cols = go.jacobian()  # first use loads the kernels; the second is timed
    blocks.free()
    cols.free()
    go.synchronize()
    t0 = time.perf_counter()
    blocks, cols = go.jacobian()
    go.synchronize()
    t_asm = time.perf_counter() - t0
    reps = 20
    go.block_matrix_apply(blocks, cols, z, y)
    go.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        go.block_matrix_apply(blocks, cols, z, y)
    go.synchronize()
    t_mb = (time.perf_counter() - t0) / reps
    t_mf = go.time_apply(z, y, warmup=3, reps=reps) * 1e-3
    mat_bytes = 7.0 * go.ncells * nloc * nloc * 8
    out = {"workload": "32^3 Q2 SIPG Poisson, 884,736 DoF", "assembly_seconds": t_asm,
           "matrix_bytes": mat_bytes, "block_matrix_apply_us": t_mb * 1e6,
           "block_matrix_apply_GBs": (mat_bytes + 16.0 * N) / t_mb / 1e9,
           "matrix_free_apply_us": t_mf * 1e6, "matrix_free_speedup": t_mb / t_mf}
    for v in (z, y, blocks, cols):
        v.free()
    go.close()
    return out


def instationary_leg(pkg, n, k):
    """Three Alexander2 steps (dt = 1e-3) of u_t - div(grad u) = f on the 64^3 Q3 exp problem from
    u = 0, each stage a CG/block-Jacobi solve to 1e-10 (OneStepMethod::apply)."""
    cfg = pkg.make_config(n=n, degree=k, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    xo, xn = go.vector(), go.vector()
    steps = []
    for _ in range(3):
        r = go.onestep_step(xo, xn, scheme=pkg.TS_ALEXANDER2, dt=1e-3, method=pkg.SOLVER_CG,
                            precond=pkg.PREC_BLOCK_JACOBI, reduction=1e-10)
        steps.append({"seconds": r["elapsed"], "linear_solver_iterations": r["linear_solver_iterations"],
                      "converged": bool(r["converged"])})
        xo.copy_from(xn)
    out = {"workload": f"{n}^3 Q{k} SIPG heat equation, Alexander2 (2 stages), dt 1e-3, CG/block Jacobi to 1e-10 "
                       f"per stage", "steps": steps}
    for v in (xo, xn):
        v.free()
    go.close()
    return out


# the 64^3 Q3 FP64 apply y = J z (constant kappa: the z-walk instance, template argument 2)
APPLY_KERNEL = "dg_apply_rowreg_kernel<4, double, false, false, false, false, false, false, 2>"
CALIB_KERNEL = "axpy_kernel<double>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "apply_traffic.json")


def _csv_rows(d, suffix):
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def profile_summary(d, n, k):
    """Summary (markdown) of the rocprofv3 runs of this benchmark kept under `d`:
    d/trace (--kernel-trace --stats), d/fetch (--pmc FETCH_SIZE), d/write (--pmc WRITE_SIZE).
    HBM bytes of the apply are the counters scaled by what they read for the axpy
    calibration launches (16 B/DoF read, 8 B/DoF written); also writes TRAFFIC_FILE."""
    N = n ** 3 * (k + 1) ** 3
    stats = _csv_rows(os.path.join(d, "trace"), "_kernel_stats.csv")
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["| kernel | calls | average (us) | total (%) |", "|---|---|---|---|"]
    for r in stats[:8]:
        lines.append(f"| `{r['Name'][:110]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['Percentage']):.1f} |")
    avg = {}
    for name, sub, cnt in (("fetch", "fetch", "FETCH_SIZE"), ("write", "write", "WRITE_SIZE")):
        for r in _csv_rows(os.path.join(d, sub), "_counter_collection.csv"):
            if r["Counter_Name"] != cnt:
                continue
            for tag, pat in (("apply", APPLY_KERNEL), ("calib", CALIB_KERNEL)):
                if pat in r["Kernel_Name"]:
                    avg.setdefault((name, tag), []).append(float(r["Counter_Value"]) * 1024.0)
    m = {key: sum(v) / len(v) for key, v in avg.items()}
    ff = 16.0 * N / m[("fetch", "calib")]
    fw = 8.0 * N / m[("write", "calib")]
    fetch = m[("fetch", "apply")] * ff
    write = m[("write", "apply")] * fw
    apply_avg = [float(r["AverageNs"]) for r in stats if APPLY_KERNEL in r["Name"]]
    out = {"kernel": APPLY_KERNEL, "n": n, "degree": k, "fetch_bytes": fetch, "write_bytes": write,
           "traffic_bytes": fetch + write, "algorithmic_bytes": 16.0 * N, "fetch_factor": ff, "write_factor": fw,
           "average_us": apply_avg[0] / 1e3 if apply_avg else None}
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(out, fh, indent=1)
    md = [f"Kernel time per launch (rocprofv3 --kernel-trace --stats), {n}^3 Q{k}:", ""] + lines + [
        "", f"HBM traffic of `{APPLY_KERNEL}` per launch (rocprofv3 --pmc, one counter per run):", "",
        "| counter | raw (KiB/launch) | scale (axpy calibration) | bytes/launch | algorithmic |", "|---|---|---|---|---|",
        f"| FETCH_SIZE | {m[('fetch', 'apply')] / 1024:.0f} | {ff:.3f} | {fetch:.4g} | {8.0 * N:.4g} |",
        f"| WRITE_SIZE | {m[('write', 'apply')] / 1024:.0f} | {fw:.3f} | {write:.4g} | {8.0 * N:.4g} |",
        "", f"traffic = {fetch + write:.4g} B/launch = {(fetch + write) / (16.0 * N):.3f} x algorithmic"]
    return "\n".join(md)


FP64_PEAK = 78.6e12  # MI355X dense FP64 (vector and matrix), FLOP/s
APPLY_WORKLOADS = {  # grid size (threads) of the apply launches of this benchmark -> workload
    "2097152": "64^3 Q3", "4194304": "64^3 Q3", "995328": "96^3 Q2", "7962624": "96^3 Q2",
    "56623104": "96^3 Q2"}  # the last: conv_var_kernel, one 64-lane wave per cell


def counter_summary(d):
    """Markdown table of the apply kernels' counters from the rocprofv3 passes kept under `d`
    (d/trace: --kernel-trace --stats; d/pmc*: --pmc passes; see the profiles/ summary for the
    counter sets). Per kernel: median duration, achieved HBM fraction (16 B/DoF FP64, 8 B/DoF
    FP32 over 8 TB/s), VALU busy (SQ_INSTS_VALU x 4 cycles over the 1024 SIMDs' cycles; GRBM_GUI_ACTIVE
    is summed over the 8 XCDs), FP64 utilisation ((2 FMA + ADD + MUL)_F64 x 64 lanes per wave
    instruction / duration / 78.6 TFLOP/s), LDS busy and bank-conflict cycles (per CU), SALU busy,
    instructions per wave, fraction of wave cycles waiting, mean resident waves per CU
    (SQ_WAVE_CYCLES counts quad-cycles: 4 SQ_WAVE_CYCLES / cycles / 256, rocprofv3's own
    OccupancyPercent formula), HBM
    traffic over algorithmic bytes (FETCH_SIZE x 2 and WRITE_SIZE x 1, the axpy calibration).
    conv_var_*kernel (the velocity field's convection terms, y += C z) is measured against its own
    algorithmic bytes: z read, y read and written (24 B/DoF FP64), and the field at the 27 volume
    and 54 face points of a Q2 cell (135 values per cell)."""
    import collections
    import re
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    meta = {}
    for r in _csv_rows(d, "_counter_collection.csv"):
        key = (r["Kernel_Name"], r["Grid_Size"])
        agg[key][r["Counter_Name"]].append(float(r["Counter_Value"]))
        meta[key] = r["VGPR_Count"]
    dur = collections.defaultdict(list)
    for r in _csv_rows(os.path.join(d, "trace"), "_kernel_trace.csv"):
        dur[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)

    def short(k):
        m = re.search(r"(dg_apply\w*kernel|conv_var\w*kernel)<[^(]*>", k)
        return m.group(0) if m else None

    out = ["| kernel | workload | us | HBM frac | VALU busy | FP64 util | LDS busy | LDS conflict | SALU busy "
           "| VALU/wave | LDS/wave | VMEM/wave | wait | waves/CU | traffic/alg |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for key, cs in sorted(agg.items(), key=lambda kv: short(kv[0][0]) or ""):
        s = short(key[0])
        if not s or key[1] not in APPLY_WORKLOADS:
            continue
        c = {k: sum(v) / len(v) for k, v in cs.items()}
        cyc = c["GRBM_GUI_ACTIVE"] / 8.0
        ts = sorted(dur.get(key[0], []))
        us = ts[len(ts) // 2] if ts else float("nan")
        wl = APPLY_WORKLOADS[key[1]]
        n, k = (64, 3) if "Q3" in wl else (96, 2)
        N = n ** 3 * (k + 1) ** 3
        bpd = 16.0 if "double" in s else 8.0
        if s.startswith("conv_var"):
            es = 8.0 if "double" in s else 4.0
            bpd = 3 * es + 135 * es / (k + 1) ** 3
        waves = c["SQ_WAVES"]
        fp64 = 2 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_MUL_F64"]
        traffic = (2.0 * c["FETCH_SIZE"] + c["WRITE_SIZE"]) * 1024.0 / (bpd * N)
        out.append(
            f"| `{s}` | {wl} | {us:.1f} | {bpd * N / (us * 1e-6) / 8e12:.2f} | "
            f"{c['SQ_INSTS_VALU'] * 4 / 1024 / cyc:.2f} | {fp64 * 64 / (us * 1e-6) / FP64_PEAK:.2f} | "
            f"{c['SQ_LDS_IDX_ACTIVE'] / 256 / cyc:.2f} | {c['SQ_LDS_BANK_CONFLICT'] / 256 / cyc:.2f} | "
            f"{c['SQ_INSTS_SALU'] / 1024 / cyc:.2f} | {c['SQ_INSTS_VALU'] / waves:.0f} | {c['SQ_INSTS_LDS'] / waves:.0f} | "
            f"{(c['SQ_INSTS_VMEM_RD'] + c['SQ_INSTS_VMEM_WR']) / waves:.0f} | {c['SQ_WAIT_ANY'] / c['SQ_WAVE_CYCLES']:.2f} | "
            f"{4.0 * c['SQ_WAVE_CYCLES'] / cyc / 256:.1f} | {traffic:.3f} |")
    return "\n".join(out)


def measured_traffic(n, k):
    """HBM bytes per apply launch from the committed rocprofv3 counter runs (or None)."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    return t["traffic_bytes"] if (t.get("n"), t.get("degree")) == (n, k) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-convection", action="store_true", help="skip the 96^3 Q2 convection-diffusion leg")
    ap.add_argument("--convection-only", action="store_true",
                    help="of the solve legs run only the 96^3 Q2 convection-diffusion one (for its profile)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the applies of the Q2 and cell-wise-kappa operators (apply_variants)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--inner-reduction", type=float, default=0.1,
                    help="mixed-precision CG: FP32 residual reduction between FP64 residual replacements")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run 5 axpy launches (reads 16 B/DoF, writes 8 B/DoF) to scale HBM counters")
    ap.add_argument("--counter-summary", metavar="DIR",
                    help="print the counter table of the apply kernels from rocprofv3 passes under DIR and exit")
    ap.add_argument("--profile-summary", metavar="DIR",
                    help="print the summary of the rocprofv3 runs of this benchmark kept under DIR and exit")
    args = ap.parse_args()
    if args.counter_summary:
        print(counter_summary(args.counter_summary))
        return
    if args.profile_summary:
        print(profile_summary(args.profile_summary, args.n, args.degree))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this library runs on one GPU only; a multi-GPU run (domain decomposition "
              "over ranks) is not implemented.", file=sys.stderr)
        sys.exit(2)

    pkg = import_package()
    n, k = args.n, args.degree
    cfg = pkg.make_config(n=n, degree=k, alpha=3.0, problem=pkg.PROBLEM_EXP)
    log(f"bench: creating 64^3-style grid n={n} k={k}")
    go = pkg.GridOperator(cfg)
    N = go.N
    z = go.vector(splitmix_uniform(N, 20261202))
    y = go.vector()

    # ---- timed region: K operator applies (y = J z, the same kernel as
    # OnTheFlyOperator::apply) back to back on the handle's stream, inputs resident in HBM.
    # HIP events on that stream bracket the K launches, so the per-launch average that the
    # roofline uses comes from the timed launches themselves.
    for _ in range(args.warmup):
        go.operator_apply(z, y)
    go.synchronize()
    # a burst of 10 launches after a second of idle, for context: sustained FP64 load lowers
    # the clock, and the timed region below (the reported figure) is the sustained rate
    time.sleep(1.0)
    burst_ms = go.time_apply(z, y, warmup=2, reps=10)  # the warm-up launches absorb the wake-up
    go.synchronize()
    t0 = time.perf_counter()
    kern_ms = go.time_apply(z, y, warmup=0, reps=args.steps)
    go.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N * args.steps / (t1 - t0) / 1e6
    log(f"bench: {args.steps} applies, {ms_per_step:.4f} ms/apply, {value:.1f} MDoF/s")
    alg_bytes = 16.0 * N
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    z32 = go.vector(splitmix_uniform(N, 20261202).astype(np.float32), dtype=np.float32)
    y32 = go.vector(dtype=np.float32)
    kern32_ms = go.time_apply(z32, y32, warmup=3, reps=max(20, args.steps))
    stream_ms = go.time_stream(z, y, warmup=3, reps=20)
    stream_gbs = 24.0 * N / (stream_ms * 1e-3) / 1e9
    log(f"bench: kernel avg {kern_ms * 1e3:.1f} us (fp64), {kern32_ms * 1e3:.1f} us (fp32); "
        f"stream reference {stream_gbs:.0f} GB/s")
    z32.free()
    y32.free()
    variants = None
    if not args.no_variants:
        variants = apply_variants_leg(pkg)
    if args.calibrate:
        for _ in range(5):
            go.axpy(0.5, z, y)
        go.synchronize()

    out = {
        "metric": "MDoF/s DG apply; solve time to 1e-10 rel err, mixed-prec vs FP64, 1 GPU",
        "value": value,
        "unit": "MDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SIPG Poisson -div(grad u)=f, unit cube, YaspGrid-style {n}^3 hexahedra, "
                               f"Q{k} DG ({(k + 1) ** 3} DoF/cell, {N} DoF), alpha 3, FP64 jacobian_apply "
                               f"on a seeded uniform[-1,1) vector",
                   "cells": n ** 3, "degree": k, "dofs": N},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n, k),
                     "kernel": (f"dg_apply_rowreg_kernel<{k + 1},double,...>" if n % {1: 64, 2: 32, 3: 16}[k] == 0
                                else f"dg_apply_kernel<{k + 1},double,...>"), "kernel_avg_us": kern_ms * 1e3,
                     "burst_kernel_avg_us": burst_ms * 1e3, "burst_frac": alg_bytes / (burst_ms * 1e-3) / 1e9 / 8000.0,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "fp32_kernel_avg_us": kern32_ms * 1e3,
                     "fp32_achieved_GBs": 8.0 * N / (kern32_ms * 1e-3) / 1e9,
                     "stream_GBs": stream_gbs, "frac_of_stream": achieved / stream_gbs,
                     "flops_per_dof": kron_flops_per_dof(k),
                     "achieved_fp64_TFLOPs": kron_flops_per_dof(k) * N / (kern_ms * 1e-3) / 1e12},
    }
    if variants:
        q2 = variants["q2_96_constant_kappa"]
        out["roofline_q2"] = {"bound": "hbm", "kernel": "dg_apply_rowreg_kernel<3,double,...>",
                              "workload": "96^3 Q2 SIPG, constant kappa, FP64 y = J z",
                              "achieved": q2["fp64_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": q2["fp64_frac"], "kernel_avg_us": q2["fp64_us"],
                              "algorithmic_bytes_per_launch": q2["algorithmic_bytes_fp64"]}
        out["apply_variants"] = variants

    checks = {}  # solutions whose oracle residual the cpu_baseline leg evaluates
    if args.convection_only and not args.no_solve:
        z.free()
        y.free()
        go.close()
        out["convection"] = convection_leg(pkg, checks)
    elif not args.no_solve:
        solve = {}
        # plain FP64 CG/Jacobi on the manufactured problem (64^3 Q3)
        x = go.vector()
        log("bench: FP64 CG/Jacobi solve (exp problem)")
        res = go.linear_problem_solve(x, method=pkg.SOLVER_CG, reduction=1e-10)
        solve["exp_fp64_cg"] = {"seconds": res["solver"]["elapsed"], "iterations": res["solver"]["iterations"],
                                "converged": bool(res["solver"]["converged"]),
                                "reduction": res["solver"]["reduction"]}
        # the 16^3 Q2 reference run on the GPU (the CPU one is in cpu_baseline)
        c1 = pkg.make_config(n=16, degree=2, alpha=3.0, problem=pkg.PROBLEM_EXP)
        go1 = pkg.GridOperator(c1)
        x1 = go1.vector()
        r1 = go1.linear_problem_solve(x1, method=pkg.SOLVER_CG, reduction=1e-10)
        solve["c1_fp64_cg"] = {"workload": "16^3 Q2 SIPG Poisson, exp manufactured solution", "seconds":
                               r1["assembler_time"] + r1["linear_solver_time"],
                               "iterations": r1["solver"]["iterations"], "converged": bool(r1["solver"]["converged"]),
                               "l2_error": float(np.sqrt(go1.l2_error_squared(x1, 10)))}
        x1.free()
        go1.close()
        x.free()
        x = go.vector()
        log("bench: FP64 CG/block Jacobi solve (exp problem)")
        res = go.linear_problem_solve(x, method=pkg.SOLVER_CG, precond=pkg.PREC_BLOCK_JACOBI, reduction=1e-10)
        solve["exp_fp64_cg_block_jacobi"] = {"seconds": res["solver"]["elapsed"],
                                             "iterations": res["solver"]["iterations"],
                                             "converged": bool(res["solver"]["converged"]),
                                             "reduction": res["solver"]["reduction"]}
        x.free()
        # lognormal kappa: plain FP64 vs mixed precision
        z.free()
        y.free()
        go.close()
        cfg3 = pkg.make_config(n=n, degree=k, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT,
                               f0=1.0, g0=0.0)
        kappa = lognormal_kappa(n ** 3, 31337)
        go3 = pkg.GridOperator(cfg3, kappa)
        x64 = go3.vector()
        log("bench: FP64 CG/Jacobi solve (lognormal kappa)")
        r64 = go3.linear_problem_solve(x64, method=pkg.SOLVER_CG, reduction=1e-10, maxit=100000)
        xm = go3.vector()
        log("bench: mixed-precision solve (lognormal kappa)")
        rm = go3.solve_mixed(xm, inner_reduction=args.inner_reduction, outer_reduction=1e-10, inner_maxit=100000)
        xr = go3.vector()
        log("bench: mixed-precision solve, restarted refinement with FP32 CG to 1e-4 (lognormal kappa)")
        rr = go3.solve_mixed(xr, variant=pkg.MIXED_RESTARTED, inner_reduction=1e-4, outer_reduction=1e-10,
                             inner_maxit=100000)
        a64 = x64.download()
        am = xm.download()
        ar = xr.download()
        rel = float(np.linalg.norm(am - a64) / np.linalg.norm(a64))
        t64 = r64["assembler_time"] + r64["linear_solver_time"]
        solve["lognormal_fp64_cg"] = {"seconds": t64, "iterations": r64["solver"]["iterations"],
                                      "converged": bool(r64["solver"]["converged"]),
                                      "reduction": r64["solver"]["reduction"]}
        solve["lognormal_mixed"] = {"seconds": rm["elapsed"], "outer": rm["outer_iterations"],
                                    "inner_iterations": rm["inner_iterations"],
                                    "inner_reduction": args.inner_reduction, "converged": bool(rm["converged"]),
                                    "reduction": rm["reduction"], "rel_diff_vs_fp64_gpu": rel,
                                    "speedup_vs_fp64": t64 / rm["elapsed"] if rm["elapsed"] > 0 else None}
        solve["lognormal_mixed_restarted"] = {
            "seconds": rr["elapsed"], "outer": rr["outer_iterations"], "inner_iterations": rr["inner_iterations"],
            "inner_reduction": 1e-4, "converged": bool(rr["converged"]), "reduction": rr["reduction"],
            "rel_diff_vs_fp64_gpu": float(np.linalg.norm(ar - a64) / np.linalg.norm(a64)),
            "speedup_vs_fp64": t64 / rr["elapsed"] if rr["elapsed"] > 0 else None,
            "note": "the stated configuration: each FP32 CG solve to 1e-4, FP64 refinement to 1e-10"}
        # the oracle's residual of these solutions is evaluated in the cpu_baseline leg
        checks["lognormal"] = {"cfg": cfg3, "kappa": kappa, "first_defect": rm["first_defect"], "target": 1e-10,
                               "workload": f"{n}^3 Q{k} lognormal-kappa problem",
                               "solutions": {"fp64_cg": a64, "mixed_reliable": am, "mixed_restarted": ar}}
        # the same pair with per-cell block Jacobi (Blocking::fixed in the reference)
        bj = pkg.PREC_BLOCK_JACOBI
        x64b = go3.vector()
        log("bench: FP64 CG/block Jacobi solve (lognormal kappa)")
        r64b = go3.linear_problem_solve(x64b, method=pkg.SOLVER_CG, precond=bj, reduction=1e-10, maxit=100000)
        xmb = go3.vector()
        log("bench: mixed-precision solve with block Jacobi (lognormal kappa)")
        rmb = go3.solve_mixed(xmb, precond=bj, inner_reduction=args.inner_reduction, outer_reduction=1e-10,
                              inner_maxit=100000)
        a64b, amb = x64b.download(), xmb.download()
        t64b = r64b["assembler_time"] + r64b["linear_solver_time"]
        solve["lognormal_fp64_cg_block_jacobi"] = {
            "seconds": t64b, "iterations": r64b["solver"]["iterations"],
            "converged": bool(r64b["solver"]["converged"]), "reduction": r64b["solver"]["reduction"],
            "rel_diff_vs_point_jacobi_fp64": float(np.linalg.norm(a64b - a64) / np.linalg.norm(a64))}
        solve["lognormal_mixed_block_jacobi"] = {
            "seconds": rmb["elapsed"], "outer": rmb["outer_iterations"], "inner_iterations": rmb["inner_iterations"],
            "inner_reduction": args.inner_reduction, "converged": bool(rmb["converged"]),
            "reduction": rmb["reduction"],
            "rel_diff_vs_fp64_gpu": float(np.linalg.norm(amb - a64b) / np.linalg.norm(a64b)),
            "speedup_vs_fp64_block_jacobi": t64b / rmb["elapsed"] if rmb["elapsed"] > 0 else None,
            "speedup_vs_fp64_point_jacobi": t64 / rmb["elapsed"] if rmb["elapsed"] > 0 else None}
        out["solve"] = solve
        for v in (x64, xm, xr, x64b, xmb):
            v.free()
        go3.close()
        if not args.no_convection:
            out["convection"] = convection_leg(pkg, checks)
        log("bench: 32^3 Q2 adjoint log-kappa gradient")
        out["adjoint_logkappa"] = adjoint_leg(pkg)
        log("bench: 32^3 Q2 assembled Jacobian and block-matrix apply")
        out["matrix_based"] = matrix_based_leg(pkg)
        log("bench: instationary Alexander2 steps")
        out["instationary"] = instationary_leg(pkg, n, k)
    else:
        z.free()
        y.free()
        go.close()

    if not args.no_cpu:
        log("bench: CPU baseline (oracle)")
        its = out.get("solve", {}).get("exp_fp64_cg", {}).get("iterations")
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds, its, checks)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
