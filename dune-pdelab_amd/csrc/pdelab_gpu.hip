// pdelab_gpu.hip -- C-ABI of the gfx950 matrix-free DG hot path (see include/pdelab_gpu.h).
//
// One handle = one GridOperator + its solver backends on one HIP device and one HIP
// stream. All vectors are device pointers in PDELab's cell-blocked DoF layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "blas_kernels.hh"
#include "handle.hh"

namespace {
thread_local std::string g_create_error;

struct Workspace {
  std::vector<void*> bufs;
};
}  // namespace

// ---------------------------------------------------------------------------
// setup of the kernel argument blocks
template <typename T>
static void fill_args(H* h, DGArgs<T>& a) {
  const Tables1D& t = h->tab;
  const int n1 = t.n1;
  std::memset(&a, 0, sizeof(a));
  for (int i = 0; i < n1; ++i) {
    for (int j = 0; j < n1; ++j) {
      long double mk = 0, mg = 0, mgt = 0;
      for (int l = 0; l < n1; ++l) {
        mk += t.Minv[i][l] * t.K[l][j];
        mg += t.Minv[i][l] * t.G[l][j];
        mgt += t.Minv[i][l] * t.G[j][l];
      }
      a.MiK[i][j] = (T)mk;
      a.MiG[i][j] = (T)mg;
      a.MiGt[i][j] = (T)mgt;
      a.M[i][j] = (T)t.M[i][j];
    }
    long double me0 = 0, me1 = 0, md0 = 0, md1 = 0;
    for (int l = 0; l < n1; ++l) {
      me0 += t.Minv[i][l] * t.e0[l];
      me1 += t.Minv[i][l] * t.e1[l];
      md0 += t.Minv[i][l] * t.d0[l];
      md1 += t.Minv[i][l] * t.d1[l];
    }
    a.Mie0[i] = (T)me0; a.Mie1[i] = (T)me1; a.Mid0[i] = (T)md0; a.Mid1[i] = (T)md1;
    a.d0[i] = (T)t.d0[i]; a.d1[i] = (T)t.d1[i];
    a.Mdiag[i] = (T)t.M[i][i]; a.Kdiag[i] = (T)t.K[i][i]; a.Gdiag[i] = (T)t.G[i][i];
    a.e0d0[i] = (T)(t.e0[i] * t.d0[i]);
    a.e1d1[i] = (T)(t.e1[i] * t.d1[i]);
  }
  const pdelab_config& c = h->cfg;
  const double hh[3] = {c.lx / c.nx, c.ly / c.ny, c.lz / c.nz};
  const double V = hh[0] * hh[1] * hh[2];
  for (int d = 0; d < 3; ++d) {
    a.S[d] = (T)(V / (hh[d] * hh[d]));
    a.Sc[d] = (T)(V / hh[d]);
    a.b[d] = (T)c.b[d];
  }
  a.V = (T)V;
  a.c = (T)c.c;
  a.vc = (T)(V * c.c);
  // convectiondiffusiondg.hh:99-101
  double theta = 1.0;
  if (c.method == PDELAB_SIPG) theta = -1.0;
  if (c.method == PDELAB_IIPG) theta = 0.0;
  a.theta = (T)theta;
  const int k = c.degree;
  a.alphak = (T)(c.alpha * (double)(k * (k + 3 - 1)));
  a.kappa0 = (T)c.kappa0;
  a.weightsOn = c.weights == PDELAB_WEIGHTS_ON ? 1 : 0;
  a.aniso = c.kappa_per_cell == 2 ? 1 : 0;
  // constant-kappa face coefficients, same formulas as face_coeffs()
  {
    const T klo = (T)c.kappa0, khi = (T)c.kappa0;
    if (a.weightsOn) {
      const T den = khi + klo + T(1e-20);
      const T omega_s = klo / den, omega_n = khi / den;
      a.wI_hi = omega_s * khi;
      a.wI_lo = omega_n * klo;
      a.PI = a.alphak * (T(2.0) * khi * klo / den);
      a.PB = a.alphak * (T)c.kappa0;
    } else {
      a.wI_hi = T(0.5) * khi;
      a.wI_lo = T(0.5) * klo;
      a.PI = a.alphak;
      a.PB = a.alphak;
    }
  }
  a.n[0] = c.nx; a.n[1] = c.ny; a.n[2] = c.nz;
  a.stride[0] = 1;
  a.stride[1] = c.nx;
  a.stride[2] = (long long)c.nx * c.ny;
  a.ncells = h->ncells;
  for (int s = 0; s < 6; ++s) a.bc[s] = c.bctype[s] == 0 ? BC_DIRICHLET : c.bctype[s];
}

// ---------------------------------------------------------------------------
// Constant-coefficient line matrices, extracted column by column from line_op in long
// double: A (both faces interior), the corrections Dlo/Dhi of a boundary face, and the
// rank-two neighbour couplings through the trace value and trace derivative.
using LD = long double;

template <int N1, bool TRANS>
static void line_eval(const DGArgs<LD>& al, int D, const LD (&u)[N1], const LD (&ul)[N1], const LD (&ur)[N1],
                      bool lo, bool hi, LD (&s)[N1]) {
  const LD k = al.kappa0;
  const FaceCoef<LD> f{al.wI_lo, al.wI_hi, al.PI};
  if (D == 0) line_op<N1, 0, LD, TRANS, true, true, false>(al, u, ul, ur, lo, hi, k, f, f, s);
  else if (D == 1) line_op<N1, 1, LD, TRANS, true, true, false>(al, u, ul, ur, lo, hi, k, f, f, s);
  else line_op<N1, 2, LD, TRANS, true, true, false>(al, u, ul, ur, lo, hi, k, f, f, s);
}

template <int N1, bool TRANS>
static bool build_line_mats_n(const DGArgs<LD>& al, LineMats<LD>& m) {
  std::memset(&m, 0, sizeof(m));
  LD zero[N1], e[N1], s0[N1], s1[N1], x[N1];
  for (int i = 0; i < N1; ++i) zero[i] = 0;
  bool ok = true;
  for (int D = 0; D < 3; ++D) {
    LD scale = 0;
    for (int j = 0; j < N1; ++j) {
      for (int i = 0; i < N1; ++i) e[i] = (i == j);
      line_eval<N1, TRANS>(al, D, e, zero, zero, true, true, s0);
      for (int i = 0; i < N1; ++i) { m.A[D][i][j] = s0[i]; scale = std::max(scale, std::fabs(s0[i])); }
      line_eval<N1, TRANS>(al, D, e, zero, zero, false, true, s1);
      for (int i = 0; i < N1; ++i) m.Dlo[D][i][j] = s1[i] - s0[i];
      line_eval<N1, TRANS>(al, D, e, zero, zero, true, false, s1);
      for (int i = 0; i < N1; ++i) m.Dhi[D][i][j] = s1[i] - s0[i];
    }
    // lower neighbour: Al x = Pl x[n-1] + Ql (d1 . x)
    int j0 = 0;
    for (int j = 1; j < N1 - 1; ++j)
      if (std::fabs(al.d1[j]) > std::fabs(al.d1[j0])) j0 = j;
    for (int i = 0; i < N1; ++i) x[i] = (i == j0) ? 1 / al.d1[j0] : 0;
    line_eval<N1, TRANS>(al, D, zero, x, zero, true, true, s0);
    for (int i = 0; i < N1; ++i) e[i] = (i == N1 - 1);
    line_eval<N1, TRANS>(al, D, zero, e, zero, true, true, s1);
    for (int i = 0; i < N1; ++i) { m.Ql[D][i] = s0[i]; m.Pl[D][i] = s1[i] - s0[i] * al.d1[N1 - 1]; }
    // upper neighbour: Ar x = Pr x[0] + Qr (d0 . x)
    j0 = 1;
    for (int j = 2; j < N1; ++j)
      if (std::fabs(al.d0[j]) > std::fabs(al.d0[j0])) j0 = j;
    for (int i = 0; i < N1; ++i) x[i] = (i == j0) ? 1 / al.d0[j0] : 0;
    line_eval<N1, TRANS>(al, D, zero, zero, x, true, true, s0);
    for (int i = 0; i < N1; ++i) e[i] = (i == 0);
    line_eval<N1, TRANS>(al, D, zero, zero, e, true, true, s1);
    for (int i = 0; i < N1; ++i) { m.Qr[D][i] = s0[i]; m.Pr[D][i] = s1[i] - s0[i] * al.d0[0]; }
    // the neighbour couplings must be exactly of that rank-two form
    for (int j = 0; j < N1; ++j) {
      for (int i = 0; i < N1; ++i) e[i] = (i == j);
      line_eval<N1, TRANS>(al, D, zero, e, zero, true, true, s0);
      line_eval<N1, TRANS>(al, D, zero, zero, e, true, true, s1);
      for (int i = 0; i < N1; ++i) {
        const LD pl = m.Pl[D][i] * (j == N1 - 1) + m.Ql[D][i] * al.d1[j];
        const LD pr = m.Pr[D][i] * (j == 0) + m.Qr[D][i] * al.d0[j];
        const LD tol = 1e-14L * (scale + std::fabs(s0[i]) + std::fabs(s1[i]));
        if (std::fabs(pl - s0[i]) > tol || std::fabs(pr - s1[i]) > tol) ok = false;
      }
    }
  }
  return ok;
}

template <typename T>
static void cast_line_mats(const LineMats<LD>& m, LineMats<T>& o) {
  const LD* src = &m.A[0][0][0];
  T* dst = &o.A[0][0][0];
  static_assert(sizeof(LineMats<LD>) / sizeof(LD) == sizeof(LineMats<T>) / sizeof(T), "layout");
  for (size_t i = 0; i < sizeof(LineMats<LD>) / sizeof(LD); ++i) dst[i] = (T)src[i];
}

// The structure used by line_op_sym: A_d = S_d Ah (+ V c I for d = 0), Pl_d = S_d Ph,
// Ql_d = S_d Qh, Pr_d[i] = S_d Ph[n-1-i], Qr_d[i] = -S_d Qh[n-1-i], Ah centro-symmetric
// and d0[j] = -d1[n-1-j]; it holds for constant kappa without convection (the
// operator commutes with the reflection x -> 1-x of every direction).
static bool symmetric_form(const DGArgs<LD>& al, const LineMats<LD>& m, int n, LD (&Ah)[4][4], LD (&Ph)[4],
                           LD (&Qh)[4]) {
  std::memset(Ah, 0, sizeof(Ah));
  std::memset(Ph, 0, sizeof(Ph));
  std::memset(Qh, 0, sizeof(Qh));
  const LD S1 = al.S[1];
  LD scale = 0, pscale = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      Ah[i][j] = m.A[1][i][j] / S1;
      scale = std::max(scale, std::fabs(Ah[i][j]));
    }
    Ph[i] = m.Pl[1][i] / S1;
    Qh[i] = m.Ql[1][i] / S1;
    pscale = std::max(pscale, std::max(std::fabs(Ph[i]), std::fabs(Qh[i])));
  }
  const LD tol = 1e-12L;
  auto close = [&](LD x, LD y, LD sc) { return std::fabs(x - y) <= tol * sc; };
  for (int j = 0; j < n; ++j)
    if (!close(al.d0[j], -al.d1[n - 1 - j], std::fabs(al.d1[n - 1 - j]) + 1)) return false;
  for (int d = 0; d < 3; ++d) {
    const LD Sd = al.S[d];
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        const LD a = m.A[d][i][j] - (d == 0 && i == j ? al.vc : 0);
        if (!close(a, Sd * Ah[i][j], Sd * scale)) return false;
        if (!close(Ah[i][j], Ah[n - 1 - i][n - 1 - j], scale)) return false;
      }
      if (!close(m.Pl[d][i], Sd * Ph[i], Sd * pscale) || !close(m.Ql[d][i], Sd * Qh[i], Sd * pscale)) return false;
      if (!close(m.Pr[d][i], Sd * Ph[n - 1 - i], Sd * pscale) || !close(m.Qr[d][i], -Sd * Qh[n - 1 - i], Sd * pscale))
        return false;
    }
  }
  return true;
}

static int build_line_mats(H* h) {
  DGArgs<LD> al;
  fill_args<LD>(h, al);
  LineMats<LD> m[2];
  bool ok = true;
  switch (h->n1) {
    case 2: ok = build_line_mats_n<2, false>(al, m[0]) && build_line_mats_n<2, true>(al, m[1]); break;
    case 3: ok = build_line_mats_n<3, false>(al, m[0]) && build_line_mats_n<3, true>(al, m[1]); break;
    default: ok = build_line_mats_n<4, false>(al, m[0]) && build_line_mats_n<4, true>(al, m[1]); break;
  }
  if (!ok) return fail(h, PDELAB_ERR_INVALID_ARG, "constant-coefficient line operator is not of rank-two face form");
  for (int t = 0; t < 2; ++t) {
    LD Ah[4][4], Ph[4], Qh[4];
    h->sym[t] = symmetric_form(al, m[t], h->n1, Ah, Ph, Qh);
    for (int i = 0; i < 4; ++i) {
      h->a64.Ph[t][i] = (double)Ph[i];
      h->a64.Qh[t][i] = (double)Qh[i];
      h->a32.Ph[t][i] = (float)Ph[i];
      h->a32.Qh[t][i] = (float)Qh[i];
      for (int j = 0; j < 4; ++j) {
        h->a64.Ah[t][i][j] = (double)Ah[i][j];
        h->a32.Ah[t][i][j] = (float)Ah[i][j];
      }
    }
  }
  LineMats<double> m64[2];
  LineMats<float> m32[2];
  for (int t = 0; t < 2; ++t) { cast_line_mats(m[t], m64[t]); cast_line_mats(m[t], m32[t]); }
  HIPCHK(h, hipMalloc(&h->d_lm64, sizeof(m64)));
  HIPCHK(h, hipMalloc(&h->d_lm32, sizeof(m32)));
  HIPCHK(h, hipMemcpy(h->d_lm64, m64, sizeof(m64), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_lm32, m32, sizeof(m32), hipMemcpyHostToDevice));
  return PDELAB_OK;
}

static void fill_load_args(H* h) {
  LoadArgs& a = h->la;
  std::memset(&a, 0, sizeof(a));
  const pdelab_config& c = h->cfg;
  const Tables1D& t = h->tab;
  a.problem = c.problem;
  a.kappa0 = c.kappa0;
  for (int d = 0; d < 3; ++d) a.b[d] = c.b[d];
  a.c = c.c;
  a.f0 = c.f0;
  a.g0 = c.g0;
  a.j0 = c.j0;
  a.o0 = c.o0;
  for (int i = 0; i < 7; ++i) a.quad[i] = c.quad[i];
  a.n_gauss = c.n_gaussians;
  for (int i = 0; i < c.n_gaussians && i < 16; ++i)
    for (int j = 0; j < 5; ++j) a.gauss[i][j] = c.gaussians[i][j];
  a.n[0] = c.nx; a.n[1] = c.ny; a.n[2] = c.nz;
  a.ncells = h->ncells;
  a.h[0] = c.lx / c.nx; a.h[1] = c.ly / c.ny; a.h[2] = c.lz / c.nz;
  a.V = a.h[0] * a.h[1] * a.h[2];
  a.m = t.m;
  for (int q = 0; q < t.m; ++q) {
    a.qx[q] = (double)t.qx[q];
    a.qw[q] = (double)t.qw[q];
    for (int i = 0; i < t.n1; ++i) a.P[q][i] = (double)t.P[q][i];
  }
  for (int i = 0; i < t.n1; ++i) {
    a.e0[i] = (double)t.e0[i]; a.e1[i] = (double)t.e1[i];
    a.d0[i] = (double)t.d0[i]; a.d1[i] = (double)t.d1[i];
  }
  a.theta = (double)h->a64.theta;
  a.alphak = (double)h->a64.alphak;
  a.weightsOn = h->a64.weightsOn;
  a.aniso = h->a64.aniso;
  for (int s = 0; s < 6; ++s) a.bc[s] = h->a64.bc[s];
  a.bface_offset[0] = 0;
  for (int s = 0; s < 6; ++s) {
    int d = s / 2, t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    a.bface_offset[s + 1] = a.bface_offset[s] + (long long)a.n[t1] * a.n[t2];
  }
  GradArgs& g = h->ga;
  std::memset(&g, 0, sizeof(g));
  for (int d = 0; d < 3; ++d) { g.n[d] = a.n[d]; g.stride[d] = h->a64.stride[d]; g.S[d] = (double)h->a64.S[d]; }
  g.ncells = h->ncells;
  for (int i = 0; i < t.n1; ++i) {
    for (int j = 0; j < t.n1; ++j) { g.M[i][j] = (double)t.M[i][j]; g.K[i][j] = (double)t.K[i][j]; }
    g.e0[i] = a.e0[i]; g.e1[i] = a.e1[i]; g.d0[i] = a.d0[i]; g.d1[i] = a.d1[i];
  }
  g.theta = a.theta;
  g.alphak = a.alphak;
  for (int s = 0; s < 6; ++s) g.bc[s] = a.bc[s];
}

// ---------------------------------------------------------------------------
// kernel dispatch
static int check_outflow(H* h) {
  if (h->varb) return PDELAB_OK;  // a velocity field is checked when it is set
  for (int s = 0; s < 6; ++s) {
    if (h->a64.bc[s] != BC_OUTFLOW) continue;
    const int d = s / 2;
    const double nf = (s % 2 ? 1.0 : -1.0) * h->cfg.b[d];
    if (nf < -1e-30) return fail(h, PDELAB_ERR_OUTFLOW_ON_INFLOW, "Outflow boundary condition on inflow!");
  }
  return PDELAB_OK;
}

template <typename T>
static int apply_mode(H* h, const T* z, T* y, bool trans, int mode, T* dotp = nullptr,
                      const int* stop = nullptr) {
  CHK(check_outflow(h));
  if (h->varb && mode == 2)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "velocity field: CG (non-symmetric operator), use BiCGSTAB");
  int st = PDELAB_ERR_UNSUPPORTED;
  switch (h->n1) {
    case 2: st = launch_apply_n<2, T>(h, z, y, fc_for<T>(h), args_for<T>(h), trans, mode, dotp, stop); break;
    case 3: st = launch_apply_n<3, T>(h, z, y, fc_for<T>(h), args_for<T>(h), trans, mode, dotp, stop); break;
    case 4: st = launch_apply_n<4, T>(h, z, y, fc_for<T>(h), args_for<T>(h), trans, mode, dotp, stop); break;
    default: return fail(h, PDELAB_ERR_UNSUPPORTED, "degree");
  }
  CHK(st);
  return vel_launch<T>(h, z, y, stop, trans);  // the velocity field's convection terms, if set
}
static int apply64(H* h, const double* z, double* y, bool trans, bool accum) {
  return apply_mode<double>(h, z, y, trans, accum ? 1 : 0);
}
static int apply32(H* h, const float* z, float* y, bool trans, bool accum) {
  return apply_mode<float>(h, z, y, trans, accum ? 1 : 0);
}
template <typename T>
static int apply_t(H* h, const T* z, T* y, bool trans, bool accum) {
  return apply_mode<T>(h, z, y, trans, accum ? 1 : 0);
}
// partials of the fused apply dot: one per apply workgroup (a walking workgroup covers
// apply_walk(h) chunks)
static long long apply_partials(H* h) { return apply_blocks(h) / apply_walk(h); }

// PDELAB_GPU_APPLY=general runs the general apply kernel on every grid (for comparison;
// by default row_kernel() chooses per launch).
static void choose_apply_shape(H* h) {
  const char* v = std::getenv("PDELAB_GPU_APPLY");
  h->apply_variant = (v && std::strcmp(v, "general") == 0) ? 1 : 0;
}

static int compute_R0(H* h) {
  if (h->R0_valid) return PDELAB_OK;
  if (!h->d_R0) HIPCHK(h, hipMalloc(&h->d_R0, sizeof(double) * h->N));
  HIPCHK(h, hipMemsetAsync(h->d_R0, 0, sizeof(double) * h->N, h->stream));
  const long long n = h->N;
  const int nb = (int)((n + 255) / 256);
  const double* bvt = h->varb ? h->d_bv64 : nullptr;  // velocity field at the volume / face points
  const double* bft = h->varb ? h->d_bf64 : nullptr;
  switch (h->n1) {
    case 2: hipLaunchKernelGGL(load_volume_kernel<2>, dim3(nb), dim3(256), 0, h->stream, h->d_R0, h->d_ftab, h->d_src, h->la, bvt); break;
    case 3: hipLaunchKernelGGL(load_volume_kernel<3>, dim3(nb), dim3(256), 0, h->stream, h->d_R0, h->d_ftab, h->d_src, h->la, bvt); break;
    case 4: hipLaunchKernelGGL(load_volume_kernel<4>, dim3(nb), dim3(256), 0, h->stream, h->d_R0, h->d_ftab, h->d_src, h->la, bvt); break;
  }
  HIPCHK(h, hipGetLastError());
  // boundary data: g on Dirichlet sides (none for the ZERO problem), j on Neumann and o on
  // outflow sides (j0 / o0 for the problems without an exact solution)
  const int prob = h->cfg.problem;
  const bool exact_fluxes = prob == PDELAB_PROBLEM_EXP || prob == PDELAB_PROBLEM_QUADRATIC || prob == PDELAB_PROBLEM_TABULATED;
  for (int s = 0; s < 6; ++s) {
    const int bct = h->a64.bc[s];
    bool data = false;
    if (bct == BC_DIRICHLET) data = prob != PDELAB_PROBLEM_ZERO;
    else if (bct == BC_NEUMANN) data = exact_fluxes || h->cfg.j0 != 0.0;
    else if (bct == BC_OUTFLOW) data = exact_fluxes || h->cfg.o0 != 0.0;
    if (!data) continue;
    const int d = s / 2, t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const long long nf = (long long)h->la.n[t1] * h->la.n[t2] * h->nloc;
    const int gb = (int)((nf + 255) / 256);
    switch (h->n1) {
      case 2: hipLaunchKernelGGL(load_boundary_kernel<2>, dim3(gb), dim3(256), 0, h->stream, h->d_R0, h->d_gtab, h->d_kappa, h->la, s, bct, bft); break;
      case 3: hipLaunchKernelGGL(load_boundary_kernel<3>, dim3(gb), dim3(256), 0, h->stream, h->d_R0, h->d_gtab, h->d_kappa, h->la, s, bct, bft); break;
      case 4: hipLaunchKernelGGL(load_boundary_kernel<4>, dim3(gb), dim3(256), 0, h->stream, h->d_R0, h->d_gtab, h->d_kappa, h->la, s, bct, bft); break;
    }
    HIPCHK(h, hipGetLastError());
  }
  h->R0_valid = true;
  return PDELAB_OK;
}

static int compute_diag(H* h) {
  if (h->diag_valid) return PDELAB_OK;
  const long long n = h->N;
  if (!h->d_diag) HIPCHK(h, hipMalloc(&h->d_diag, sizeof(double) * n));
  if (!h->d_dinv) HIPCHK(h, hipMalloc(&h->d_dinv, sizeof(double) * n));
  if (!h->d_dinv32) HIPCHK(h, hipMalloc(&h->d_dinv32, sizeof(float) * n));
  HIPCHK(h, hipMemsetAsync(h->d_diag, 0, sizeof(double) * n, h->stream));
  const int nb = (int)((n + 255) / 256);
#define DIAGL(N1_)                                                                                           \
  if (h->vark) hipLaunchKernelGGL((dg_diag_kernel<N1_, double, true>), dim3(nb), dim3(256), 0, h->stream, h->d_diag, h->d_kappa, h->a64); \
  else hipLaunchKernelGGL((dg_diag_kernel<N1_, double, false>), dim3(nb), dim3(256), 0, h->stream, h->d_diag, h->d_kappa, h->a64);
  switch (h->n1) {
    case 2: DIAGL(2); break;
    case 3: DIAGL(3); break;
    case 4: DIAGL(4); break;
  }
#undef DIAGL
  HIPCHK(h, hipGetLastError());
  CHK(vel_diag(h, h->d_diag));
  hipLaunchKernelGGL(invert_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, h->d_diag, h->d_dinv);
  hipLaunchKernelGGL(invert_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, h->d_diag, h->d_dinv32);
  HIPCHK(h, hipGetLastError());
  h->diag_valid = true;
  return PDELAB_OK;
}

// ---------------------------------------------------------------------------
// reductions
template <typename A>
static A* partial_buf(H* h);
template <>
double* partial_buf<double>(H* h) { return h->d_part64; }
template <>
float* partial_buf<float>(H* h) { return h->d_part32; }
template <typename A>
static A* scal_buf(H* h);
template <>
double* scal_buf<double>(H* h) { return h->d_scal64; }
template <>
float* scal_buf<float>(H* h) { return h->d_scal32; }
template <typename A>
static A* hscal_buf(H* h);
template <>
double* hscal_buf<double>(H* h) { return h->h_scal64; }
template <>
float* hscal_buf<float>(H* h) { return h->h_scal32; }

template <typename A, int NR>
static int finish_reduce(H* h, int slot) {
  hipLaunchKernelGGL((final_reduce_kernel<A, NR>), dim3(1), dim3(kFinalThreads), 0, h->stream, partial_buf<A>(h),
                     kRedBlocks, scal_buf<A>(h) + slot, (const int*)nullptr);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <typename A>
static int read_scalars(H* h, A* out, int count) {
  HIPCHK(h, hipMemcpyAsync(hscal_buf<A>(h), scal_buf<A>(h), sizeof(A) * count, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::memcpy(out, hscal_buf<A>(h), sizeof(A) * count);
  return PDELAB_OK;
}

template <typename T, typename A>
static int dot_dev(H* h, const T* x, const T* y, long long n, int slot) {
  hipLaunchKernelGGL((dot_partial_kernel<T, A>), dim3(kRedBlocks), dim3(kRedThreads), 0, h->stream, x, y, n,
                     partial_buf<A>(h));
  HIPCHK(h, hipGetLastError());
  return finish_reduce<A, 1>(h, slot);
}

// ---------------------------------------------------------------------------
// dune-istl IterativeSolver::Iteration semantics
struct Iter {
  double def0 = 0, def = 0, red, i = 0;
  bool converged = false, aborted = false;
  explicit Iter(double r) : red(r) {}
  bool step(double it, double d) {
    if (!std::isfinite(d)) { aborted = true; def = d; i = it; converged = false; return true; }
    if (it == 0) def0 = d;
    def = d;
    i = it;
    converged = (d < def0 * red) || (d < 1e-30);
    return converged;
  }
  void finalize(pdelab_linear_solver_result* r, double el) const {
    r->converged = converged ? 1 : 0;
    r->iterations = (unsigned)std::ceil(i);
    r->reduction = def0 > 0 ? def / def0 : 0.0;
    r->conv_rate = std::pow(r->reduction, 1.0 / std::max(1.0, std::ceil(i)));
    r->elapsed = el;
  }
};

template <typename T>
static T** ws(H* h);
template <>
double** ws<double>(H* h) { return h->w64; }
template <>
float** ws<float>(H* h) { return h->w32; }

template <typename T>
static int ensure_ws(H* h, int count) {
  T** w = ws<T>(h);
  for (int i = 0; i < count; ++i)
    if (!w[i]) HIPCHK(h, hipMalloc(&w[i], sizeof(T) * h->N));
  return PDELAB_OK;
}

template <typename T>
static const T* dinv_for(H* h, int precond);
template <>
const double* dinv_for<double>(H* h, int precond) { return precond == PDELAB_PREC_JACOBI ? h->d_dinv : nullptr; }
template <>
const float* dinv_for<float>(H* h, int precond) { return precond == PDELAB_PREC_JACOBI ? h->d_dinv32 : nullptr; }

// Preconditioner W of a Krylov solve: none, point Jacobi (dinv), per-cell block Jacobi
// of J (or of J^T for the transposed operator), one block-SOR sweep of J, or block Jacobi of
// J with iteratively solved blocks.
template <typename T>
struct Prec {
  int kind = PDELAB_PREC_NONE;
  const T* dinv = nullptr;
  bool trans = false;
  bool block() const { return kind == PDELAB_PREC_BLOCK_JACOBI; }
  bool sor() const { return kind == PDELAB_PREC_BLOCK_SOR; }
  bool ibj() const { return kind == PDELAB_PREC_ITERATIVE_BLOCK_JACOBI; }
};

// The cached point diagonal and block Jacobi data belong to the operator with the current
// reaction shift (one-step stage operators); dropped when the shift has changed since.
static void sync_prec_cache(H* h) {
  if (h->prec_shift == h->shift) return;
  h->diag_valid = false;
  bj_invalidate(h);
  h->prec_shift = h->shift;
}

// Checks the preconditioner and prepares its data (point diagonal or block data).
template <typename T>
static int make_prec(H* h, int precond, bool trans, Prec<T>& P) {
  sync_prec_cache(h);
  P.kind = precond;
  P.trans = trans;
  if (precond == PDELAB_PREC_JACOBI) CHK(compute_diag(h));
  else if (precond == PDELAB_PREC_BLOCK_JACOBI) CHK(bj_prepare(h, trans, sizeof(T) == 8 ? 64 : 32));
  else if (precond == PDELAB_PREC_BLOCK_SOR) {
    if (trans) return fail(h, PDELAB_ERR_UNSUPPORTED, "block SOR: not for the transposed operator");
    CHK(sor_prepare(h, sizeof(T) == 8 ? 64 : 32));
  } else if (precond == PDELAB_PREC_ITERATIVE_BLOCK_JACOBI) {
    if (trans) return fail(h, PDELAB_ERR_UNSUPPORTED, "iterative block Jacobi: not for the transposed operator");
    CHK(ibj_prepare(h, sizeof(T) == 8 ? 64 : 32));
  } else if (precond != PDELAB_PREC_NONE) return fail(h, PDELAB_ERR_INVALID_ARG, "unknown preconditioner");
  P.dinv = dinv_for<T>(h, precond);
  return PDELAB_OK;
}

// out = W^{-1} in; out.in into scal[slot] unless slot < 0
template <typename T, typename A>
static int prec_dot(H* h, const Prec<T>& P, const T* in, T* out, int slot) {
  if (P.block()) {
    const BJVec<T, A> v{in, out, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nb = 0;
    CHK((bj_launch<T, A>(h, P.trans, BJ_PREC, v, &nb)));
    if (slot >= 0)
      hipLaunchKernelGGL((final_reduce_kernel<A, 1>), dim3(1), dim3(kFinalThreads), 0, h->stream,
                         (const A*)h->d_bjpart, nb, scal_buf<A>(h) + slot, (const int*)nullptr);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  }
  hipLaunchKernelGGL((prec_dot_kernel<T, A>), dim3(kRedBlocks), dim3(kRedThreads), 0, h->stream, h->N, P.dinv, in, out,
                     partial_buf<A>(h));
  HIPCHK(h, hipGetLastError());
  if (slot >= 0) CHK((finish_reduce<A, 1>(h, slot)));
  return PDELAB_OK;
}

// Iterations between host reads of the device CG loop's stop flag.
constexpr int kCheckEvery = 16;

// Outcome of one run of the device CG loop.
struct CgLoopOut {
  int iters = 0;          // iterations done (the stopping one included)
  double def = 0;         // defect at the last iteration
  bool stopped = false;   // the stopping rule fired (else the iteration limit was reached)
  bool aborted = false;   // non-finite defect
  int rho_slot = 7;       // scalar slot of the rho the last iteration used (beta's denominator)
};

// The device CG loop from iteration 1: p holds the search direction and scal[7] the rho
// of iteration 1. Stops where dune-istl's rule fires against def_ref:
// def < def_ref * red || def < 1e-30.
template <typename T, typename A>
static int cg_loop(H* h, T* x, T* b, T* p, T* q, T* z, const Prec<T>& P, bool trans, int maxit, double def_ref,
                   double red, CgLoopOut& out) {
  const long long n = h->N;
  const long long nbapply = apply_partials(h);
  // sized for either kernel: the choice can change with set_kappa
  if (!h->d_applydot) {
    const long long nmax = std::max(std::max(apply_blocks_of(h, true), apply_blocks_of(h, false)), 1LL);
    HIPCHK(h, hipMalloc(&h->d_applydot, sizeof(double) * (size_t)nmax));
  }
  static_assert(std::is_same<T, A>::value, "the fused apply dot is accumulated in the vector type");
  A* scal = scal_buf<A>(h);
  T* dotp = reinterpret_cast<T*>(h->d_applydot);
  h->h_state[0] = h->h_state[1] = h->h_state[2] = h->h_state[3] = 0;
  h->h_dstate[0] = def_ref;
  h->h_dstate[1] = def_ref;
  h->h_dstate[2] = red;
  HIPCHK(h, hipMemcpyAsync(h->d_state, h->h_state, sizeof(int) * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_dstate, h->h_dstate, sizeof(double) * 4, hipMemcpyHostToDevice, h->stream));
  bool stopped = false;
  int last = 0;
  for (int i = 1; i <= maxit && !stopped; ++i) {
    const int rslot = (i & 1) ? 7 : 5;  // slot holding the current rho
    const int oslot = (i & 1) ? 4 : 6;  // (def^2, rho_new) written here
    // q = J p with the p.q partials -> scal[0]; r -= lambda q with the (r.r, z.r)
    // partials -> scal[oslot..] and the defect check; x += lambda p, p = z + beta p
    CHK(apply_mode<T>(h, p, q, trans, 2, dotp, h->d_state));
    hipLaunchKernelGGL((cg_pq_kernel<A>), dim3(1), dim3(kFinalThreads), 0, h->stream, (const A*)dotp, (int)nbapply,
                       scal, rslot, (const int*)h->d_state);
    const A* parts = partial_buf<A>(h);
    int nparts = kRedBlocks;
    if (P.block()) {
      // r -= lambda q and z = W^{-1} r in one pass over the cells (z is kept for the p update)
      const BJVec<T, A> v{b, z, b, q, scal + kSlotLambda, nullptr, h->d_state};
      CHK((bj_launch<T, A>(h, P.trans, BJ_CGR, v, &nparts)));
      parts = static_cast<const A*>(h->d_bjpart);
    } else {
      hipLaunchKernelGGL((cg_rupdate_kernel<T, A>), dim3(kRedBlocks), dim3(kRedThreads), 0, h->stream, n, b,
                         (const T*)q, P.dinv, (const A*)scal, partial_buf<A>(h), (const int*)h->d_state);
    }
    hipLaunchKernelGGL((cg_check_kernel<A>), dim3(1), dim3(kFinalThreads), 0, h->stream, parts, nparts, scal + oslot,
                       h->d_state, h->d_dstate, i, (const A*)(scal + rslot), scal + kSlotBeta);
    {
      const dim3 grid((unsigned)((n + 256LL * xp_ept<T>() - 1) / (256LL * xp_ept<T>())));
      const T* zsrc = P.block() ? z : nullptr;
      if (P.block())
        hipLaunchKernelGGL((cg_xpupdate_kernel<T, A, XP_Z>), grid, dim3(256), 0, h->stream, n, x, p, (const T*)b, P.dinv,
                           zsrc, (const A*)scal, (const int*)h->d_state, i);
      else if (P.dinv)
        hipLaunchKernelGGL((cg_xpupdate_kernel<T, A, XP_DINV_R>), grid, dim3(256), 0, h->stream, n, x, p, (const T*)b,
                           P.dinv, zsrc, (const A*)scal, (const int*)h->d_state, i);
      else
        hipLaunchKernelGGL((cg_xpupdate_kernel<T, A, XP_R>), grid, dim3(256), 0, h->stream, n, x, p, (const T*)b, P.dinv,
                           zsrc, (const A*)scal, (const int*)h->d_state, i);
    }
    HIPCHK(h, hipGetLastError());
    last = i;
    if (i % kCheckEvery == 0 || i == maxit) {
      HIPCHK(h, hipMemcpyAsync(h->h_state, h->d_state, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      stopped = h->h_state[0] != 0;
    }
  }
  HIPCHK(h, hipMemcpyAsync(h->h_state, h->d_state, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_dstate, h->d_dstate, sizeof(double) * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  out.stopped = h->h_state[0] != 0;
  out.aborted = h->h_state[1] != 0;
  out.iters = out.stopped ? h->h_state[2] : last;
  out.def = h->h_dstate[0];
  out.rho_slot = (out.iters & 1) ? 7 : 5;
  return PDELAB_OK;
}

// CGSolver::apply (dune-istl): solves J z = b from the given z; b becomes the defect.
// Workspace slots base, base+1 (p, q).
//
// The loop runs on the device: per iteration the apply (with the fused p.q partials),
// the update of r (with the r.r and z.r partials), the Iteration::step check and the
// update of x and the direction are enqueued without waiting; the check kernel raises a
// device flag at the iteration where dune-istl's stopping rule fires and every later
// kernel of the loop returns immediately, so the state is exactly the one at that
// iteration. The host reads the flag back every kCheckEvery iterations.
template <typename T, typename A>
static int cg_solve(H* h, T* x, T* b, double reduction, int maxit, int precond, bool trans, int base,
                    pdelab_linear_solver_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  const long long n = h->N;
  Prec<T> P;
  CHK(make_prec<T>(h, precond, trans, P));
  CHK(ensure_ws<T>(h, base + (P.block() ? 3 : 2)));
  T* p = ws<T>(h)[base];
  T* q = ws<T>(h)[base + 1];
  T* z = P.block() ? ws<T>(h)[base + 2] : nullptr;
  Iter it(reduction);
  // b -= J x  (applyscaleadd(-1, x, b))
  CHK(apply_t<T>(h, x, q, trans, false));
  hipLaunchKernelGGL(axpy_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, T(-1), (const T*)q, b);
  CHK(dot_dev<T, A>(h, b, b, n, 1));
  A sc[8];
  CHK(read_scalars<A>(h, sc, 2));
  const double def0 = std::sqrt((double)sc[1]);
  if (!it.step(0, def0)) {
    // p = W^{-1} b; rho = p.b  -> slot 7
    CHK((prec_dot<T, A>(h, P, b, p, 7)));
    CgLoopOut lo;
    CHK((cg_loop<T, A>(h, x, b, p, q, z, P, trans, maxit, def0, reduction, lo)));
    it.def = lo.def;
    it.i = lo.iters;
    it.aborted = lo.aborted;
    it.converged = !lo.aborted && ((lo.def < it.def0 * it.red) || (lo.def < 1e-30));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  it.finalize(res, el);
  if (it.aborted) return fail(h, PDELAB_ERR_SOLVER_ABORT, "CG: defect is infinite or NaN");
  return PDELAB_OK;
}

// Iterations of the device BiCGSTAB loop between host reads of its stop flag.
constexpr int kBicgCheckEvery = 8;

// BiCGSTABSolver::apply (dune-istl) with half-step iteration counting, as a device-side
// loop: the scalars, the Iteration::step test and the breakdown checks run in
// single-workgroup kernels (bicg_scalar_kernel), so an iteration enqueues its 13 launches
// without waiting; the host reads the stop flag every kBicgCheckEvery iterations.
// Workspace slots base .. base+4 (p, v, t, y, rt).
template <typename T, typename A>
static int bicgstab_solve(H* h, T* x, T* r, double reduction, int maxit, int precond, bool trans, int base,
                          pdelab_linear_solver_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  const long long n = h->N;
  CHK(ensure_ws<T>(h, base + 5));
  T *p = ws<T>(h)[base], *v = ws<T>(h)[base + 1], *t = ws<T>(h)[base + 2], *y = ws<T>(h)[base + 3],
    *rt = ws<T>(h)[base + 4];
  Prec<T> P;
  CHK(make_prec<T>(h, precond, trans, P));
  if (!h->d_bscal) HIPCHK(h, hipMalloc(&h->d_bscal, sizeof(double) * 8));
  Iter iter(reduction);
  int status = PDELAB_OK;
  CHK(apply_t<T>(h, x, t, trans, false));
  hipLaunchKernelGGL(axpy_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, T(-1), (const T*)t, r);
  HIPCHK(h, hipMemcpyAsync(rt, r, sizeof(T) * n, hipMemcpyDeviceToDevice, h->stream));
  CHK(dot_dev<T, A>(h, r, r, n, 0));
  A sc[4];
  CHK(read_scalars<A>(h, sc, 1));
  const double norm0 = std::sqrt((double)sc[0]);
  if (!iter.step(0, norm0)) {
    int* st = h->d_state;
    double* ds = h->d_dstate;
    double* bs = h->d_bscal;
    // rho = alpha = omega = 1, rho_new = rt.r = r.r (rt = r), beta = 0
    hipLaunchKernelGGL(bicg_init_kernel, dim3(1), dim3(64), 0, h->stream, bs, (double)sc[0]);
    h->h_state[0] = h->h_state[1] = h->h_state[2] = h->h_state[3] = 0;
    h->h_dstate[0] = norm0;
    h->h_dstate[1] = norm0;
    h->h_dstate[2] = reduction;
    h->h_dstate[3] = 0.0;
    HIPCHK(h, hipMemcpyAsync(st, h->h_state, sizeof(int) * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(ds, h->h_dstate, sizeof(double) * 4, hipMemcpyHostToDevice, h->stream));
    A* part = partial_buf<A>(h);
    const dim3 rg(kRedBlocks), rb(kRedThreads), fg(1), fb(kFinalThreads);
    if (P.ibj()) HIPCHK(h, hipMemsetAsync(h->d_ibj_stat, 0, 16, h->stream));
    auto prec = [&](const T* in, T* out) -> int {
      if (P.sor()) return sor_launch<T>(h, in, out, st, false);
      if (P.ibj()) return ibj_launch<T>(h, in, out, static_cast<int*>(h->d_ibj_stat), nullptr, st, false);
      if (P.block()) {
        const BJVec<T, A> bv{in, out, nullptr, nullptr, nullptr, nullptr, st};
        return bj_launch<T, A>(h, P.trans, BJ_PREC, bv, nullptr);
      }
      hipLaunchKernelGGL(prec_stop_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, P.dinv, in, out,
                         (const int*)st);
      HIPCHK(h, hipGetLastError());
      return PDELAB_OK;
    };
    bool stopped = false;
    for (int i = 1; i <= maxit && !stopped; ++i) {
      // first half step: p, y = W^{-1} p, v = J y, alpha, x += alpha y, r -= alpha v
      hipLaunchKernelGGL(bicg_p_dev_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, p, (const T*)r,
                         (const T*)v, (const double*)bs, (const int*)st, i == 1 ? 1 : 0);
      CHK(prec(p, y));
      CHK(apply_mode<T>(h, y, v, trans, 0, nullptr, st));
      hipLaunchKernelGGL((dot_partial_kernel<T, A>), rg, rb, 0, h->stream, (const T*)rt, (const T*)v, n, part);
      hipLaunchKernelGGL((bicg_scalar_kernel<A, 0>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                         2 * i - 1);
      hipLaunchKernelGGL((bicg_xr_kernel<T, A, 1>), rg, rb, 0, h->stream, n, x, (const T*)y, r, (const T*)v,
                         (const T*)nullptr, (const double*)bs, kBsAlpha, part, (const int*)st);
      hipLaunchKernelGGL((bicg_scalar_kernel<A, 1>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                         2 * i - 1);
      // second half step: y = W^{-1} r, t = J y, omega, x += omega y, r -= omega t, rho
      CHK(prec(r, y));
      CHK(apply_mode<T>(h, y, t, trans, 0, nullptr, st));
      hipLaunchKernelGGL((dot2_kernel<T, A>), rg, rb, 0, h->stream, n, (const T*)t, (const T*)r, part);
      hipLaunchKernelGGL((bicg_scalar_kernel<A, 2>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                         2 * i);
      hipLaunchKernelGGL((bicg_xr_kernel<T, A, 2>), rg, rb, 0, h->stream, n, x, (const T*)y, r, (const T*)t,
                         (const T*)rt, (const double*)bs, kBsOmega, part, (const int*)st);
      hipLaunchKernelGGL((bicg_scalar_kernel<A, 3>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                         2 * i);
      HIPCHK(h, hipGetLastError());
      if (i % kBicgCheckEvery == 0 || i == maxit) {
        HIPCHK(h, hipMemcpyAsync(h->h_state, st, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        stopped = h->h_state[0] != 0;
      }
    }
    HIPCHK(h, hipMemcpyAsync(h->h_state, st, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->h_dstate, ds, sizeof(double) * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int code = h->h_state[1];
    iter.def = h->h_dstate[0];
    iter.i = 0.5 * h->h_state[2];
    iter.aborted = code == BICG_NAN;
    iter.converged = h->h_state[0] != 0 && code == BICG_RUNNING;
    if (P.ibj()) {
      int ierr = 0;
      HIPCHK(h, hipMemcpy(&ierr, h->d_ibj_stat, sizeof(int), hipMemcpyDeviceToHost));
      if (ierr != 0)
        status = fail(h, PDELAB_ERR_SOLVER_ABORT, "iterative block Jacobi: breakdown in a block solve (code " +
                                                      std::to_string(ierr) + ")");
    }
    if (code == BICG_RHO) status = fail(h, PDELAB_ERR_SOLVER_ABORT, "breakdown in BiCGSTAB - rho");
    else if (code == BICG_OMEGA) status = fail(h, PDELAB_ERR_SOLVER_ABORT, "breakdown in BiCGSTAB - omega");
    else if (code == BICG_H) status = fail(h, PDELAB_ERR_SOLVER_ABORT, "abs(h) < EPSILON in BiCGSTAB");
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  iter.finalize(res, el);
  if (status != PDELAB_OK) { res->converged = 0; return status; }
  if (iter.aborted) return fail(h, PDELAB_ERR_SOLVER_ABORT, "BiCGSTAB: defect is infinite or NaN");
  return PDELAB_OK;
}

template <typename T, typename A>
static int krylov(H* h, int method, int precond, double reduction, int maxit, bool trans, T* z, T* r, int base,
                  pdelab_linear_solver_result* res) {
  if (maxit <= 0) maxit = 5000;
  if (method == PDELAB_SOLVER_CG && precond == PDELAB_PREC_BLOCK_SOR)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "block SOR: BiCGSTAB only (the preconditioner is not symmetric)");
  if (method == PDELAB_SOLVER_CG && precond == PDELAB_PREC_ITERATIVE_BLOCK_JACOBI)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "iterative block Jacobi: BiCGSTAB only (the preconditioner is an inner iteration)");
  if (method == PDELAB_SOLVER_CG) {
    if constexpr (std::is_same<T, A>::value) return cg_solve<T, A>(h, z, r, reduction, maxit, precond, trans, base, res);
    else return fail(h, PDELAB_ERR_UNSUPPORTED, "CG accumulates in its vector type");
  }
  if (method == PDELAB_SOLVER_BICGSTAB) return bicgstab_solve<T, A>(h, z, r, reduction, maxit, precond, trans, base, res);
  return fail(h, PDELAB_ERR_INVALID_ARG, "unknown Krylov method");
}

static int validate(const pdelab_config* c, std::string& msg) {
  if (!c) { msg = "null config"; return PDELAB_ERR_INVALID_ARG; }
  if (c->abi_version != PDELAB_GPU_ABI_VERSION) { msg = "ABI version mismatch"; return PDELAB_ERR_INVALID_ARG; }
  if (c->nx < 1 || c->ny < 1 || c->nz < 1) { msg = "cells per direction must be >= 1"; return PDELAB_ERR_INVALID_ARG; }
  if (!(c->lx > 0 && c->ly > 0 && c->lz > 0)) { msg = "domain extents must be > 0"; return PDELAB_ERR_INVALID_ARG; }
  if (c->degree < 1 || c->degree > 3) { msg = "degree must be 1, 2 or 3"; return PDELAB_ERR_UNSUPPORTED; }
  if (c->method < 0 || c->method > 2) { msg = "unknown DG method"; return PDELAB_ERR_INVALID_ARG; }
  if (c->weights < 0 || c->weights > 1) { msg = "unknown weights"; return PDELAB_ERR_INVALID_ARG; }
  if (c->intorderadd < 0) { msg = "intorderadd must be >= 0"; return PDELAB_ERR_INVALID_ARG; }
  if (c->problem < 0 || c->problem > PDELAB_PROBLEM_TABULATED) { msg = "unknown problem"; return PDELAB_ERR_INVALID_ARG; }
  if (c->n_gaussians < 0 || c->n_gaussians > PDELAB_MAX_GAUSSIANS) { msg = "n_gaussians"; return PDELAB_ERR_INVALID_ARG; }
  for (int s = 0; s < 6; ++s) {
    int b = c->bctype[s];
    if (!(b == 0 || b == PDELAB_BC_DIRICHLET || b == PDELAB_BC_NEUMANN || b == PDELAB_BC_OUTFLOW || b == PDELAB_BC_NONE)) {
      msg = "unknown boundary condition type";
      return PDELAB_ERR_INVALID_ARG;
    }
  }
  if (c->kappa_per_cell < 0 || c->kappa_per_cell > 2) { msg = "kappa_per_cell must be 0, 1 or 2"; return PDELAB_ERR_INVALID_ARG; }
  if (c->kappa_per_cell == 0 && !(c->kappa0 > 0)) { msg = "kappa0 must be > 0"; return PDELAB_ERR_INVALID_ARG; }
  return PDELAB_OK;
}

// per-cell face coefficients of the current kappa (FP64 and FP32) for the apply kernels
static int update_face_coeffs(H* h) {
  const long long nc = h->ncells;
  if (!h->d_fc64) HIPCHK(h, hipMalloc(&h->d_fc64, sizeof(double) * FC * nc));
  if (!h->d_fc32) HIPCHK(h, hipMalloc(&h->d_fc32, sizeof(float) * FC * nc));
  const int nb = (int)((nc + 255) / 256);
  hipLaunchKernelGGL(face_coeff_kernel<double>, dim3(nb), dim3(256), 0, h->stream, (const double*)h->d_kappa,
                     h->d_fc64, h->a64);
  hipLaunchKernelGGL(face_coeff_kernel<float>, dim3(nb), dim3(256), 0, h->stream, (const double*)h->d_kappa,
                     h->d_fc32, h->a64);
  HIPCHK(h, hipGetLastError());
  bj_invalidate(h);
  return PDELAB_OK;
}

static int upload_kappa(H* h, const double* kappa_host) {
  // one value per cell, or a_x, a_y, a_z per cell (kappa_per_cell == 2)
  const size_t nc = (size_t)h->ncells * (h->a64.aniso ? 3 : 1);
  for (size_t i = 0; i < nc; ++i)
    if (!(kappa_host[i] > 0) || !std::isfinite(kappa_host[i]))
      return fail(h, PDELAB_ERR_INVALID_ARG, "kappa must be finite and > 0");
  if (!h->d_kappa) HIPCHK(h, hipMalloc(&h->d_kappa, sizeof(double) * nc));
  if (!h->d_kappa32) HIPCHK(h, hipMalloc(&h->d_kappa32, sizeof(float) * nc));
  HIPCHK(h, hipMemcpyAsync(h->d_kappa, kappa_host, sizeof(double) * nc, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for((long long)nc)), dim3(256), 0, h->stream,
                     (long long)nc, (const double*)h->d_kappa, h->d_kappa32);
  HIPCHK(h, hipGetLastError());
  CHK(update_face_coeffs(h));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->vark = true;
  h->R0_valid = false;
  h->diag_valid = false;
  return PDELAB_OK;
}

// ===========================================================================
// Mixed precision with a flexible outer Krylov iteration: FGMRES(m) (Saad 1993, "A flexible
// inner-outer preconditioned GMRES algorithm") in FP64 on J d = R(x), right-preconditioned
// by the FP32 Krylov solve (CG or BiCGSTAB with its own preconditioner) of J32 z = v to
// inner_reduction from z = 0. v has unit norm, so the FP32 solve needs no scaling. Unlike a
// restarted refinement, the outer iteration keeps every FP32 correction in its Krylov space:
// each outer step minimises the FP64 residual over all corrections of the cycle. The Arnoldi
// basis is orthogonalised by modified Gram-Schmidt in FP64. x -= d at the end of a cycle, and
// each cycle starts from the true FP64 residual R(x) = J x + R(0), which also decides
// convergence (||R(x)|| <= outer_reduction ||R(x0)||). outer_iterations counts the FP64
// outer steps (one FP64 apply and one FP32 solve each).
static int fgmres_mixed(H* h, int method, int precond, double inner_reduction, int inner_maxit,
                        double outer_reduction, int outer_maxit, double* x, pdelab_mixed_result* res) {
  constexpr int m = H::kFgmresRestart;
  const long long n = h->N;
  for (double*& v : h->fg)
    if (!v) HIPCHK(h, hipMalloc(&v, sizeof(double) * n));
  CHK(ensure_ws<float>(h, 2));
  double** V = h->fg;      // V[0..m]
  double** Z = h->fg + m + 1;  // Z[0..m-1]
  float* r32 = h->w32[0];
  float* z32 = h->w32[1];
  const dim3 g(grid_for(n)), b(256);
  double def0 = -1, def = 0;
  int outer = 0, inner = 0;
  bool conv = false;
  std::vector<double> Hm((m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
  for (;;) {
    // true FP64 residual r = J x + R(0) into V[0]
    CHK(apply64(h, x, V[0], false, false));
    hipLaunchKernelGGL(axpy_kernel<double>, g, b, 0, h->stream, n, 1.0, (const double*)h->d_R0, V[0]);
    CHK(pdelab_gpu_norm(h, V[0], &def));
    if (!std::isfinite(def)) {
      res->defect = def;
      return fail(h, PDELAB_ERR_SOLVER_ABORT, "flexible refinement: residual is infinite or NaN");
    }
    if (def0 < 0) def0 = def;
    if (def <= outer_reduction * def0 || def < 1e-30) { conv = true; break; }
    if (outer >= outer_maxit) break;
    hipLaunchKernelGGL(scale_kernel<double>, g, b, 0, h->stream, n, 1.0 / def, V[0]);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = def;
    const double target = outer_reduction * def0;
    int kdone = 0;
    for (int j = 0; j < m && outer < outer_maxit; ++j) {
      // z_j = J32^{-1} v_j in FP32, then w = J z_j in FP64 (into V[j+1])
      hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, h->stream, n, (const double*)V[j], r32);
      hipLaunchKernelGGL(fill_kernel<float>, g, b, 0, h->stream, n, 0.0f, z32);
      HIPCHK(h, hipGetLastError());
      pdelab_linear_solver_result ires;
      const int st = krylov<float, float>(h, method, precond, inner_reduction, inner_maxit, false, z32, r32, 2, &ires);
      inner += (int)ires.iterations;
      if (st == PDELAB_ERR_SOLVER_ABORT) h->err.clear();  // a breakdown still leaves a usable direction
      else CHK(st);
      hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, h->stream, n, (const float*)z32, Z[j]);
      HIPCHK(h, hipGetLastError());
      double* w = V[j + 1];
      CHK(apply64(h, Z[j], w, false, false));
      // modified Gram-Schmidt
      for (int i = 0; i <= j; ++i) {
        double hij = 0;
        CHK(pdelab_gpu_dot(h, w, V[i], &hij));
        Hm[i * m + j] = hij;
        hipLaunchKernelGGL(axpy_kernel<double>, g, b, 0, h->stream, n, -hij, (const double*)V[i], w);
      }
      double hn = 0;
      CHK(pdelab_gpu_norm(h, w, &hn));
      Hm[(j + 1) * m + j] = hn;
      if (hn > 0) hipLaunchKernelGGL(scale_kernel<double>, g, b, 0, h->stream, n, 1.0 / hn, w);
      HIPCHK(h, hipGetLastError());
      // Givens rotations of column j and the residual estimate |g_{j+1}|
      for (int i = 0; i < j; ++i) {
        const double a0 = Hm[i * m + j], a1 = Hm[(i + 1) * m + j];
        Hm[i * m + j] = cs[i] * a0 + sn[i] * a1;
        Hm[(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
      }
      const double a0 = Hm[j * m + j], a1 = Hm[(j + 1) * m + j];
      const double rr = std::hypot(a0, a1);
      cs[j] = rr > 0 ? a0 / rr : 1.0;
      sn[j] = rr > 0 ? a1 / rr : 0.0;
      Hm[j * m + j] = rr;
      Hm[(j + 1) * m + j] = 0.0;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      ++outer;
      kdone = j + 1;
      if (std::fabs(gv[j + 1]) <= 0.5 * target || hn == 0.0) break;
    }
    // y = H^{-1} g (upper triangular), x -= sum_i y_i z_i
    for (int i = kdone - 1; i >= 0; --i) {
      double sum = gv[i];
      for (int k2 = i + 1; k2 < kdone; ++k2) sum -= Hm[i * m + k2] * y[k2];
      y[i] = Hm[i * m + i] != 0.0 ? sum / Hm[i * m + i] : 0.0;
    }
    for (int i = 0; i < kdone; ++i)
      hipLaunchKernelGGL(axpy_kernel<double>, g, b, 0, h->stream, n, -y[i], (const double*)Z[i], x);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  res->converged = conv ? 1 : 0;
  res->outer_iterations = outer;
  res->inner_iterations = inner;
  res->first_defect = def0;
  res->defect = def;
  res->reduction = def0 > 0 ? def / def0 : 0.0;
  return PDELAB_OK;
}

// ===========================================================================
// Mixed-precision CG with reliable updates (residual replacement): one FP32 CG run on
// J z = R(x) whose recursive residual is replaced by the FP64 residual whenever it has
// dropped by inner_reduction (or by the reduction still missing). At each replacement
// the FP32 correction is subtracted from x in FP64, R(x) is recomputed in FP64 and
// converted, and the search direction is kept (beta from the replaced residual), so
// the Krylov space is not rebuilt as a restarted refinement would rebuild it.
static int cg_mixed_reliable(H* h, int precond, double inner_reduction, int inner_maxit, double outer_reduction,
                             int outer_maxit, double* x, pdelab_mixed_result* res) {
  const long long n = h->N;
  Prec<float> P;
  CHK(make_prec<float>(h, precond, false, P));
  CHK(ensure_ws<double>(h, 6));
  CHK(ensure_ws<float>(h, 5));
  double* r64 = h->w64[5];
  float *r32 = h->w32[0], *x32 = h->w32[1], *p32 = h->w32[2], *q32 = h->w32[3], *z32 = h->w32[4];
  auto residual64 = [&](double& d) -> int {
    CHK(apply64(h, x, r64, false, false));
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0,
                       (const double*)h->d_R0, r64);
    CHK(pdelab_gpu_norm(h, r64, &d));
    if (!std::isfinite(d)) return fail(h, PDELAB_ERR_SOLVER_ABORT, "refinement: residual is infinite or NaN");
    return PDELAB_OK;
  };
  // the FP32 recursion runs on the residual at unit norm: r32 = r / scale with scale the FP64
  // defect at the last replacement, so its dot products stay far from the FP32 underflow
  // threshold however small the outer residual has become; x32 holds the correction / scale
  auto to_f32 = [&](double scale) -> int {
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0 / scale,
                       (const double*)r64, r32);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  double def0 = 0, def = 0;
  CHK(residual64(def0));
  def = def0;
  res->first_defect = def0;
  int outer = 0, inner = 0;
  bool conv = def0 < 1e-30;
  if (!conv) {
    double scale = def0;
    CHK(to_f32(scale));
    hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 0.0f, x32);
    // p = W^{-1} r; rho -> slot 7
    CHK((prec_dot<float, float>(h, P, r32, p32, 7)));
    for (;;) {
      const double thr = std::max(inner_reduction * def, 0.5 * outer_reduction * def0);
      CgLoopOut lo;
      CHK((cg_loop<float, float>(h, x32, r32, p32, q32, z32, P, false, std::max(1, inner_maxit - inner), 1.0,
                                 thr / scale, lo)));
      inner += lo.iters;
      if (lo.aborted) {
        res->inner_iterations = inner;
        return fail(h, PDELAB_ERR_SOLVER_ABORT, "mixed CG: FP32 defect is infinite or NaN");
      }
      // residual replacement: x -= scale x32, r = R(x) in FP64
      hipLaunchKernelGGL(sub_scaled_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, scale,
                         (const float*)x32, x);
      hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 0.0f, x32);
      HIPCHK(h, hipGetLastError());
      CHK(residual64(def));
      ++outer;
      if (def <= outer_reduction * def0 || def < 1e-30) {
        conv = true;
        break;
      }
      if (outer >= outer_maxit || inner >= inner_maxit) break;
      // continue the direction in the new scale: with r32 = r / scale the CG recurrence
      // p = W^{-1} r + (rho_new / rho_old) p reads p_new = q + (scale_new / scale_old)
      // (rho32_new / rho32_old) p_old in scaled units
      float sc[8];
      CHK((read_scalars<float>(h, sc, 8)));
      const double rho_old = sc[lo.rho_slot];
      const double scale_old = scale;
      scale = def;
      CHK(to_f32(scale));
      CHK((prec_dot<float, float>(h, P, r32, q32, 7)));
      CHK((read_scalars<float>(h, sc, 8)));
      const float beta = rho_old != 0 ? (float)((scale / scale_old) * (sc[7] / rho_old)) : 0.0f;
      // p = q + beta p with q = W^{-1} r
      hipLaunchKernelGGL(cg_pupdate_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, p32,
                         (const float*)q32, (const float*)nullptr, beta);
      HIPCHK(h, hipGetLastError());
    }
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  res->converged = conv ? 1 : 0;
  res->outer_iterations = outer;
  res->inner_iterations = inner;
  res->defect = def;
  res->reduction = def0 > 0 ? def / def0 : 0.0;
  return PDELAB_OK;
}

// Mixed-precision BiCGSTAB with reliable updates: one FP32 BiCGSTAB run on J z = R(x) whose
// recursive residual is replaced by the FP64 residual whenever it has dropped by
// inner_reduction (or by the reduction still missing), as cg_mixed_reliable does for CG.
// The FP32 vectors hold the residual at unit norm (r32 = r / scale); at a replacement the
// correction is subtracted from x in FP64, R(x) is recomputed and rescaled, and the recurrence
// goes on: p, v, the shadow residual and rho_old stay in the old scale, which cancels in
// beta = (rho_new / rho_old)(alpha / omega) applied to p - omega v. A run stopped after the
// first half step resumes with its second half (omega is scale-free); after a full step
// rho_new = rt.r is taken from the replaced residual. A breakdown restarts the recurrence
// with the shadow residual set to the current residual.
static int bicgstab_mixed_reliable(H* h, int precond, double inner_reduction, int inner_maxit, double outer_reduction,
                                   int outer_maxit, double* x, pdelab_mixed_result* res) {
  using T = float;
  using A = float;
  const long long n = h->N;
  Prec<T> P;
  CHK(make_prec<T>(h, precond, false, P));
  CHK(ensure_ws<double>(h, 6));
  CHK(ensure_ws<T>(h, 7));
  double* r64 = h->w64[5];
  T *r = h->w32[0], *xc = h->w32[1], *p = h->w32[2], *v = h->w32[3], *t = h->w32[4], *y = h->w32[5],
    *rt = h->w32[6];
  if (!h->d_bscal) HIPCHK(h, hipMalloc(&h->d_bscal, sizeof(double) * 8));
  int* st = h->d_state;
  double* ds = h->d_dstate;
  double* bs = h->d_bscal;
  A* part = partial_buf<A>(h);
  const dim3 rg(kRedBlocks), rb(kRedThreads), fg(1), fb(kFinalThreads);
  auto residual64 = [&](double& d) -> int {
    CHK(apply64(h, x, r64, false, false));
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0,
                       (const double*)h->d_R0, r64);
    CHK(pdelab_gpu_norm(h, r64, &d));
    if (!std::isfinite(d)) return fail(h, PDELAB_ERR_SOLVER_ABORT, "refinement: residual is infinite or NaN");
    return PDELAB_OK;
  };
  auto to_f32 = [&](double scale) -> int {
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0 / scale,
                       (const double*)r64, r);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  auto prec = [&](const T* in, T* out) -> int {
    if (P.sor()) return sor_launch<T>(h, in, out, st, false);
    if (P.ibj()) return ibj_launch<T>(h, in, out, static_cast<int*>(h->d_ibj_stat), nullptr, st, false);
    if (P.block()) {
      const BJVec<T, A> bv{in, out, nullptr, nullptr, nullptr, nullptr, st};
      return bj_launch<T, A>(h, false, BJ_PREC, bv, nullptr);
    }
    hipLaunchKernelGGL(prec_stop_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, P.dinv, in, out,
                       (const int*)st);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  // a fresh recurrence: rt = r, rho_new = rt.r, p = r at the first step
  auto start = [&]() -> int {
    HIPCHK(h, hipMemcpyAsync(rt, r, sizeof(T) * n, hipMemcpyDeviceToDevice, h->stream));
    CHK((dot_dev<T, A>(h, r, r, n, 0)));
    A sc[1];
    CHK((read_scalars<A>(h, sc, 1)));
    hipLaunchKernelGGL(bicg_init_kernel, dim3(1), dim3(64), 0, h->stream, bs, (double)sc[0]);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  if (P.ibj()) HIPCHK(h, hipMemsetAsync(h->d_ibj_stat, 0, 16, h->stream));
  double def0 = 0, def = 0;
  CHK(residual64(def0));
  def = def0;
  res->first_defect = def0;
  int outer = 0, it = 0;
  long long halves = 0;
  bool conv = def0 < 1e-30;
  if (!conv) {
    double scale = def0;
    CHK(to_f32(scale));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, T(0), xc);
    CHK(start());
    bool first = true, resume_second = false;
    for (;;) {
      // one run of the FP32 recurrence, stopped by the device when its (unit-scaled) defect
      // falls below thr / scale
      const double thr = std::max(inner_reduction * def, 0.5 * outer_reduction * def0);
      h->h_state[0] = h->h_state[1] = h->h_state[2] = h->h_state[3] = 0;
      h->h_dstate[0] = 1.0;
      h->h_dstate[1] = 1.0;
      h->h_dstate[2] = thr / scale;
      h->h_dstate[3] = 0.0;
      HIPCHK(h, hipMemcpyAsync(st, h->h_state, sizeof(int) * 4, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(ds, h->h_dstate, sizeof(double) * 4, hipMemcpyHostToDevice, h->stream));
      const bool resumed = resume_second;
      bool stopped = false;
      int seg = 0;
      while (!stopped && it < inner_maxit) {
        ++it;
        ++seg;
        if (!resume_second) {
          hipLaunchKernelGGL(bicg_p_dev_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, p, (const T*)r,
                             (const T*)v, (const double*)bs, (const int*)st, first ? 1 : 0);
          CHK(prec(p, y));
          CHK(apply_mode<T>(h, y, v, false, 0, nullptr, st));
          hipLaunchKernelGGL((dot_partial_kernel<T, A>), rg, rb, 0, h->stream, (const T*)rt, (const T*)v, n, part);
          hipLaunchKernelGGL((bicg_scalar_kernel<A, 0>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                             2 * seg - 1);
          hipLaunchKernelGGL((bicg_xr_kernel<T, A, 1>), rg, rb, 0, h->stream, n, xc, (const T*)y, r, (const T*)v,
                             (const T*)nullptr, (const double*)bs, kBsAlpha, part, (const int*)st);
          hipLaunchKernelGGL((bicg_scalar_kernel<A, 1>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                             2 * seg - 1);
          first = false;
        }
        resume_second = false;
        CHK(prec(r, y));
        CHK(apply_mode<T>(h, y, t, false, 0, nullptr, st));
        hipLaunchKernelGGL((dot2_kernel<T, A>), rg, rb, 0, h->stream, n, (const T*)t, (const T*)r, part);
        hipLaunchKernelGGL((bicg_scalar_kernel<A, 2>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                           2 * seg);
        hipLaunchKernelGGL((bicg_xr_kernel<T, A, 2>), rg, rb, 0, h->stream, n, xc, (const T*)y, r, (const T*)t,
                           (const T*)rt, (const double*)bs, kBsOmega, part, (const int*)st);
        hipLaunchKernelGGL((bicg_scalar_kernel<A, 3>), fg, fb, 0, h->stream, (const A*)part, kRedBlocks, bs, st, ds,
                           2 * seg);
        HIPCHK(h, hipGetLastError());
        if (seg % kBicgCheckEvery == 0 || it >= inner_maxit) {
          HIPCHK(h, hipMemcpyAsync(h->h_state, st, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
          HIPCHK(h, hipStreamSynchronize(h->stream));
          stopped = h->h_state[0] != 0;
        }
      }
      HIPCHK(h, hipMemcpyAsync(h->h_state, st, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
      stopped = h->h_state[0] != 0;
      const int code = h->h_state[1];
      const int half = stopped ? h->h_state[2] : 2 * seg;  // the half step at which the run stopped
      halves += half - (resumed ? 1 : 0);
      if (stopped) it -= seg - (half + 1) / 2;  // iterations queued after the stop did nothing
      if (code == BICG_NAN) {
        res->inner_iterations = it;
        return fail(h, PDELAB_ERR_SOLVER_ABORT, "mixed BiCGSTAB: FP32 defect is infinite or NaN");
      }
      // residual replacement: x -= scale xc, r = R(x) in FP64
      hipLaunchKernelGGL(sub_scaled_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, scale, (const T*)xc,
                         x);
      hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, T(0), xc);
      HIPCHK(h, hipGetLastError());
      CHK(residual64(def));
      ++outer;
      if (def <= outer_reduction * def0 || def < 1e-30) {
        conv = true;
        break;
      }
      if (outer >= outer_maxit || it >= inner_maxit) break;
      scale = def;
      CHK(to_f32(scale));
      if (code != BICG_RUNNING) {  // a breakdown of the FP32 recurrence: start it again
        CHK(start());
        first = true;
        continue;
      }
      if (half % 2 == 1) {  // stopped after the first half step: its second half follows
        resume_second = true;
        continue;
      }
      // stopped after a full step: rho_new from the replaced residual, then beta
      CHK((dot_dev<T, A>(h, rt, r, n, 0)));
      A sc[1];
      CHK((read_scalars<A>(h, sc, 1)));
      double b[5];
      HIPCHK(h, hipMemcpy(b, bs, sizeof(b), hipMemcpyDeviceToHost));
      b[kBsRhoNew] = (double)sc[0];
      if (std::fabs(b[kBsRho]) <= kBicgEps || std::fabs(b[kBsOmega]) <= kBicgEps) {
        CHK(start());
        first = true;
        continue;
      }
      b[kBsBeta] = (b[kBsRhoNew] / b[kBsRho]) * (b[kBsAlpha] / b[kBsOmega]);
      HIPCHK(h, hipMemcpy(bs, b, sizeof(b), hipMemcpyHostToDevice));
    }
  }
  if (P.ibj()) {
    int ierr = 0;
    HIPCHK(h, hipMemcpy(&ierr, h->d_ibj_stat, sizeof(int), hipMemcpyDeviceToHost));
    if (ierr != 0)
      return fail(h, PDELAB_ERR_SOLVER_ABORT,
                  "iterative block Jacobi: breakdown in a block solve (code " + std::to_string(ierr) + ")");
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  res->converged = conv ? 1 : 0;
  res->outer_iterations = outer;
  res->inner_iterations = (int)((halves + 1) / 2);
  res->defect = def;
  res->reduction = def0 > 0 ? def / def0 : 0.0;
  return PDELAB_OK;
}

extern "C" {

void pdelab_gpu_default_config(pdelab_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->abi_version = PDELAB_GPU_ABI_VERSION;
  cfg->nx = cfg->ny = cfg->nz = 4;
  cfg->lx = cfg->ly = cfg->lz = 1.0;
  cfg->degree = 1;
  cfg->method = PDELAB_SIPG;
  cfg->weights = PDELAB_WEIGHTS_ON;
  cfg->alpha = 1.0;
  cfg->intorderadd = 0;
  cfg->kappa_per_cell = 0;
  cfg->kappa0 = 1.0;
  cfg->problem = PDELAB_PROBLEM_ZERO;
}

static int init_scratch(H* h);  // device address of the no-allocation scratch array

int pdelab_gpu_create(const pdelab_config* cfg, const double* kappa_host, pdelab_gpu_gridoperator** out) {
  g_create_error.clear();
  if (!out) { g_create_error = "null output pointer"; return PDELAB_ERR_INVALID_ARG; }
  *out = nullptr;
  std::string msg;
  int st = validate(cfg, msg);
  if (st) { g_create_error = msg; return st; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_create_error = "no HIP device"; return PDELAB_ERR_NO_DEVICE; }
  H* h = new H;
  h->cfg = *cfg;
  if (!build_tables(cfg->degree, cfg->intorderadd, h->tab)) {
    g_create_error = "table construction failed";
    delete h;
    return PDELAB_ERR_UNSUPPORTED;
  }
  h->n1 = cfg->degree + 1;
  h->ncells = (long long)cfg->nx * cfg->ny * cfg->nz;
  h->nloc = (long long)h->n1 * h->n1 * h->n1;
  h->N = h->ncells * h->nloc;
  h->conv = cfg->b[0] != 0.0 || cfg->b[1] != 0.0 || cfg->b[2] != 0.0;
  (void)hipGetDevice(&h->device);
  auto bail = [&](const std::string& m) {
    g_create_error = m;
    pdelab_gpu_destroy(h);
    return PDELAB_ERR_HIP;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail("hipStreamCreate failed");
  fill_args<double>(h, h->a64);
  fill_args<float>(h, h->a32);
  fill_load_args(h);
  choose_apply_shape(h);
  if (build_line_mats(h) != PDELAB_OK) {
    const int code = h->err.find("rank-two") != std::string::npos ? PDELAB_ERR_INVALID_ARG : PDELAB_ERR_HIP;
    g_create_error = h->err;
    pdelab_gpu_destroy(h);
    return code;
  }
  if (hipMalloc(&h->d_part64, sizeof(double) * kRedBlocks * 4) != hipSuccess) return bail("hipMalloc failed");
  if (hipMalloc(&h->d_part32, sizeof(float) * kRedBlocks * 4) != hipSuccess) return bail("hipMalloc failed");
  if (hipMalloc(&h->d_scal64, sizeof(double) * 16) != hipSuccess) return bail("hipMalloc failed");
  if (hipMalloc(&h->d_scal32, sizeof(float) * 16) != hipSuccess) return bail("hipMalloc failed");
  if (hipHostMalloc(&h->h_scal64, sizeof(double) * 16) != hipSuccess) return bail("hipHostMalloc failed");
  if (hipHostMalloc(&h->h_scal32, sizeof(float) * 16) != hipSuccess) return bail("hipHostMalloc failed");
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return bail("hipEventCreate failed");
  if (hipMalloc(&h->d_state, sizeof(int) * 4) != hipSuccess) return bail("hipMalloc failed");
  if (hipMalloc(&h->d_dstate, sizeof(double) * 4) != hipSuccess) return bail("hipMalloc failed");
  if (hipHostMalloc(&h->h_state, sizeof(int) * 4) != hipSuccess) return bail("hipHostMalloc failed");
  if (hipHostMalloc(&h->h_dstate, sizeof(double) * 4) != hipSuccess) return bail("hipHostMalloc failed");
  if (init_scratch(h) != PDELAB_OK) return bail(h->err);
  if (cfg->kappa_per_cell) {
    if (!kappa_host) { g_create_error = "kappa_per_cell requires a kappa array"; pdelab_gpu_destroy(h); return PDELAB_ERR_INVALID_ARG; }
    if (upload_kappa(h, kappa_host) != PDELAB_OK) {
      g_create_error = h->err;
      pdelab_gpu_destroy(h);
      return PDELAB_ERR_INVALID_ARG;
    }
  }
  *out = h;
  return PDELAB_OK;
}

int pdelab_gpu_destroy(pdelab_gpu_gridoperator* h) {
  if (!h) return PDELAB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ptrs[] = {h->d_kappa, h->d_kappa32, h->d_R0, h->d_src, h->d_ftab, h->d_gtab, h->d_diag, h->d_dinv,
                  h->d_dinv32, h->d_part64, h->d_part32, h->d_scal64, h->d_scal32, h->d_applydot,
                  h->d_state, h->d_dstate, h->d_bscal, h->d_lm64, h->d_lm32, h->d_fc64, h->d_fc32};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (double* p : h->os)
    if (p) (void)hipFree(p);
  for (double* p : h->fg)
    if (p) (void)hipFree(p);
  bj_free(h);
  vel_free(h);
  for (int i = 0; i < 8; ++i) {
    if (h->w64[i]) (void)hipFree(h->w64[i]);
    if (h->w32[i]) (void)hipFree(h->w32[i]);
  }
  if (h->h_scal64) (void)hipHostFree(h->h_scal64);
  if (h->h_scal32) (void)hipHostFree(h->h_scal32);
  if (h->h_state) (void)hipHostFree(h->h_state);
  if (h->h_dstate) (void)hipHostFree(h->h_dstate);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PDELAB_OK;
}

const char* pdelab_gpu_last_error(const pdelab_gpu_gridoperator* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}
size_t pdelab_gpu_size(const pdelab_gpu_gridoperator* h) { return h ? (size_t)h->N : 0; }
size_t pdelab_gpu_num_cells(const pdelab_gpu_gridoperator* h) { return h ? (size_t)h->ncells : 0; }
int pdelab_gpu_is_linear(const pdelab_gpu_gridoperator* h) { return h ? 1 : 0; }
int pdelab_gpu_quadrature_points(const pdelab_gpu_gridoperator* h) { return h ? h->tab.m : 0; }

int pdelab_gpu_set_velocity_field(pdelab_gpu_gridoperator* h, const double* b_vol, const double* b_face) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  return vel_set(h, b_vol, b_face);
}

int pdelab_gpu_set_kappa(pdelab_gpu_gridoperator* h, const double* kappa_host) {
  if (!h || !kappa_host) return PDELAB_ERR_INVALID_ARG;
  return upload_kappa(h, kappa_host);
}

int pdelab_gpu_set_log_kappa_device(pdelab_gpu_gridoperator* h, const double* logk) {
  if (!h || !logk) return PDELAB_ERR_INVALID_ARG;
  if (h->a64.aniso) return fail(h, PDELAB_ERR_UNSUPPORTED, "log kappa: isotropic A = kappa I only");
  const long long nc = h->ncells;
  if (!h->d_kappa) HIPCHK(h, hipMalloc(&h->d_kappa, sizeof(double) * nc));
  if (!h->d_kappa32) HIPCHK(h, hipMalloc(&h->d_kappa32, sizeof(float) * nc));
  hipLaunchKernelGGL(exp_kernel, dim3(grid_for(nc)), dim3(256), 0, h->stream, nc, logk, h->d_kappa, h->d_kappa32);
  HIPCHK(h, hipGetLastError());
  CHK(update_face_coeffs(h));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->vark = true;
  h->R0_valid = false;
  h->diag_valid = false;
  return PDELAB_OK;
}

int pdelab_gpu_set_cell_source(pdelab_gpu_gridoperator* h, const double* src) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  if (!src) {
    if (h->d_src) { (void)hipFree(h->d_src); h->d_src = nullptr; }
  } else {
    if (!h->d_src) HIPCHK(h, hipMalloc(&h->d_src, sizeof(double) * h->ncells));
    HIPCHK(h, hipMemcpyAsync(h->d_src, src, sizeof(double) * h->ncells, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->R0_valid = false;
  return PDELAB_OK;
}

int pdelab_gpu_set_tabulated_data(pdelab_gpu_gridoperator* h, const double* fq, const double* gq) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  if (h->cfg.problem != PDELAB_PROBLEM_TABULATED)
    return fail(h, PDELAB_ERR_INVALID_ARG, "tabulated data needs problem PDELAB_PROBLEM_TABULATED");
  const long long m = h->tab.m;
  const long long nf = h->ncells * m * m * m;
  const long long ng = h->la.bface_offset[6] * m * m;
  if (fq) {
    if (!h->d_ftab) HIPCHK(h, hipMalloc(&h->d_ftab, sizeof(double) * nf));
    HIPCHK(h, hipMemcpyAsync(h->d_ftab, fq, sizeof(double) * nf, hipMemcpyHostToDevice, h->stream));
  }
  if (gq) {
    if (!h->d_gtab) HIPCHK(h, hipMalloc(&h->d_gtab, sizeof(double) * ng));
    HIPCHK(h, hipMemcpyAsync(h->d_gtab, gq, sizeof(double) * ng, hipMemcpyHostToDevice, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->R0_valid = false;
  return PDELAB_OK;
}

int pdelab_gpu_malloc(pdelab_gpu_gridoperator* h, void** dptr, size_t bytes) {
  if (!h || !dptr) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipMalloc(dptr, bytes ? bytes : 1));
  return PDELAB_OK;
}
int pdelab_gpu_free(pdelab_gpu_gridoperator* h, void* dptr) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (dptr) HIPCHK(h, hipFree(dptr));
  return PDELAB_OK;
}
int pdelab_gpu_memcpy_htod(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PDELAB_OK;
}
int pdelab_gpu_memcpy_dtoh(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PDELAB_OK;
}
int pdelab_gpu_memcpy_dtod(pdelab_gpu_gridoperator* h, void* dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!dst || !src))) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
  return PDELAB_OK;
}
int pdelab_gpu_memset_zero(pdelab_gpu_gridoperator* h, void* dptr, size_t bytes) {
  if (!h || (bytes && !dptr)) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipMemsetAsync(dptr, 0, bytes, h->stream));
  return PDELAB_OK;
}
int pdelab_gpu_synchronize(pdelab_gpu_gridoperator* h) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PDELAB_OK;
}

int pdelab_gpu_residual(pdelab_gpu_gridoperator* h, const double* x, double* r) {
  if (!h || !x || !r) return PDELAB_ERR_INVALID_ARG;
  CHK(compute_R0(h));
  CHK(apply64(h, x, r, false, true));
  hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(h->N)), dim3(256), 0, h->stream, h->N, 1.0,
                     (const double*)h->d_R0, r);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_jacobian_apply(pdelab_gpu_gridoperator* h, const double* z, double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  return apply64(h, z, y, false, true);
}

int pdelab_gpu_operator_apply(pdelab_gpu_gridoperator* h, const double* z, double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  return apply64(h, z, y, false, false);
}

int pdelab_gpu_jacobian_apply_f32(pdelab_gpu_gridoperator* h, const float* z, float* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  return apply32(h, z, y, false, true);
}

int pdelab_gpu_jacobian_apply_nonlinear(pdelab_gpu_gridoperator* h, const double*, const double*, double*) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  return fail(h, PDELAB_ERR_LINEARITY_MISMATCH, "Your trying to use a non linear jacobian apply for a linear problem.");
}

int pdelab_gpu_jacobian_apply_transpose(pdelab_gpu_gridoperator* h, const double* z, double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  return apply64(h, z, y, true, true);
}

int pdelab_gpu_point_diagonal(pdelab_gpu_gridoperator* h, double* d) {
  if (!h || !d) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  sync_prec_cache(h);
  CHK(compute_diag(h));
  hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(h->N)), dim3(256), 0, h->stream, h->N, 1.0,
                     (const double*)h->d_diag, d);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_block_diagonal(pdelab_gpu_gridoperator* h, int transpose, double* blocks) {
  if (!h || !blocks) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  return bj_block_diagonal(h, transpose != 0, blocks);
}

int pdelab_gpu_jacobian(pdelab_gpu_gridoperator* h, int transpose, double* blocks, long long* cols) {
  if (!h || !blocks || !cols) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  return jac_assemble(h, transpose != 0, blocks, cols);
}

int pdelab_gpu_block_matrix_apply(pdelab_gpu_gridoperator* h, const double* blocks, const long long* cols,
                                  const double* z, double* y) {
  if (!h || !blocks || !cols || !z || !y) return PDELAB_ERR_INVALID_ARG;
  return jac_block_apply(h, blocks, cols, z, y);
}

int pdelab_gpu_block_jacobi_apply(pdelab_gpu_gridoperator* h, const double* z, double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  sync_prec_cache(h);
  const BJVec<double, double> v{z, y, nullptr, nullptr, nullptr, nullptr, nullptr};
  return bj_launch<double, double>(h, false, BJ_ADD, v, nullptr);
}

int pdelab_gpu_block_jacobi_apply_f32(pdelab_gpu_gridoperator* h, const float* z, float* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  sync_prec_cache(h);
  const BJVec<float, float> v{z, y, nullptr, nullptr, nullptr, nullptr, nullptr};
  return bj_launch<float, float>(h, false, BJ_ADD, v, nullptr);
}

int pdelab_gpu_block_sor_apply(pdelab_gpu_gridoperator* h, const double* d, double* v) {
  if (!h || !d || !v || d == v) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  sync_prec_cache(h);
  return sor_launch<double>(h, d, v, nullptr, true);
}

int pdelab_gpu_iterative_block_jacobi_apply(pdelab_gpu_gridoperator* h, const double* d, double* v,
                                            unsigned long long* iterations) {
  if (!h || !d || !v || d == v) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  sync_prec_cache(h);
  CHK(ibj_prepare(h, 64));
  HIPCHK(h, hipMemsetAsync(h->d_ibj_stat, 0, 16, h->stream));
  CHK(ibj_launch<double>(h, d, v, static_cast<int*>(h->d_ibj_stat),
                         reinterpret_cast<unsigned long long*>(static_cast<char*>(h->d_ibj_stat) + 8), nullptr, true));
  struct {
    int err, pad;
    unsigned long long its;
  } st;
  HIPCHK(h, hipMemcpyAsync(&st, h->d_ibj_stat, 16, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (iterations) *iterations = st.its;
  if (st.err != 0)
    return fail(h, PDELAB_ERR_SOLVER_ABORT,
                "iterative block Jacobi: breakdown in a block solve (code " + std::to_string(st.err) + ")");
  return PDELAB_OK;
}

int pdelab_gpu_set_block_solver_options(pdelab_gpu_gridoperator* h, double reduction, int maxiter, int precondition,
                                        double weight) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  if (!std::isfinite(reduction) || reduction <= 0.0 || maxiter < 1 || !std::isfinite(weight) || weight == 0.0)
    return fail(h, PDELAB_ERR_INVALID_ARG, "block solver options: reduction > 0, maxiter >= 1, weight finite and nonzero");
  h->ibj_opt = IBJOptions{reduction, maxiter, precondition != 0 ? 1 : 0, weight};
  return PDELAB_OK;
}

int pdelab_gpu_set_sor_relaxation(pdelab_gpu_gridoperator* h, double omega) {
  if (!h) return PDELAB_ERR_INVALID_ARG;
  if (!std::isfinite(omega) || omega <= 0.0 || omega >= 2.0)
    return fail(h, PDELAB_ERR_INVALID_ARG, "block SOR: omega must lie in (0, 2)");
  h->sor_omega = omega;
  return PDELAB_OK;
}

int pdelab_gpu_mass_apply(pdelab_gpu_gridoperator* h, const double* z, double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  const int nb = (int)((h->N + 255) / 256);
  switch (h->n1) {
    case 2: hipLaunchKernelGGL((dg_mass_kernel<2, double>), dim3(nb), dim3(256), 0, h->stream, z, y, h->a64); break;
    case 3: hipLaunchKernelGGL((dg_mass_kernel<3, double>), dim3(nb), dim3(256), 0, h->stream, z, y, h->a64); break;
    case 4: hipLaunchKernelGGL((dg_mass_kernel<4, double>), dim3(nb), dim3(256), 0, h->stream, z, y, h->a64); break;
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_dot(pdelab_gpu_gridoperator* h, const double* x, const double* y, double* out) {
  if (!h || !x || !y || !out) return PDELAB_ERR_INVALID_ARG;
  CHK((dot_dev<double, double>(h, x, y, h->N, 0)));
  return read_scalars<double>(h, out, 1);
}
int pdelab_gpu_norm(pdelab_gpu_gridoperator* h, const double* x, double* out) {
  CHK(pdelab_gpu_dot(h, x, x, out));
  *out = std::sqrt(*out);
  return PDELAB_OK;
}
int pdelab_gpu_axpy(pdelab_gpu_gridoperator* h, double a, const double* x, double* y) {
  if (!h || !x || !y) return PDELAB_ERR_INVALID_ARG;
  hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(h->N)), dim3(256), 0, h->stream, h->N, a, x, y);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_solve(pdelab_gpu_gridoperator* h, int method, int precond, double reduction, int maxit, double* z,
                     double* r, pdelab_linear_solver_result* res) {
  if (!h || !z || !r || !res) return PDELAB_ERR_INVALID_ARG;
  std::memset(res, 0, sizeof(*res));
  CHK(check_outflow(h));
  return krylov<double, double>(h, method, precond, reduction, maxit, false, z, r, 0, res);
}

int pdelab_gpu_solve_transpose(pdelab_gpu_gridoperator* h, int method, int precond, double reduction, int maxit,
                               double* z, double* r, pdelab_linear_solver_result* res) {
  if (!h || !z || !r || !res) return PDELAB_ERR_INVALID_ARG;
  std::memset(res, 0, sizeof(*res));
  CHK(check_outflow(h));
  return krylov<double, double>(h, method, precond, reduction, maxit, true, z, r, 0, res);
}

int pdelab_gpu_linear_problem_solve(pdelab_gpu_gridoperator* h, int method, int precond, double reduction,
                                    double min_defect, int maxit, double* x, pdelab_linear_problem_result* res) {
  if (!h || !x || !res) return PDELAB_ERR_INVALID_ARG;
  std::memset(res, 0, sizeof(*res));
  const long long n = h->N;
  CHK(ensure_ws<double>(h, 7));
  double* r = h->w64[5];
  double* z = h->w64[6];
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(h, hipMemsetAsync(r, 0, sizeof(double) * n, h->stream));
  CHK(pdelab_gpu_residual(h, x, r));  // residual is additive
  double defect = 0;
  CHK(pdelab_gpu_norm(h, r, &defect));
  res->assembler_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  HIPCHK(h, hipMemsetAsync(z, 0, sizeof(double) * n, h->stream));
  const double red = std::max(reduction, min_defect / defect);
  auto t1 = std::chrono::steady_clock::now();
  int st = krylov<double, double>(h, method, precond, red, maxit, false, z, r, 0, &res->solver);
  res->linear_solver_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  res->first_defect = defect;
  res->defect = defect * res->solver.reduction;
  res->linear_solver_iterations = (int)res->solver.iterations;
  if (st) return st;
  hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, -1.0, (const double*)z, x);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PDELAB_OK;
}

int pdelab_gpu_solve_mixed(pdelab_gpu_gridoperator* h, int method, int precond, double inner_reduction,
                           int inner_maxit, double outer_reduction, int outer_maxit, double* x,
                           pdelab_mixed_result* res) {
  if (!h || !x || !res) return PDELAB_ERR_INVALID_ARG;
  std::memset(res, 0, sizeof(*res));
  if (outer_maxit <= 0) outer_maxit = 50;
  if (inner_maxit <= 0) inner_maxit = 5000;
  const int variant = method & (PDELAB_MIXED_RESTARTED | PDELAB_MIXED_FLEXIBLE);
  method &= ~(PDELAB_MIXED_RESTARTED | PDELAB_MIXED_FLEXIBLE);
  if (method != PDELAB_SOLVER_CG && method != PDELAB_SOLVER_BICGSTAB)
    return fail(h, PDELAB_ERR_INVALID_ARG, "unknown Krylov method");
  if (variant == (PDELAB_MIXED_RESTARTED | PDELAB_MIXED_FLEXIBLE))
    return fail(h, PDELAB_ERR_INVALID_ARG, "PDELAB_MIXED_RESTARTED and PDELAB_MIXED_FLEXIBLE exclude each other");
  if (method == PDELAB_SOLVER_CG && precond == PDELAB_PREC_BLOCK_SOR)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "block SOR: BiCGSTAB only (the preconditioner is not symmetric)");
  if (method == PDELAB_SOLVER_CG && precond == PDELAB_PREC_ITERATIVE_BLOCK_JACOBI)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "iterative block Jacobi: BiCGSTAB only (the preconditioner is an inner iteration)");
  if (variant & PDELAB_MIXED_FLEXIBLE) {
    auto t0 = std::chrono::steady_clock::now();
    CHK(compute_R0(h));
    CHK(fgmres_mixed(h, method, precond, inner_reduction, inner_maxit, outer_reduction, outer_maxit, x, res));
    res->elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PDELAB_OK;
  }
  if (!(variant & PDELAB_MIXED_RESTARTED)) {
    auto t0 = std::chrono::steady_clock::now();
    CHK(compute_R0(h));
    if (method == PDELAB_SOLVER_CG)
      CHK(cg_mixed_reliable(h, precond, inner_reduction, inner_maxit, outer_reduction, outer_maxit, x, res));
    else
      CHK(bicgstab_mixed_reliable(h, precond, inner_reduction, inner_maxit, outer_reduction, outer_maxit, x, res));
    res->elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PDELAB_OK;
  }

  const long long n = h->N;
  CHK(ensure_ws<double>(h, 6));
  CHK(ensure_ws<float>(h, 2));
  double* r64 = h->w64[5];
  float* r32 = h->w32[0];
  float* z32 = h->w32[1];
  auto t0 = std::chrono::steady_clock::now();
  CHK(compute_R0(h));
  double def0 = 0, def = 0;
  int breakdowns = 0;
  for (int outer = 0;; ++outer) {
    // FP64 outer residual r = R(x) = J x + R(0)
    CHK(apply64(h, x, r64, false, false));
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0,
                       (const double*)h->d_R0, r64);
    CHK(pdelab_gpu_norm(h, r64, &def));
    if (outer == 0) def0 = def;
    res->outer_iterations = outer;
    if (!std::isfinite(def)) {
      res->defect = def;
      return fail(h, PDELAB_ERR_SOLVER_ABORT, "refinement: residual is infinite or NaN");
    }
    if (def <= outer_reduction * def0 || def < 1e-30) { res->converged = 1; break; }
    if (outer >= outer_maxit) break;
    // the FP32 solve sees the residual at unit norm (its dot products then stay far from
    // the FP32 underflow threshold however small the outer residual has become)
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 1.0 / def,
                       (const double*)r64, r32);
    HIPCHK(h, hipMemsetAsync(z32, 0, sizeof(float) * n, h->stream));
    // the inner solve reduces by inner_reduction, or by a quarter of the reduction
    // still missing when that is the weaker requirement (the last refinement step)
    const double need = def0 > 0 ? outer_reduction * def0 / def : 1.0;
    const double ired = std::min(0.5, std::max(inner_reduction, 0.25 * need));
    pdelab_linear_solver_result ires;
    const int st = krylov<float, float>(h, method, precond, ired, inner_maxit, false, z32, r32, 2, &ires);
    res->inner_iterations += (int)ires.iterations;
    if (st == PDELAB_ERR_SOLVER_ABORT) {
      // an FP32 breakdown ends this inner solve: its iterate so far is still a correction
      // (a non-finite one makes the next outer residual non-finite, which aborts)
      if (++breakdowns > 3) return fail(h, PDELAB_ERR_SOLVER_ABORT, "refinement: repeated FP32 BiCGSTAB breakdown");
      h->err.clear();
    } else {
      CHK(st);
    }
    hipLaunchKernelGGL(sub_scaled_f32_kernel, dim3(grid_for(n)), dim3(256), 0, h->stream, n, def, (const float*)z32, x);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  res->first_defect = def0;
  res->defect = def;
  res->reduction = def0 > 0 ? def / def0 : 0.0;
  res->elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PDELAB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Output: values at tensor sets of local points, and the L2 error integral.
static PointSet point_set(const H* h, int m, const long double* xi, const long double* w) {
  PointSet ps;
  std::memset(&ps, 0, sizeof(ps));
  ps.m = m;
  const int k = h->cfg.degree;
  for (int q = 0; q < m; ++q) {
    ps.xi[q] = (double)xi[q];
    ps.w[q] = w ? (double)w[q] : 0.0;
    for (int a = 0; a <= k; ++a) {  // QkDG Lagrange basis, nodes j/k
      long double r = 1;
      for (int j = 0; j <= k; ++j)
        if (j != a) r *= ((long double)k * xi[q] - j) / (long double)(a - j);
      ps.P[q][a] = (double)r;
    }
  }
  return ps;
}

static CellGeom cell_geom(const H* h) {
  CellGeom g;
  g.n[0] = h->cfg.nx; g.n[1] = h->cfg.ny; g.n[2] = h->cfg.nz;
  g.h[0] = h->cfg.lx / h->cfg.nx; g.h[1] = h->cfg.ly / h->cfg.ny; g.h[2] = h->cfg.lz / h->cfg.nz;
  g.ncells = h->ncells;
  return g;
}

template <int MODE>
static int launch_points(H* h, const double* x, const PointSet& ps, double* out) {
  const dim3 grid((unsigned)h->ncells), block(256);
  const CellGeom g = cell_geom(h);
  switch (h->n1) {
    case 2: hipLaunchKernelGGL((cell_points_kernel<2, MODE>), grid, block, 0, h->stream, x, ps, g, h->la, out); break;
    case 3: hipLaunchKernelGGL((cell_points_kernel<3, MODE>), grid, block, 0, h->stream, x, ps, g, h->la, out); break;
    default: hipLaunchKernelGGL((cell_points_kernel<4, MODE>), grid, block, 0, h->stream, x, ps, g, h->la, out); break;
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

// Checks of the L2 error integral's arguments, shared by both entry points.
static int l2_error_check(H* h, const double* x, int order, double* err) {
  if (!h || !x || !err || order < 0) return PDELAB_ERR_INVALID_ARG;
  if (order / 2 + 1 > kMaxPts) return fail(h, PDELAB_ERR_INVALID_ARG, "l2 error: quadrature order above 23");
  if (h->cfg.problem != PDELAB_PROBLEM_EXP && h->cfg.problem != PDELAB_PROBLEM_QUADRATIC)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "l2 error: the problem has no built-in exact solution");
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "l2 error: more than 2^31 - 1 cells");
  return PDELAB_OK;
}

// The L2 error integral into part[0 .. ncells] (per-cell partials, then their fixed-order
// sum in part[ncells]), copied to *err. Every entry of part is written before it is read.
static int l2_error_into(H* h, const double* x, int order, double* part, double* err) {
  const int m = order / 2 + 1;
  long double xi[kMaxPts], w[kMaxPts];
  gauss01(m, xi, w);
  CHK(launch_points<1>(h, x, point_set(h, m, xi, w), part));
  hipLaunchKernelGGL((final_reduce_kernel<double, 1>), dim3(1), dim3(kFinalThreads), 0, h->stream,
                     (const double*)part, (int)h->ncells, part + h->ncells, (const int*)nullptr);
  HIPCHK(h, hipGetLastError());
  double v = 0;
  if (hipMemcpyAsync(&v, part + h->ncells, sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return fail(h, PDELAB_ERR_HIP, "l2 error: copy of the result failed");
  *err = v;
  return PDELAB_OK;
}

// Scratch of pdelab_gpu_l2_error_squared_noalloc: the per-cell partials and their sum. Each
// device holds its own copy for the life of the process. Host code never passes the
// symbol itself to a kernel: it uses the device address that pdelab_gpu_create obtains
// once per device with hipGetSymbolAddress (g_scratch_ptr).
static __device__ __align__(256) double scratch[PDELAB_NOALLOC_MAX_CELLS + 1];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

// With h->device current: the device address of that device's scratch array, looked up
// once per device.
static int init_scratch(H* h) {
  if (h->device < 0 || h->device >= kMaxDevices)
    return fail(h, PDELAB_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[h->device]) return PDELAB_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(h, PDELAB_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[h->device] = (double*)p;
  return PDELAB_OK;
}

extern "C" {

int pdelab_gpu_l2_error_squared(pdelab_gpu_gridoperator* h, const double* x, int order, double* err) {
  CHK(l2_error_check(h, x, order, err));
  double* part = nullptr;
  HIPCHK(h, hipMallocAsync((void**)&part, sizeof(double) * (h->ncells + 1), h->stream));
  const int st = l2_error_into(h, x, order, part, err);
  (void)hipFreeAsync(part, h->stream);
  return st;
}

int pdelab_gpu_l2_error_squared_noalloc(pdelab_gpu_gridoperator* h, const double* x, int order, double* err) {
  CHK(l2_error_check(h, x, order, err));
  if (h->ncells > PDELAB_NOALLOC_MAX_CELLS)
    return fail(h, PDELAB_ERR_INVALID_ARG,
                "l2 error (no allocation): " + std::to_string(h->ncells) + " cells, more than the scratch maximum of " +
                    std::to_string((long long)PDELAB_NOALLOC_MAX_CELLS));
  double* part = h->device >= 0 && h->device < kMaxDevices ? g_scratch_ptr[h->device] : nullptr;
  if (!part) return fail(h, PDELAB_ERR_HIP, "l2 error (no allocation): scratch address of the device not set up");
  return l2_error_into(h, x, order, part, err);
}

int pdelab_gpu_sample(pdelab_gpu_gridoperator* h, const double* x, int intervals, double* values) {
  if (!h || !x || !values || intervals < 1) return PDELAB_ERR_INVALID_ARG;
  if (intervals + 1 > kMaxPts) return fail(h, PDELAB_ERR_INVALID_ARG, "sample: at most 11 intervals");
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "sample: more than 2^31 - 1 cells");
  long double xi[kMaxPts];
  for (int q = 0; q <= intervals; ++q) xi[q] = (long double)q / intervals;
  return launch_points<0>(h, x, point_set(h, intervals + 1, xi, nullptr), values);
}

// ---------------------------------------------------------------------------
// One-step methods. A stage operator a M + b J equals b (J + (a/b) M), and J + s M is J with
// the reaction coefficient c + s: the reaction term c u v is the DG mass on the same
// quadrature. The stage solves therefore run the apply kernels with a shifted reaction.
static void set_reaction_shift(H* h, double s) {
  if (s == h->shift) return;
  h->shift = s;
  const double c = h->cfg.c + s;
  h->a64.c = c;
  h->a64.vc = h->a64.V * c;
  h->a32.c = (float)c;
  h->a32.vc = (float)(h->a64.V * c);
}

// Coefficients a(r, i) = A[r-1][i], b(r, i) = B[r-1][i] of the schemes
// (R:dune/pdelab/instationary/onestepparameter.hh: ImplicitEulerParameter :193,
// OneStepThetaParameter :96-101 with theta = 1/2, Alexander2Parameter :449-461).
struct StepScheme {
  int s;
  double A[2][3], B[2][3];
};

static bool step_scheme(int scheme, StepScheme& p) {
  std::memset(&p, 0, sizeof(p));
  if (scheme == PDELAB_TS_IMPLICIT_EULER) {
    p.s = 1;
    p.A[0][0] = -1.0; p.A[0][1] = 1.0;
    p.B[0][0] = 0.0; p.B[0][1] = 1.0;
  } else if (scheme == PDELAB_TS_CRANK_NICOLSON) {
    p.s = 1;
    p.A[0][0] = -1.0; p.A[0][1] = 1.0;
    p.B[0][0] = 0.5; p.B[0][1] = 0.5;
  } else if (scheme == PDELAB_TS_ALEXANDER2) {
    const double alpha = 1.0 - 0.5 * std::sqrt(2.0);
    p.s = 2;
    p.A[0][0] = -1.0; p.A[0][1] = 1.0; p.A[0][2] = 0.0;
    p.A[1][0] = -1.0; p.A[1][1] = 0.0; p.A[1][2] = 1.0;
    p.B[0][0] = 0.0; p.B[0][1] = alpha; p.B[0][2] = 0.0;
    p.B[1][0] = 0.0; p.B[1][1] = 1.0 - alpha; p.B[1][2] = alpha;
  } else {
    return false;
  }
  return true;
}

static int ensure_os(H* h) {
  for (int i = 0; i < 4; ++i)
    if (!h->os[i]) HIPCHK(h, hipMalloc(&h->os[i], sizeof(double) * h->N));
  return PDELAB_OK;
}

// y += M z with the DG mass
static int mass_add(H* h, const double* z, double* y) { return pdelab_gpu_mass_apply(h, z, y); }

int pdelab_gpu_onestep_jacobian_apply(pdelab_gpu_gridoperator* h, double a_mass, double b_dt, const double* z,
                                      double* y) {
  if (!h || !z || !y) return PDELAB_ERR_INVALID_ARG;
  CHK(check_outflow(h));
  CHK(ensure_os(h));
  const long long n = h->N;
  double* t = h->os[3];
  if (b_dt != 0.0) {
    // t = (J + (a/b) M) z, y += b t
    set_reaction_shift(h, a_mass / b_dt);
    const int st = apply64(h, z, t, false, false);
    set_reaction_shift(h, 0.0);
    CHK(st);
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, b_dt, (const double*)t, y);
  } else if (a_mass != 0.0) {
    HIPCHK(h, hipMemsetAsync(t, 0, sizeof(double) * n, h->stream));
    CHK(mass_add(h, z, t));
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, a_mass, (const double*)t, y);
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_onestep_step(pdelab_gpu_gridoperator* h, int scheme, int method, int precond, double dt,
                            double reduction, int maxit, const double* xold, double* xnew,
                            pdelab_onestep_result* res) {
  if (!h || !xold || !xnew || !res || !(dt > 0)) return PDELAB_ERR_INVALID_ARG;
  std::memset(res, 0, sizeof(*res));
  StepScheme P;
  if (!step_scheme(scheme, P)) return fail(h, PDELAB_ERR_INVALID_ARG, "unknown one-step scheme");
  CHK(check_outflow(h));
  auto t0 = std::chrono::steady_clock::now();
  const long long n = h->N;
  set_reaction_shift(h, 0.0);
  CHK(compute_R0(h));
  CHK(ensure_os(h));
  CHK(ensure_ws<double>(h, 7));
  double *C = h->os[0], *r = h->os[1], *stage = h->os[2], *t = h->os[3], *z = h->w64[6];
  const dim3 gv(grid_for(n)), bv(256);
  // adds w R(u) = w (J u + R(0)) and w_m M u to y
  auto add_terms = [&](const double* u, double w, double wm, double* y) -> int {
    if (w != 0.0) {
      CHK(apply64(h, u, t, false, false));
      hipLaunchKernelGGL(axpy_kernel<double>, gv, bv, 0, h->stream, n, w, (const double*)t, y);
      hipLaunchKernelGGL(axpy_kernel<double>, gv, bv, 0, h->stream, n, w, (const double*)h->d_R0, y);
    }
    if (wm != 0.0) {
      HIPCHK(h, hipMemsetAsync(t, 0, sizeof(double) * n, h->stream));
      CHK(mass_add(h, u, t));
      hipLaunchKernelGGL(axpy_kernel<double>, gv, bv, 0, h->stream, n, wm, (const double*)t, y);
    }
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  res->converged = 1;
  res->stages = P.s;
  for (int rr = 1; rr <= P.s; ++rr) {
    const double* A = P.A[rr - 1];
    const double* B = P.B[rr - 1];
    // constant part of the stage residual from the earlier stages
    HIPCHK(h, hipMemsetAsync(C, 0, sizeof(double) * n, h->stream));
    for (int j = 0; j < rr; ++j) CHK(add_terms(j == 0 ? xold : stage, B[j] * dt, A[j], C));
    // the stage unknown and its initial guess (implicitonestep.hh:170-183)
    double* x = rr == P.s ? xnew : stage;
    if (rr > 1) HIPCHK(h, hipMemcpyAsync(x, stage, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    else if (x != xnew) HIPCHK(h, hipMemcpyAsync(x, xnew, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    const double bdt = B[rr] * dt, am = A[rr];
    if (!(bdt > 0.0)) return fail(h, PDELAB_ERR_UNSUPPORTED, "one-step: explicit stage");
    // StationaryLinearProblemSolver::apply on the stage operator am M + bdt J
    HIPCHK(h, hipMemcpyAsync(r, C, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    CHK(add_terms(x, bdt, am, r));
    double defect = 0;
    CHK(pdelab_gpu_norm(h, r, &defect));
    const double red = std::max(reduction, 1e-99 / defect);
    hipLaunchKernelGGL(scale_kernel<double>, gv, bv, 0, h->stream, n, 1.0 / bdt, r);
    HIPCHK(h, hipMemsetAsync(z, 0, sizeof(double) * n, h->stream));
    set_reaction_shift(h, am / bdt);
    pdelab_linear_solver_result ls;
    auto t1 = std::chrono::steady_clock::now();
    const int st = krylov<double, double>(h, method, precond, red, maxit, false, z, r, 0, &ls);
    res->linear_solver_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    set_reaction_shift(h, 0.0);
    res->linear_solver_iterations += (int)ls.iterations;
    if (!ls.converged) res->converged = 0;
    CHK(st);
    hipLaunchKernelGGL(axpy_kernel<double>, gv, bv, 0, h->stream, n, -1.0, (const double*)z, x);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  res->elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PDELAB_OK;
}

int pdelab_gpu_grad_logkappa(pdelab_gpu_gridoperator* h, const double* u, const double* lambda, double* g) {
  if (!h || !u || !lambda || !g) return PDELAB_ERR_INVALID_ARG;
  if (h->a64.aniso) return fail(h, PDELAB_ERR_UNSUPPORTED, "kappa gradient: isotropic A = kappa I only");
  if (!h->d_kappa) {
    std::vector<double> k(h->ncells, h->cfg.kappa0);
    CHK(upload_kappa(h, k.data()));
  }
  const bool won = h->cfg.weights == PDELAB_WEIGHTS_ON;
  const int nb = (int)((h->ncells + 127) / 128);
#define GRADL(N1_)                                                                                                    \
  if (won) hipLaunchKernelGGL((grad_logkappa_kernel<N1_, true>), dim3(nb), dim3(128), 0, h->stream, u, lambda, h->d_kappa, g, h->ga); \
  else hipLaunchKernelGGL((grad_logkappa_kernel<N1_, false>), dim3(nb), dim3(128), 0, h->stream, u, lambda, h->d_kappa, g, h->ga);
  switch (h->n1) {
    case 2: GRADL(2); break;
    case 3: GRADL(3); break;
    case 4: GRADL(4); break;
  }
#undef GRADL
  HIPCHK(h, hipGetLastError());
  // the Dirichlet data terms of the residual scale with kappa_c too (g != 0)
  if (h->cfg.problem != PDELAB_PROBLEM_ZERO) {
    for (int s = 0; s < 6; ++s) {
      if (h->a64.bc[s] != BC_DIRICHLET) continue;
      const int d = s / 2, t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      const long long nf = (long long)h->la.n[t1] * h->la.n[t2];
      const int gb = (int)((nf + 127) / 128);
      switch (h->n1) {
        case 2: hipLaunchKernelGGL(grad_logkappa_dirichlet_kernel<2>, dim3(gb), dim3(128), 0, h->stream, lambda, h->d_gtab, h->d_kappa, g, h->la, s); break;
        case 3: hipLaunchKernelGGL(grad_logkappa_dirichlet_kernel<3>, dim3(gb), dim3(128), 0, h->stream, lambda, h->d_gtab, h->d_kappa, g, h->la, s); break;
        case 4: hipLaunchKernelGGL(grad_logkappa_dirichlet_kernel<4>, dim3(gb), dim3(128), 0, h->stream, lambda, h->d_gtab, h->d_kappa, g, h->la, s); break;
      }
      HIPCHK(h, hipGetLastError());
    }
  }
  return PDELAB_OK;
}

int pdelab_gpu_grad_source(pdelab_gpu_gridoperator* h, const double* lambda, double* g) {
  if (!h || !lambda || !g) return PDELAB_ERR_INVALID_ARG;
  const Tables1D& t = h->tab;
  double w[4] = {0, 0, 0, 0};
  for (int i = 0; i < t.n1; ++i) w[i] = (double)t.wint[i];
  const int nb = (int)((h->ncells + 255) / 256);
  const double V = h->la.V;
  switch (h->n1) {
    case 2: hipLaunchKernelGGL(grad_source_kernel<2>, dim3(nb), dim3(256), 0, h->stream, lambda, g, h->ncells, V, nullptr, w[0], w[1], w[2], w[3]); break;
    case 3: hipLaunchKernelGGL(grad_source_kernel<3>, dim3(nb), dim3(256), 0, h->stream, lambda, g, h->ncells, V, nullptr, w[0], w[1], w[2], w[3]); break;
    case 4: hipLaunchKernelGGL(grad_source_kernel<4>, dim3(nb), dim3(256), 0, h->stream, lambda, g, h->ncells, V, nullptr, w[0], w[1], w[2], w[3]); break;
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int pdelab_gpu_time_apply(pdelab_gpu_gridoperator* h, int precision, const void* z, void* y, int warmup, int reps,
                          double* avg_ms) {
  if (!h || !z || !y || !avg_ms || reps < 1 || (precision != 64 && precision != 32)) return PDELAB_ERR_INVALID_ARG;
  auto one = [&]() -> int {
    if (precision == 64) return apply64(h, (const double*)z, (double*)y, false, false);
    return apply32(h, (const float*)z, (float*)y, false, false);
  };
  for (int i = 0; i < warmup; ++i) CHK(one());
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < reps; ++i) CHK(one());
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *avg_ms = (double)ms / reps;
  return PDELAB_OK;
}

int pdelab_gpu_time_stream(pdelab_gpu_gridoperator* h, const double* x, double* y, int warmup, int reps,
                           double* avg_ms) {
  if (!h || !x || !y || !avg_ms || reps < 1) return PDELAB_ERR_INVALID_ARG;
  const long long n = h->N;
  auto one = [&]() -> int {
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, n, 0.5, x, y);
    HIPCHK(h, hipGetLastError());
    return PDELAB_OK;
  };
  for (int i = 0; i < warmup; ++i) CHK(one());
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < reps; ++i) CHK(one());
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *avg_ms = (double)ms / reps;
  return PDELAB_OK;
}

}  // extern "C"
