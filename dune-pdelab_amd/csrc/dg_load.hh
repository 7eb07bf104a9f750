// dg_load.hh -- load vector R(0) (source f and Dirichlet data g) and the adjoint
// parameter-gradient kernels of the tensor-product DG operator. FP64 only.
#ifndef PDELAB_DG_LOAD_HH
#define PDELAB_DG_LOAD_HH

#include <hip/hip_runtime.h>

#include "dg_kernels.hh"
#include "dg_tables.hh"

namespace pdelab_amd {

enum { PROB_ZERO = 0, PROB_EXP = 1, PROB_CONSTANT = 2, PROB_GAUSSIANS = 3, PROB_QUADRATIC = 4, PROB_TABULATED = 5 };

struct LoadArgs {
  int problem;
  double kappa0, b[3], c, f0, g0, quad[7], j0, o0;
  int n_gauss;
  double gauss[16][5];
  int n[3];
  long long ncells;
  double h[3], V;
  int m;
  double qx[kMaxQ], qw[kMaxQ], P[kMaxQ][kMaxN1];
  double e0[kMaxN1], e1[kMaxN1], d0[kMaxN1], d1[kMaxN1];
  double theta, alphak;
  int weightsOn;
  int bc[6];
  long long bface_offset[7];
  int aniso;  // kappa holds a_x, a_y, a_z per cell (kappa_index)
};

__device__ inline double exact_u(const LoadArgs& a, const double x[3]) {
  if (a.problem == PROB_EXP) {
    double r2 = 0;
    for (int d = 0; d < 3; ++d) r2 += (x[d] - 0.5) * (x[d] - 0.5);
    return exp(-r2);
  }
  if (a.problem == PROB_QUADRATIC) {
    const double* q = a.quad;
    return q[0] + q[1] * x[0] + q[2] * x[1] + q[3] * x[2] + q[4] * x[0] * x[0] + q[5] * x[1] * x[1] + q[6] * x[2] * x[2];
  }
  return 0.0;
}

// Source term f of the built-in problems (ConvectionDiffusionModelProblem::f,
// R:dune/pdelab/localoperator/convectiondiffusionparameter.hh:174-179); b is the velocity at x
__device__ inline double eval_f(const LoadArgs& a, const double x[3], const double b[3]) {
  switch (a.problem) {
    case PROB_EXP: {
      double r2 = 0, bg = 0;
      for (int d = 0; d < 3; ++d) r2 += (x[d] - 0.5) * (x[d] - 0.5);
      const double u = exp(-r2);
      for (int d = 0; d < 3; ++d) bg += b[d] * (-2.0 * (x[d] - 0.5) * u);
      return a.kappa0 * (6.0 - 4.0 * r2) * u + bg + a.c * u;
    }
    case PROB_CONSTANT: return a.f0;
    case PROB_GAUSSIANS: {
      double v = 0;
      for (int i = 0; i < a.n_gauss; ++i) {
        double r2 = 0;
        for (int d = 0; d < 3; ++d) r2 += (x[d] - a.gauss[i][d]) * (x[d] - a.gauss[i][d]);
        v += a.gauss[i][4] * exp(-r2 / (2.0 * a.gauss[i][3] * a.gauss[i][3]));
      }
      return v;
    }
    case PROB_QUADRATIC: {
      const double* q = a.quad;
      const double lap = 2.0 * (q[4] + q[5] + q[6]);
      const double grad[3] = {q[1] + 2 * q[4] * x[0], q[2] + 2 * q[5] * x[1], q[3] + 2 * q[6] * x[2]};
      return -a.kappa0 * lap + b[0] * grad[0] + b[1] * grad[1] + b[2] * grad[2] + a.c * exact_u(a, x);
    }
    default: return 0.0;
  }
}

// Neumann flux j ((b u - A grad u).n = j) and outflow data o (-(A grad u).n = o) of the
// built-in problems at x on a boundary side with outer normal nsign e_d
// (convectiondiffusionparameter.hh:121-124, 196-208): the fluxes of the exact solution
// (A = kappa0 I) for EXP and QUADRATIC, the constants j0 / o0 otherwise. bn = b . n at x.
__device__ inline double eval_flux(const LoadArgs& a, const double x[3], int d, double nsign, double bn, bool neumann) {
  if (a.problem == PROB_EXP || a.problem == PROB_QUADRATIC) {
    double gd;
    if (a.problem == PROB_EXP) gd = -2.0 * (x[d] - 0.5) * exact_u(a, x);
    else gd = a.quad[1 + d] + 2.0 * a.quad[4 + d] * x[d];
    const double diff = -a.kappa0 * nsign * gd;
    return neumann ? bn * exact_u(a, x) + diff : diff;
  }
  return neumann ? a.j0 : a.o0;
}

__device__ inline double eval_g(const LoadArgs& a, const double x[3]) {
  switch (a.problem) {
    case PROB_EXP:
    case PROB_QUADRATIC: return exact_u(a, x);
    case PROB_CONSTANT:
    case PROB_GAUSSIANS: return a.g0;
    default: return 0.0;
  }
}

// R0 += lambda_volume contribution -int f phi_i (convectiondiffusiondg.hh:1048-1075),
// sum-factorised over the m^3 Gauss points. One thread per DoF. bvt: the tabulated velocity
// field (per cell, component d of the m^3 points at d m^3), or null for the constant b.
template <int N1>
__global__ __launch_bounds__(256) void load_volume_kernel(double* __restrict__ R0, const double* __restrict__ ftab,
                                                          const double* __restrict__ src, const LoadArgs a,
                                                          const double* __restrict__ bvt) {
  constexpr int L = N1 * N1, NL = L * N1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= a.ncells * NL) return;
  const long long c = gid / NL;
  const int li = (int)(gid % NL);
  const int i0 = li % N1, i1 = (li / N1) % N1, i2 = li / L;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  double lo[3], ext[3];
  for (int d = 0; d < 3; ++d) { lo[d] = coord[d] * a.h[d]; ext[d] = (coord[d] + 1) * a.h[d] - lo[d]; }
  const int m = a.m;
  const double s = src ? src[c] : 0.0;
  double acc = 0.0;
  for (int q2 = 0; q2 < m; ++q2)
    for (int q1 = 0; q1 < m; ++q1) {
      const double w12 = a.qw[q1] * a.qw[q2] * a.P[q1][i1] * a.P[q2][i2];
      for (int q0 = 0; q0 < m; ++q0) {
        double f;
        if (a.problem == PROB_TABULATED) {
          f = ftab ? ftab[c * (long long)(m * m * m) + q0 + m * (q1 + m * q2)] : 0.0;
        } else {
          const double x[3] = {lo[0] + ext[0] * a.qx[q0], lo[1] + ext[1] * a.qx[q1], lo[2] + ext[2] * a.qx[q2]};
          if (bvt) {
            const long long m3 = (long long)m * m * m, q = q0 + m * (q1 + m * q2);
            const double* bc = bvt + c * 3 * m3 + q;
            const double bq[3] = {bc[0], bc[m3], bc[2 * m3]};
            f = eval_f(a, x, bq);
          } else {
            f = eval_f(a, x, a.b);
          }
        }
        f += s;
        acc += f * a.qw[q0] * a.P[q0][i0] * w12;
      }
    }
  R0[gid] += -(ext[0] * ext[1] * ext[2]) * acc;
}

// R0 += data terms of residual_boundary_integral (u = 0): on Dirichlet sides
// (convectiondiffusiondg.hh:839-877) inflow convection g (b.n) psi, -theta g (A n . grad psi),
// -penalty g psi; on Neumann sides (:776-783) j psi; on outflow sides (:810-815) o psi.
// One launch per boundary side (no two threads of a launch write the same entry); one
// thread per (boundary cell, DoF). bct is the side's boundary type. bft: the tabulated
// velocity field (per cell and face f = 2 d + upper, b_d at the m^2 points), or null for the
// constant b; with it the upwind choice of the inflow term is made per point.
template <int N1>
__global__ __launch_bounds__(256) void load_boundary_kernel(double* __restrict__ R0, const double* __restrict__ gtab,
                                                            const double* __restrict__ kappa, const LoadArgs a,
                                                            int side, int bct, const double* __restrict__ bft) {
  constexpr int L = N1 * N1, NL = L * N1;
  const int d = side / 2, upper = side % 2;
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  const long long nfaces = (long long)a.n[t1] * a.n[t2];
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nfaces * NL) return;
  const long long fidx = gid / NL;
  const int li = (int)(gid % NL);
  const int ii[3] = {li % N1, (li / N1) % N1, li / L};
  int coord[3];
  coord[d] = upper ? a.n[d] - 1 : 0;
  coord[t1] = (int)(fidx % a.n[t1]);
  coord[t2] = (int)(fidx / a.n[t1]);
  const long long c = coord[0] + (long long)a.n[0] * (coord[1] + (long long)a.n[1] * coord[2]);
  double lo[3], ext[3];
  for (int e = 0; e < 3; ++e) { lo[e] = coord[e] * a.h[e]; ext[e] = (coord[e] + 1) * a.h[e] - lo[e]; }
  const int m = a.m;
  const double nsign = upper ? 1.0 : -1.0;
  const double* bfc = bft ? bft + (c * 6 + side) * (long long)(m * m) : nullptr;
  double gt = 0.0, gconv = 0.0;
  for (int q2 = 0; q2 < m; ++q2)
    for (int q1 = 0; q1 < m; ++q1) {
      const double bn = bfc ? nsign * bfc[q1 + m * q2] : nsign * a.b[d];
      double g;
      if (a.problem == PROB_TABULATED) {
        g = gtab ? gtab[(a.bface_offset[side] + fidx) * (m * m) + q1 + m * q2] : 0.0;
      } else {
        double x[3];
        x[d] = lo[d] + ext[d] * (upper ? 1.0 : 0.0);
        x[t1] = lo[t1] + ext[t1] * a.qx[q1];
        x[t2] = lo[t2] + ext[t2] * a.qx[q2];
        g = bct == BC_DIRICHLET ? eval_g(a, x) : eval_flux(a, x, d, nsign, bn, bct == BC_NEUMANN);
      }
      const double gw = g * a.qw[q1] * a.qw[q2] * a.P[q1][ii[t1]] * a.P[q2][ii[t2]];
      gt += gw;
      if (bn < 0.0) gconv += bn * gw;  // inflow: omega_n = 1 picks g (:861-868)
    }
  const double vol = ext[0] * ext[1] * ext[2];
  const double area = vol / ext[d];
  const int i = ii[d];
  const double es = upper ? a.e1[i] : a.e0[i];
  if (bct != BC_DIRICHLET) {  // Neumann flux or outflow data: + int data psi
    R0[c * NL + li] += area * gt * es;
    return;
  }
  const double hF = vol / area;
  const double kc = kappa ? kappa[kappa_index(a.aniso, c, d)] : a.kappa0;
  const double nf = nsign * a.b[d];
  const double ds = upper ? a.d1[i] : a.d0[i];
  const double Pb = (a.alphak / hF) * (a.weightsOn ? kc : 1.0);
  double coef = 0.0;
  if (!bfc && nf < 0.0) coef += nf * es;
  coef += -a.theta * kc * nsign * ds / ext[d];
  coef += -Pb * es;
  if (bfc) R0[c * NL + li] += area * (gt * coef + gconv * es);
  else R0[c * NL + li] += area * gt * coef;
}

// Tangential-mass contraction sum_{t,t'} (M (x) M)[t][t'] A[t] B[t'] of two face arrays.
template <int N1>
__device__ inline double face_mass_dot(const double (&M)[kMaxN1][kMaxN1], const double* A, const double* B) {
  double s = 0.0;
  for (int i2 = 0; i2 < N1; ++i2)
    for (int i1 = 0; i1 < N1; ++i1) {
      double r = 0.0;
      for (int j2 = 0; j2 < N1; ++j2)
        for (int j1 = 0; j1 < N1; ++j1) r += M[i1][j1] * M[i2][j2] * B[j1 + N1 * j2];
      s += A[i1 + N1 * i2] * r;
    }
  return s;
}

struct GradArgs {
  int n[3];
  long long stride[3], ncells;
  double S[3];
  double M[kMaxN1][kMaxN1], K[kMaxN1][kMaxN1];
  double e0[kMaxN1], e1[kMaxN1], d0[kMaxN1], d1[kMaxN1];
  double theta, alphak;
  int bc[6];
};

// g[c] = -lambda^T (dJ/dlog kappa_c) u for SIPG/NIPG/IIPG diffusion. All kappa dependence
// is through kappa_c (volume, boundary faces) and through the face weights:
// * weightsOn: w = kappa_lo*kappa_hi/S and H = 2w on interior faces
//   (convectiondiffusiondg.hh:326-333, 346): d w / d log kappa_c = D and
//   d(penalty)/d log kappa_c = 2 alpha k(k+2) D with D = kappa_c kappa_n (kappa_n + 1e-20)/S^2;
//   the boundary penalty is proportional to kappa_c (:717-734);
// * weightsOff: the weight of each side's normal derivative is kappa/2 of that side, the
//   penalties do not depend on kappa.
// One thread per cell; every interior face is evaluated by both of its cells.
template <int N1, bool WON>
__global__ __launch_bounds__(128) void grad_logkappa_kernel(const double* __restrict__ u, const double* __restrict__ lam,
                                                            const double* __restrict__ kappa, double* __restrict__ g,
                                                            const GradArgs a) {
  constexpr int L = N1 * N1, NL = L * N1;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  const double kc = kappa[c];
  const double* uc = u + c * NL;
  const double* lc = lam + c * NL;
  double acc = 0.0;
  // volume: kappa_c * sum_d S_d lambda^T (M (x) M (x) K_d) u
  double vol = 0.0;
  for (int i = 0; i < NL; ++i) {
    const int ii[3] = {i % N1, (i / N1) % N1, i / L};
    double row = 0.0;
    for (int j = 0; j < NL; ++j) {
      const int jj[3] = {j % N1, (j / N1) % N1, j / L};
      const double m0 = a.M[ii[0]][jj[0]], m1 = a.M[ii[1]][jj[1]], m2 = a.M[ii[2]][jj[2]];
      const double k0 = a.K[ii[0]][jj[0]], k1 = a.K[ii[1]][jj[1]], k2 = a.K[ii[2]][jj[2]];
      row += (a.S[0] * k0 * m1 * m2 + a.S[1] * m0 * k1 * m2 + a.S[2] * m0 * m1 * k2) * uc[j];
    }
    vol += lc[i] * row;
  }
  acc += kc * vol;
  double Ju[L], Gu[L], Jl[L], Gl[L];
  for (int D = 0; D < 3; ++D) {
    const double Sd = a.S[D];
    for (int side = 0; side < 2; ++side) {
      const bool upper = side == 1;
      const bool interior = upper ? coord[D] < a.n[D] - 1 : coord[D] > 0;
      if (interior) {
        const long long cn = upper ? c + a.stride[D] : c - a.stride[D];
        const double kn = kappa[cn];
        const double* ulo = upper ? uc : u + cn * NL;
        const double* uhi = upper ? u + cn * NL : uc;
        const double* llo = upper ? lc : lam + cn * NL;
        const double* lhi = upper ? lam + cn * NL : lc;
        for (int t = 0; t < L; ++t) {
          const int p = t % N1, q = t / N1;
          double ju = 0, gu = 0, jl = 0, gl = 0;
          for (int e = 0; e < N1; ++e) {
            const int ix = D == 0 ? e + N1 * p + L * q : (D == 1 ? p + N1 * e + L * q : p + N1 * q + L * e);
            ju += a.e1[e] * ulo[ix] - a.e0[e] * uhi[ix];
            jl += a.e1[e] * llo[ix] - a.e0[e] * lhi[ix];
            // weightsOn: both sides' normal derivatives carry w; weightsOff: only this
            // cell's (the lower cell's d1 part or the upper cell's d0 part) depends on kappa_c
            if (WON || upper) { gu += a.d1[e] * ulo[ix]; gl += a.d1[e] * llo[ix]; }
            if (WON || !upper) { gu += a.d0[e] * uhi[ix]; gl += a.d0[e] * lhi[ix]; }
          }
          Ju[t] = ju; Gu[t] = gu; Jl[t] = jl; Gl[t] = gl;
        }
        if (WON) {
          const double S = kc + kn + 1e-20;
          const double Dw = kc * kn * (kn + 1e-20) / (S * S);
          // face form with w = 1 and penalty 2*alpha*k(k+2): -G(u)J(l) + theta J(u)G(l) + P J(u)J(l)
          const double f = -face_mass_dot<N1>(a.M, Jl, Gu) + a.theta * face_mass_dot<N1>(a.M, Gl, Ju) +
                           2.0 * a.alphak * face_mass_dot<N1>(a.M, Jl, Ju);
          acc += Dw * Sd * f;
        } else {
          const double f = -face_mass_dot<N1>(a.M, Jl, Gu) + a.theta * face_mass_dot<N1>(a.M, Gl, Ju);
          acc += 0.5 * kc * Sd * f;
        }
      } else if (a.bc[2 * D + side] == BC_DIRICHLET) {
        for (int t = 0; t < L; ++t) {
          const int p = t % N1, q = t / N1;
          double vu = 0, du = 0, vl = 0, dl = 0;
          for (int e = 0; e < N1; ++e) {
            const int ix = D == 0 ? e + N1 * p + L * q : (D == 1 ? p + N1 * e + L * q : p + N1 * q + L * e);
            const double ev = upper ? a.e1[e] : a.e0[e], dv = upper ? a.d1[e] : a.d0[e];
            vu += ev * uc[ix]; du += dv * uc[ix]; vl += ev * lc[ix]; dl += dv * lc[ix];
          }
          Ju[t] = vu; Gu[t] = du; Jl[t] = vl; Gl[t] = dl;
        }
        // lower: (l0)(du0) + P l0 u0 - theta (dl0)(u0); upper: -(l1)(du1) + P l1 u1 + theta (dl1)(u1);
        // the penalty P is proportional to kappa_c with weightsOn only
        const double sg = upper ? -1.0 : 1.0;
        double f = sg * face_mass_dot<N1>(a.M, Jl, Gu) - sg * a.theta * face_mass_dot<N1>(a.M, Gl, Ju);
        if (WON) f += a.alphak * face_mass_dot<N1>(a.M, Jl, Ju);
        acc += kc * Sd * f;
      }
    }
  }
  g[c] = -acc;
}

// g[c] -= lambda_c^T dR0_c / d log kappa_c for the Dirichlet data terms of one boundary side
// (load_boundary_kernel): -theta g kappa_c (n . grad psi) and, with weightsOn, the penalty
// term -(alpha k(k+2) / h_F) kappa_c g psi. One launch per Dirichlet side, one thread per
// boundary cell of the side (a cell on several sides is updated by successive launches).
template <int N1>
__global__ __launch_bounds__(128) void grad_logkappa_dirichlet_kernel(const double* __restrict__ lam,
                                                                      const double* __restrict__ gtab,
                                                                      const double* __restrict__ kappa,
                                                                      double* __restrict__ g, const LoadArgs a,
                                                                      int side) {
  constexpr int L = N1 * N1, NL = L * N1;
  const int d = side / 2, upper = side % 2;
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  const long long nfaces = (long long)a.n[t1] * a.n[t2];
  const long long fidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (fidx >= nfaces) return;
  int coord[3];
  coord[d] = upper ? a.n[d] - 1 : 0;
  coord[t1] = (int)(fidx % a.n[t1]);
  coord[t2] = (int)(fidx / a.n[t1]);
  const long long c = coord[0] + (long long)a.n[0] * (coord[1] + (long long)a.n[1] * coord[2]);
  double lo[3], ext[3];
  for (int e = 0; e < 3; ++e) { lo[e] = coord[e] * a.h[e]; ext[e] = (coord[e] + 1) * a.h[e] - lo[e]; }
  const int m = a.m;
  const double vol = ext[0] * ext[1] * ext[2];
  const double area = vol / ext[d];
  const double hF = vol / area;
  const double kc = kappa[c];
  const double nsign = upper ? 1.0 : -1.0;
  double acc = 0.0;
  for (int li = 0; li < NL; ++li) {
    const int ii[3] = {li % N1, (li / N1) % N1, li / L};
    double gt = 0.0;
    for (int q2 = 0; q2 < m; ++q2)
      for (int q1 = 0; q1 < m; ++q1) {
        double gv;
        if (a.problem == PROB_TABULATED) {
          gv = gtab ? gtab[(a.bface_offset[side] + fidx) * (m * m) + q1 + m * q2] : 0.0;
        } else {
          double x[3];
          x[d] = lo[d] + ext[d] * (upper ? 1.0 : 0.0);
          x[t1] = lo[t1] + ext[t1] * a.qx[q1];
          x[t2] = lo[t2] + ext[t2] * a.qx[q2];
          gv = eval_g(a, x);
        }
        gt += gv * a.qw[q1] * a.qw[q2] * a.P[q1][ii[t1]] * a.P[q2][ii[t2]];
      }
    const int i = ii[d];
    const double es = upper ? a.e1[i] : a.e0[i];
    const double ds = upper ? a.d1[i] : a.d0[i];
    double coef = -a.theta * kc * nsign * ds / ext[d];
    if (a.weightsOn) coef += -(a.alphak / hF) * kc * es;
    acc += lam[c * NL + li] * area * gt * coef;
  }
  g[c] -= acc;
}

// g[c] = sum_i lambda_{c,i} int_c phi_i (derivative of -R w.r.t. a per-cell source)
template <int N1>
__global__ __launch_bounds__(256) void grad_source_kernel(const double* __restrict__ lam, double* __restrict__ g,
                                                          long long ncells, double V, const double* wint_dev0,
                                                          double w0, double w1, double w2, double w3) {
  constexpr int L = N1 * N1, NL = L * N1;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double w[4] = {w0, w1, w2, w3};
  double acc = 0.0;
  for (int i = 0; i < NL; ++i) acc += lam[c * NL + i] * w[i % N1] * w[(i / N1) % N1] * w[i / L];
  g[c] = V * acc;
}

}  // namespace pdelab_amd

#endif
