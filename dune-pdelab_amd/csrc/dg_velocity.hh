// dg_velocity.hh -- the convection terms of ConvectionDiffusionDG for a velocity field b(x)
// given at the quadrature points.
//
// With b varying inside a cell the convection integrals no longer factor into constant
// 1D matrices, so they are evaluated at the Gauss points as the reference does
// (R:dune/pdelab/localoperator/convectiondiffusiondg.hh): the volume term
// -int u (b . grad psi_i) (:178-186), the upwind skeleton term
// (omega_s u_s + omega_n u_n) (b . n) psi (:425-446; omega_s = 1 where b . n >= 0) and the
// boundary terms (:797-826 outflow, :885-899 Dirichlet with g = 0 in the apply). The rule is
// the tensor Gauss rule of order 2k + intorderadd with M points per direction, and every
// contraction is sum-factorised through LDS:
//   u(q) = (B (x) B (x) B) z,  V_d(q) = -w_q |K| / h_d b_d(q) u(q),
//   r = (D^T (x) B^T (x) B^T) V_0 + (B^T (x) D^T (x) B^T) V_1 + (B^T (x) B^T (x) D^T) V_2
// with B[q][i] = phi_i(x_q) and D[q][i] = phi_i'(x_q) on [0, 1], and per face the traces
// interpolated to the M^2 face points with B (x) B and the upwinded flux returned with its
// transpose. The equidistant Lagrange nodes j/k include 0 and 1, so a trace is exactly the
// DoF layer at that end of the cell.
//
// The kernels add these terms to what the constant-coefficient apply computed with b = 0.
#ifndef PDELAB_DG_VELOCITY_HH
#define PDELAB_DG_VELOCITY_HH

#include <hip/hip_runtime.h>

#include "dg_kernels.hh"
#include "dg_tables.hh"

namespace pdelab_amd {

template <typename T>
struct VelArgs {
  int n[3];
  long long stride[3];
  long long ncells;
  T B[kMaxQ][kMaxN1];   // phi_i at the Gauss points
  T Dq[kMaxQ][kMaxN1];  // phi_i' at the Gauss points
  T w[kMaxQ];           // Gauss weights on [0, 1]
  T vs[3];              // -|K| / h_d
  T area[3];            // |K| / h_d, the area of a face normal to d
  int bc[6];
};

// Device layout of the tabulated field (filled by the library from the caller's arrays):
//   bv: per cell 3 M^3 values, component d of the M^3 volume points (x fastest) at d M^3;
//   bf: per cell 6 M^2 values, face f = 2 d + upper of the cell, the component b_d along
//       +e_d at its M^2 points (lower tangential direction fastest).
template <int N1>
__device__ inline int dof3(const int (&i)[3]) {
  return i[0] + N1 * (i[1] + N1 * i[2]);
}

// y_c += (convection terms of cell c)(z). One wave per cell; `stop` as in the apply kernels
// (a device-side Krylov loop that has stopped makes the launch a no-op).
template <int N1, int M, typename T>
__global__ __launch_bounds__(64) void conv_var_kernel(const T* __restrict__ z, T* __restrict__ y,
                                                      const T* __restrict__ bv, const T* __restrict__ bf,
                                                      const VelArgs<T> a, const int* __restrict__ stop) {
  constexpr int L = N1 * N1, NL = L * N1, M2 = M * M, M3 = M2 * M;
  __shared__ T zs[NL], r[NL];
  __shared__ T A1[M * L], A2[M2 * N1];
  __shared__ T V[3][M3];
  __shared__ T W[3][N1 * M2];
  __shared__ T S[2][L * M];
  __shared__ T F[2][M * N1], fv[M2];
  if (stop && *stop) return;
  const long long c = blockIdx.x;
  const int t = threadIdx.x;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  for (int o = t; o < NL; o += 64) zs[o] = z[c * NL + o];
  __syncthreads();
  // ---- volume: interpolate to the points, x then y then z
  for (int o = t; o < M * L; o += 64) {
    const int q0 = o % M, jk = o / M;
    T acc = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) acc += a.B[q0][e] * zs[e + N1 * jk];
    A1[o] = acc;
  }
  __syncthreads();
  for (int o = t; o < M2 * N1; o += 64) {
    const int q0 = o % M, q1 = (o / M) % M, k = o / M2;
    T acc = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) acc += a.B[q1][e] * A1[q0 + M * (e + N1 * k)];
    A2[o] = acc;
  }
  __syncthreads();
  const T* bc = bv + c * 3 * M3;
  for (int o = t; o < M3; o += 64) {
    const int q0 = o % M, q1 = (o / M) % M, q2 = o / M2;
    T u = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) u += a.B[q2][e] * A2[q0 + M * (q1 + M * e)];
    const T wu = a.w[q0] * a.w[q1] * a.w[q2] * u;
#pragma unroll
    for (int d = 0; d < 3; ++d) V[d][o] = a.vs[d] * bc[d * M3 + o] * wu;
  }
  __syncthreads();
  // ---- volume: test with the gradients, x then y then z
  for (int o = t; o < N1 * M2; o += 64) {
    const int i = o % N1, q12 = o / N1;
    T w0 = T(0), w1 = T(0), w2 = T(0);
#pragma unroll
    for (int q0 = 0; q0 < M; ++q0) {
      w0 += a.Dq[q0][i] * V[0][q0 + M * q12];
      w1 += a.B[q0][i] * V[1][q0 + M * q12];
      w2 += a.B[q0][i] * V[2][q0 + M * q12];
    }
    W[0][o] = w0;
    W[1][o] = w1;
    W[2][o] = w2;
  }
  __syncthreads();
  for (int o = t; o < L * M; o += 64) {
    const int i = o % N1, j = (o / N1) % N1, q2 = o / L;
    T s0 = T(0), s1 = T(0);
#pragma unroll
    for (int q1 = 0; q1 < M; ++q1) {
      const int wi = i + N1 * (q1 + M * q2);
      s0 += a.B[q1][j] * W[0][wi] + a.Dq[q1][j] * W[1][wi];
      s1 += a.B[q1][j] * W[2][wi];
    }
    S[0][o] = s0;
    S[1][o] = s1;
  }
  __syncthreads();
  for (int o = t; o < NL; o += 64) {
    const int ij = o % L, k = o / L;
    T acc = T(0);
#pragma unroll
    for (int q2 = 0; q2 < M; ++q2) acc += a.B[q2][k] * S[0][ij + L * q2] + a.Dq[q2][k] * S[1][ij + L * q2];
    r[o] = acc;
  }
  __syncthreads();
  // ---- faces (the same for every thread of the block: the branches below are uniform)
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, up = f % 2;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const bool interior = up ? coord[d] < a.n[d] - 1 : coord[d] > 0;
    const int bct = interior ? 0 : a.bc[f];
    if (!interior && bct != BC_DIRICHLET && bct != BC_OUTFLOW) continue;  // Neumann, None: no apply term
    const long long nb = interior ? c + (up ? a.stride[d] : -a.stride[d]) : c;
    const T* zn = z + nb * NL;
    // traces along t1 interpolated: F[s][q1 + M p2] = sum_p1 B[q1][p1] u_s(p1, p2)
    for (int o = t; o < M * N1; o += 64) {
      const int q1 = o % M, p2 = o / M;
      T as = T(0), an = T(0);
#pragma unroll
      for (int p1 = 0; p1 < N1; ++p1) {
        int idx[3];
        idx[d] = up ? N1 - 1 : 0;
        idx[t1] = p1;
        idx[t2] = p2;
        as += a.B[q1][p1] * zs[dof3<N1>(idx)];
        if (interior) {
          idx[d] = up ? 0 : N1 - 1;
          an += a.B[q1][p1] * zn[dof3<N1>(idx)];
        }
      }
      F[0][o] = as;
      F[1][o] = an;
    }
    __syncthreads();
    const T* bfc = bf + (c * 6 + f) * M2;
    const T ns = up ? T(1) : T(-1);
    for (int o = t; o < M2; o += 64) {
      const int q1 = o % M, q2 = o / M;
      T us = T(0), un = T(0);
#pragma unroll
      for (int p2 = 0; p2 < N1; ++p2) {
        us += a.B[q2][p2] * F[0][q1 + M * p2];
        un += a.B[q2][p2] * F[1][q1 + M * p2];
      }
      const T cf = a.w[q1] * a.w[q2] * a.area[d] * (ns * bfc[o]);  // (b . n) times the point's weight
      T v;
      if (bct == BC_OUTFLOW) v = cf * us;
      else v = cf >= T(0) ? cf * us : cf * un;  // a Dirichlet side has un = g = 0 in the apply
      fv[o] = v;
    }
    __syncthreads();
    for (int o = t; o < N1 * M; o += 64) {
      const int p1 = o % N1, q2 = o / N1;
      T acc = T(0);
#pragma unroll
      for (int q1 = 0; q1 < M; ++q1) acc += a.B[q1][p1] * fv[q1 + M * q2];
      F[0][o] = acc;
    }
    __syncthreads();
    for (int o = t; o < L; o += 64) {
      const int p1 = o % N1, p2 = o / N1;
      T acc = T(0);
#pragma unroll
      for (int q2 = 0; q2 < M; ++q2) acc += a.B[q2][p2] * F[0][p1 + N1 * q2];
      int idx[3];
      idx[d] = up ? N1 - 1 : 0;
      idx[t1] = p1;
      idx[t2] = p2;
      r[dof3<N1>(idx)] += acc;
    }
    __syncthreads();
  }
  for (int o = t; o < NL; o += 64) y[c * NL + o] += r[o];
}

// diag += the convection terms' diagonal entries (the coefficient of z_i in the row of
// DoF i); one thread per DoF.
template <int N1, int M>
__global__ __launch_bounds__(256) void conv_var_diag_kernel(double* __restrict__ diag, const double* __restrict__ bv,
                                                            const double* __restrict__ bf, const VelArgs<double> a) {
  constexpr int L = N1 * N1, NL = L * N1, M2 = M * M, M3 = M2 * M;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= a.ncells * NL) return;
  const long long c = gid / NL;
  const int li = (int)(gid % NL);
  const int ii[3] = {li % N1, (li / N1) % N1, li / L};
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  const double* bc = bv + c * 3 * M3;
  double acc = 0.0;
  for (int q2 = 0; q2 < M; ++q2)
    for (int q1 = 0; q1 < M; ++q1)
      for (int q0 = 0; q0 < M; ++q0) {
        const int q = q0 + M * (q1 + M * q2);
        const double b0 = a.B[q0][ii[0]], b1 = a.B[q1][ii[1]], b2 = a.B[q2][ii[2]];
        const double phi = b0 * b1 * b2;
        const double w = a.w[q0] * a.w[q1] * a.w[q2];
        const double g = a.vs[0] * bc[q] * a.Dq[q0][ii[0]] * b1 * b2 +
                         a.vs[1] * bc[M3 + q] * b0 * a.Dq[q1][ii[1]] * b2 +
                         a.vs[2] * bc[2 * M3 + q] * b0 * b1 * a.Dq[q2][ii[2]];
        acc += w * phi * g;
      }
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, up = f % 2;
    if (ii[d] != (up ? N1 - 1 : 0)) continue;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const bool interior = up ? coord[d] < a.n[d] - 1 : coord[d] > 0;
    const int bct = interior ? 0 : a.bc[f];
    if (!interior && bct != BC_DIRICHLET && bct != BC_OUTFLOW) continue;
    const double* bfc = bf + (c * 6 + f) * M2;
    const double ns = up ? 1.0 : -1.0;
    for (int q2 = 0; q2 < M; ++q2)
      for (int q1 = 0; q1 < M; ++q1) {
        const double psi = a.B[q1][ii[t1]] * a.B[q2][ii[t2]];
        const double cf = a.w[q1] * a.w[q2] * a.area[d] * (ns * bfc[q1 + M * q2]);
        if (bct == BC_OUTFLOW || cf >= 0.0) acc += cf * psi * psi;
      }
  }
  diag[gid] += acc;
}

}  // namespace pdelab_amd

#endif
