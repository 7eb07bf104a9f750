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
  int m;                // Gauss points per direction
};

// Device layout of the tabulated field (filled by the library from the caller's arrays):
//   bv: per cell 3 M^3 values, component d of the M^3 volume points (x fastest) at d M^3;
//   bf: per cell 6 M^2 values, face f = 2 d + upper of the cell, the component b_d along
//       +e_d at its M^2 points (lower tangential direction fastest).
template <int N1>
__device__ inline int dof3(const int (&i)[3]) {
  return i[0] + N1 * (i[1] + N1 * i[2]);
}

// Index of the DoF of face f's layer at tangential position (p1, p2).
template <int N1>
__device__ inline int face_dof(int f, int p1, int p2) {
  const int d = f / 2;
  int idx[3];
  idx[d] = (f % 2) ? N1 - 1 : 0;
  idx[d == 0 ? 1 : 0] = p1;
  idx[d == 2 ? 1 : 2] = p2;
  return dof3<N1>(idx);
}

// y_c += (convection terms of cell c)(z), or with TRANS their transpose. One wave per cell.
// Every global load (the cell's values, the six neighbour layers facing it, the field at the
// volume and face points) is issued before the first barrier; the six faces are evaluated
// alongside the volume contractions, so the cell takes six barriers. `stop` as in the apply
// kernels (a device-side Krylov loop that has stopped makes the launch a no-op).
//
// The transpose: the volume term -int phi_j b . grad z, i.e. the gradient of z interpolated to
// the points (D in direction d, B elsewhere), P = -w |K| sum_d b_d (dz/dx_d), tested with
// B (x) B (x) B; a face, whose forward entries are [cf >= 0] cf psi_s psi_s (own column) and
// [cf > 0] (-cf) psi_n psi_s (the neighbour's row), gives [cf >= 0] cf (z_s - z_n) at each
// point (z_n = 0 on a Dirichlet side; an outflow side always takes cf z_s).
template <int N1, int M, typename T, bool TRANS = false>
__global__ __launch_bounds__(64) void conv_var_kernel(const T* __restrict__ z, T* __restrict__ y,
                                                      const T* __restrict__ bv, const T* __restrict__ bf,
                                                      const VelArgs<T> a, const int* __restrict__ stop) {
  constexpr int L = N1 * N1, NL = L * N1, M2 = M * M, M3 = M2 * M;
  __shared__ T zs[NL];
  __shared__ T zt[6][L];        // neighbour layers (zero without a neighbour)
  __shared__ T V[3][M3];        // the field at the volume points, then V_d
  __shared__ T bq[6][M2];       // b_d at the face points
  __shared__ T A1[TRANS ? 2 : 1][M * L], A2[TRANS ? 3 : 1][M2 * N1];
  __shared__ T W[3][N1 * M2];
  __shared__ T S[2][L * M];
  __shared__ T F[6][2][M * N1];  // traces interpolated along t1 (own, neighbour)
  __shared__ T fv[6][M2];        // upwinded flux at the face points
  __shared__ T G[6][N1 * M];     // flux tested along t1
  __shared__ T rf[6][L];         // face contributions to the layer DoFs
  __shared__ T tB[M][N1], tD[M][N1], tw[M];  // the 1D tables (lane-indexed reads from LDS)
  if (stop && *stop) return;
  const long long c = blockIdx.x;
  const int t = threadIdx.x;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  // face f takes part when it has a neighbour or a Dirichlet / outflow side (Neumann and
  // None sides have no term in the apply); the same for every thread
  unsigned interior = 0, active = 0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2;
    const bool in = (f % 2) ? coord[d] < a.n[d] - 1 : coord[d] > 0;
    if (in) interior |= 1u << f;
    if (in || a.bc[f] == BC_DIRICHLET || a.bc[f] == BC_OUTFLOW) active |= 1u << f;
  }
  // ---- all global loads
  for (int o = t; o < M * N1; o += 64) {
    tB[o / N1][o % N1] = a.B[o / N1][o % N1];
    tD[o / N1][o % N1] = a.Dq[o / N1][o % N1];
  }
  if (t < M) tw[t] = a.w[t];
  for (int o = t; o < NL; o += 64) zs[o] = z[c * NL + o];
  for (int o = t; o < 6 * L; o += 64) {
    const int f = o / L, pq = o % L;
    T v = T(0);
    if (interior & (1u << f)) {
      const int d = f / 2;
      const long long nb = c + ((f % 2) ? a.stride[d] : -a.stride[d]);
      v = z[nb * NL + face_dof<N1>(f ^ 1, pq % N1, pq / N1)];
    }
    zt[f][pq] = v;
  }
  const T* bc = bv + c * 3 * M3;
  for (int o = t; o < 3 * M3; o += 64) V[o / M3][o % M3] = bc[o];
  const T* bfc = bf + c * 6 * M2;
  for (int o = t; o < 6 * M2; o += 64) bq[o / M2][o % M2] = bfc[o];
  __syncthreads();
  // ---- phase 1: volume interpolation along x; face traces along t1
  for (int o = t; o < M * L; o += 64) {
    const int q0 = o % M, jk = o / M;
    T acc = T(0), accd = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      acc += tB[q0][e] * zs[e + N1 * jk];
      if (TRANS) accd += tD[q0][e] * zs[e + N1 * jk];
    }
    A1[0][o] = acc;
    if (TRANS) A1[1][o] = accd;
  }
  for (int o = t; o < 6 * M * N1; o += 64) {
    const int f = o / (M * N1), r = o % (M * N1), q1 = r % M, p2 = r / M;
    T as = T(0), an = T(0);
#pragma unroll
    for (int p1 = 0; p1 < N1; ++p1) {
      as += tB[q1][p1] * zs[face_dof<N1>(f, p1, p2)];
      an += tB[q1][p1] * zt[f][p1 + N1 * p2];
    }
    F[f][0][r] = as;
    F[f][1][r] = an;
  }
  __syncthreads();
  // ---- phase 2: along y; the upwinded flux at the face points
  for (int o = t; o < M2 * N1; o += 64) {
    const int q0 = o % M, q1 = (o / M) % M, k = o / M2;
    T acc = T(0), g0 = T(0), g1 = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int ai = q0 + M * (e + N1 * k);
      acc += tB[q1][e] * A1[0][ai];
      if (TRANS) {
        g0 += tB[q1][e] * A1[1][ai];  // d/dx
        g1 += tD[q1][e] * A1[0][ai];  // d/dy
      }
    }
    A2[0][o] = acc;
    if (TRANS) {
      A2[1][o] = g0;
      A2[2][o] = g1;
    }
  }
  for (int o = t; o < 6 * M2; o += 64) {
    const int f = o / M2, q = o % M2, q1 = q % M, q2 = q / M;
    T v = T(0);
    if (active & (1u << f)) {
      T us = T(0), un = T(0);
#pragma unroll
      for (int p2 = 0; p2 < N1; ++p2) {
        us += tB[q2][p2] * F[f][0][q1 + M * p2];
        un += tB[q2][p2] * F[f][1][q1 + M * p2];
      }
      // (b . n) times the point's weight; a Dirichlet side has un = g = 0 in the apply, an
      // outflow side takes the inside trace whatever the sign
      const T cf = tw[q1] * tw[q2] * a.area[f / 2] * ((f % 2) ? bq[f][q] : -bq[f][q]);
      const bool outflow = !(interior & (1u << f)) && a.bc[f] == BC_OUTFLOW;
      if (TRANS) v = outflow ? cf * us : (cf >= T(0) ? cf * (us - un) : T(0));
      else v = (outflow || cf >= T(0)) ? cf * us : cf * un;
    }
    fv[f][q] = v;
  }
  __syncthreads();
  // ---- phase 3: along z and the pointwise factors; the flux tested along t1
  for (int o = t; o < M3; o += 64) {
    const int q0 = o % M, q1 = (o / M) % M, q2 = o / M2;
    const T wq = tw[q0] * tw[q1] * tw[q2];
    if (TRANS) {
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int e = 0; e < N1; ++e) {
        const int ai = q0 + M * (q1 + M * e);
        gx += tB[q2][e] * A2[1][ai];
        gy += tB[q2][e] * A2[2][ai];
        gz += tD[q2][e] * A2[0][ai];
      }
      V[0][o] = wq * (a.vs[0] * V[0][o] * gx + a.vs[1] * V[1][o] * gy + a.vs[2] * V[2][o] * gz);
    } else {
      T u = T(0);
#pragma unroll
      for (int e = 0; e < N1; ++e) u += tB[q2][e] * A2[0][q0 + M * (q1 + M * e)];
      const T wu = wq * u;
#pragma unroll
      for (int d = 0; d < 3; ++d) V[d][o] = a.vs[d] * V[d][o] * wu;
    }
  }
  for (int o = t; o < 6 * N1 * M; o += 64) {
    const int f = o / (N1 * M), r = o % (N1 * M), p1 = r % N1, q2 = r / N1;
    T acc = T(0);
#pragma unroll
    for (int q1 = 0; q1 < M; ++q1) acc += tB[q1][p1] * fv[f][q1 + M * q2];
    G[f][r] = acc;
  }
  __syncthreads();
  // ---- phase 4: volume test along x; the flux tested along t2
  for (int o = t; o < N1 * M2; o += 64) {
    const int i = o % N1, q12 = o / N1;
    T w0 = T(0), w1 = T(0), w2 = T(0);
#pragma unroll
    for (int q0 = 0; q0 < M; ++q0) {
      if (TRANS) {
        w0 += tB[q0][i] * V[0][q0 + M * q12];
      } else {
        w0 += tD[q0][i] * V[0][q0 + M * q12];
        w1 += tB[q0][i] * V[1][q0 + M * q12];
        w2 += tB[q0][i] * V[2][q0 + M * q12];
      }
    }
    W[0][o] = w0;
    if (!TRANS) {
      W[1][o] = w1;
      W[2][o] = w2;
    }
  }
  for (int o = t; o < 6 * L; o += 64) {
    const int f = o / L, pq = o % L, p1 = pq % N1, p2 = pq / N1;
    T acc = T(0);
#pragma unroll
    for (int q2 = 0; q2 < M; ++q2) acc += tB[q2][p2] * G[f][p1 + N1 * q2];
    rf[f][pq] = acc;
  }
  __syncthreads();
  // ---- phase 5: along y
  for (int o = t; o < L * M; o += 64) {
    const int i = o % N1, j = (o / N1) % N1, q2 = o / L;
    T s0 = T(0), s1 = T(0);
#pragma unroll
    for (int q1 = 0; q1 < M; ++q1) {
      const int wi = i + N1 * (q1 + M * q2);
      if (TRANS) {
        s0 += tB[q1][j] * W[0][wi];
      } else {
        s0 += tB[q1][j] * W[0][wi] + tD[q1][j] * W[1][wi];
        s1 += tB[q1][j] * W[2][wi];
      }
    }
    S[0][o] = s0;
    S[1][o] = s1;
  }
  __syncthreads();
  // ---- phase 6: along z, plus the faces whose layer holds the DoF
  for (int o = t; o < NL; o += 64) {
    const int ij = o % L, k = o / L;
    const int ii[3] = {o % N1, (o / N1) % N1, k};
    T acc = T(0);
#pragma unroll
    for (int q2 = 0; q2 < M; ++q2) {
      acc += tB[q2][k] * S[0][ij + L * q2];
      if (!TRANS) acc += tD[q2][k] * S[1][ij + L * q2];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int d = f / 2;
      if (ii[d] != ((f % 2) ? N1 - 1 : 0)) continue;
      const int p1 = ii[d == 0 ? 1 : 0], p2 = ii[d == 2 ? 1 : 2];
      acc += rf[f][p1 + N1 * p2];
    }
    y[c * NL + o] += acc;
  }
}

// The same terms for the rule with N1 points per direction (intorderadd 0 or 1), the common
// case. Every volume phase then has the N1^3 items (i0, i1, i2) and every face phase the
// 6 N1^2 items (f, r1, r2), so each lane decomposes its items once and keeps the table rows
// and columns it needs, and its values of the field, in registers: the phases read LDS only
// for the data they exchange.
template <int N1, typename T, bool TRANS = false, bool REGTAB = (sizeof(T) == 4)>
__global__ __launch_bounds__(64) void conv_var_sq_kernel(const T* __restrict__ z, T* __restrict__ y,
                                                         const T* __restrict__ bv, const T* __restrict__ bf,
                                                         const VelArgs<T> a, const int* __restrict__ stop) {
  constexpr int L = N1 * N1, NL = L * N1;
  constexpr int FI = (6 * L + 63) / 64;  // face items per lane
  __shared__ T zs[NL];
  __shared__ T zt[6][L];  // neighbour layers (zero without a neighbour)
  __shared__ T V[3][NL];
  __shared__ T A1[TRANS ? 2 : 1][NL], A2[TRANS ? 3 : 1][NL];
  __shared__ T W[3][NL];
  __shared__ T S[2][NL];
  __shared__ T F[6][2][L];
  __shared__ T fv[6][L];
  __shared__ T G[6][L];
  __shared__ T rf[6][L];
  __shared__ T tB[N1][N1], tD[N1][N1];  // the tables when they are not held in registers
  if (stop && *stop) return;
  const long long c = blockIdx.x;
  const int t = threadIdx.x;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  // ---- the lane's volume item and its table rows (B[i_k][.]) and columns (B[.][i_k])
  const bool vl = t < NL;
  const int iv = vl ? t : 0;
  const int ii[3] = {iv % N1, (iv / N1) % N1, iv / L};
  // REGTAB (FP32): the rows and columns in registers; FP64 reads them from LDS at the
  // lane's row / column with compile-time offsets (registers would halve the occupancy)
  T Br[3][N1], Dr[3][N1], Bc[3][N1], Dc[3][N1];
  if constexpr (REGTAB) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int e = 0; e < N1; ++e) {
        Br[k][e] = a.B[ii[k]][e];
        Dr[k][e] = a.Dq[ii[k]][e];
        Bc[k][e] = a.B[e][ii[k]];
        Dc[k][e] = a.Dq[e][ii[k]];
      }
  } else if (t < L) {
    tB[t / N1][t % N1] = a.B[t / N1][t % N1];
    tD[t / N1][t % N1] = a.Dq[t / N1][t % N1];
  }
#define VBR(k_, e_) (REGTAB ? Br[k_][e_] : tB[ii[k_]][e_])
#define VDR(k_, e_) (REGTAB ? Dr[k_][e_] : tD[ii[k_]][e_])
#define VBC(k_, e_) (REGTAB ? Bc[k_][e_] : tB[e_][ii[k_]])
#define VDC(k_, e_) (REGTAB ? Dc[k_][e_] : tD[e_][ii[k_]])
  const T wq = a.w[ii[0]] * a.w[ii[1]] * a.w[ii[2]];
  // ---- the lane's face items (f, r1, r2)
  int fo[FI], fr1[FI], fr2[FI], fbo[FI], fbn[FI], fs1[FI], fs2[FI];
  bool fin[FI], fact[FI], fout[FI];
  T fcw[FI], fR1[FI][N1], fR2[FI][N1], fC1[FI][N1], fC2[FI][N1];
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    const int o = t + 64 * k;
    const bool ok = o < 6 * L;
    const int oo = ok ? o : 0, f = oo / L, r = oo % L, d = f / 2, up = f % 2;
    fo[k] = ok ? o : -1;
    fr1[k] = r % N1;
    fr2[k] = r / N1;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int str[3] = {1, N1, L};
    fbo[k] = (up ? N1 - 1 : 0) * str[d];  // this cell's layer on face f
    fbn[k] = (up ? 0 : N1 - 1) * str[d];  // the neighbour's layer facing it
    fs1[k] = str[t1];
    fs2[k] = str[t2];
    fin[k] = up ? coord[d] < a.n[d] - 1 : coord[d] > 0;
    fout[k] = !fin[k] && a.bc[f] == BC_OUTFLOW;
    fact[k] = ok && (fin[k] || a.bc[f] == BC_DIRICHLET || fout[k]);
    fcw[k] = a.w[fr1[k]] * a.w[fr2[k]] * a.area[d] * (up ? T(1) : T(-1));
    if constexpr (REGTAB) {
#pragma unroll
      for (int e = 0; e < N1; ++e) {
        fR1[k][e] = a.B[fr1[k]][e];
        fR2[k][e] = a.B[fr2[k]][e];
        fC1[k][e] = a.B[e][fr1[k]];
        fC2[k][e] = a.B[e][fr2[k]];
      }
    }
  }
#define FR1(k_, e_) (REGTAB ? fR1[k_][e_] : tB[fr1[k_]][e_])
#define FR2(k_, e_) (REGTAB ? fR2[k_][e_] : tB[fr2[k_]][e_])
#define FC1(k_, e_) (REGTAB ? fC1[k_][e_] : tB[e_][fr1[k_]])
#define FC2(k_, e_) (REGTAB ? fC2[k_][e_] : tB[e_][fr2[k_]])
  // ---- all global loads: the cell, the neighbour layers, the field (into registers)
  if (vl) zs[t] = z[c * NL + t];
  T bvr[3] = {T(0), T(0), T(0)};
  if (vl)
#pragma unroll
    for (int d = 0; d < 3; ++d) bvr[d] = bv[c * 3 * NL + d * NL + t];
  T bfr[FI];
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    bfr[k] = T(0);
    if (fo[k] < 0) continue;
    const int f = fo[k] / L, r = fo[k] % L;
    bfr[k] = bf[c * 6 * L + fo[k]];
    T v = T(0);
    if (fin[k]) {
      const int d = f / 2;
      const long long nb = c + ((f % 2) ? a.stride[d] : -a.stride[d]);
      v = z[nb * NL + fbn[k] + fr1[k] * fs1[k] + fr2[k] * fs2[k]];
    }
    zt[f][r] = v;
  }
  __syncthreads();
  // ---- phase 1: along x; face traces along t1
  if (vl) {
    const int base = N1 * (ii[1] + N1 * ii[2]);
    T acc = T(0), accd = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      acc += VBR(0, e) * zs[e + base];
      if (TRANS) accd += VDR(0, e) * zs[e + base];
    }
    A1[0][t] = acc;
    if (TRANS) A1[1][t] = accd;
  }
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    if (fo[k] < 0) continue;
    const int f = fo[k] / L;
    T as = T(0), an = T(0);
#pragma unroll
    for (int p1 = 0; p1 < N1; ++p1) {
      as += FR1(k, p1) * zs[fbo[k] + p1 * fs1[k] + fr2[k] * fs2[k]];
      an += FR1(k, p1) * zt[f][p1 + N1 * fr2[k]];
    }
    F[f][0][fo[k] % L] = as;
    F[f][1][fo[k] % L] = an;
  }
  __syncthreads();
  // ---- phase 2: along y; the upwinded flux at the face points
  if (vl) {
    T acc = T(0), g0 = T(0), g1 = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int ai = ii[0] + N1 * (e + N1 * ii[2]);
      acc += VBR(1, e) * A1[0][ai];
      if (TRANS) {
        g0 += VBR(1, e) * A1[1][ai];
        g1 += VDR(1, e) * A1[0][ai];
      }
    }
    A2[0][t] = acc;
    if (TRANS) {
      A2[1][t] = g0;
      A2[2][t] = g1;
    }
  }
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    if (fo[k] < 0) continue;
    const int f = fo[k] / L;
    T v = T(0);
    if (fact[k]) {
      T us = T(0), un = T(0);
#pragma unroll
      for (int p2 = 0; p2 < N1; ++p2) {
        us += FR2(k, p2) * F[f][0][fr1[k] + N1 * p2];
        un += FR2(k, p2) * F[f][1][fr1[k] + N1 * p2];
      }
      const T cf = fcw[k] * bfr[k];
      if (TRANS) v = fout[k] ? cf * us : (cf >= T(0) ? cf * (us - un) : T(0));
      else v = (fout[k] || cf >= T(0)) ? cf * us : cf * un;
    }
    fv[f][fo[k] % L] = v;
  }
  __syncthreads();
  // ---- phase 3: along z and the pointwise factors; the flux tested along t1
  if (vl) {
    if (TRANS) {
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int e = 0; e < N1; ++e) {
        const int ai = ii[0] + N1 * (ii[1] + N1 * e);
        gx += VBR(2, e) * A2[1][ai];
        gy += VBR(2, e) * A2[2][ai];
        gz += VDR(2, e) * A2[0][ai];
      }
      V[0][t] = wq * (a.vs[0] * bvr[0] * gx + a.vs[1] * bvr[1] * gy + a.vs[2] * bvr[2] * gz);
    } else {
      T u = T(0);
#pragma unroll
      for (int e = 0; e < N1; ++e) u += VBR(2, e) * A2[0][ii[0] + N1 * (ii[1] + N1 * e)];
      const T wu = wq * u;
#pragma unroll
      for (int d = 0; d < 3; ++d) V[d][t] = a.vs[d] * bvr[d] * wu;
    }
  }
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    if (fo[k] < 0) continue;
    const int f = fo[k] / L;
    T acc = T(0);
#pragma unroll
    for (int q1 = 0; q1 < N1; ++q1) acc += FC1(k, q1) * fv[f][q1 + N1 * fr2[k]];
    G[f][fo[k] % L] = acc;
  }
  __syncthreads();
  // ---- phase 4: volume test along x; the flux tested along t2
  if (vl) {
    const int base = N1 * (ii[1] + N1 * ii[2]);
    T w0 = T(0), w1 = T(0), w2 = T(0);
#pragma unroll
    for (int q0 = 0; q0 < N1; ++q0) {
      if (TRANS) {
        w0 += VBC(0, q0) * V[0][q0 + base];
      } else {
        w0 += VDC(0, q0) * V[0][q0 + base];
        w1 += VBC(0, q0) * V[1][q0 + base];
        w2 += VBC(0, q0) * V[2][q0 + base];
      }
    }
    W[0][t] = w0;
    if (!TRANS) {
      W[1][t] = w1;
      W[2][t] = w2;
    }
  }
#pragma unroll
  for (int k = 0; k < FI; ++k) {
    if (fo[k] < 0) continue;
    const int f = fo[k] / L;
    T acc = T(0);
#pragma unroll
    for (int q2 = 0; q2 < N1; ++q2) acc += FC2(k, q2) * G[f][fr1[k] + N1 * q2];
    rf[f][fo[k] % L] = acc;
  }
  __syncthreads();
  // ---- phase 5: along y
  if (vl) {
    T s0 = T(0), s1 = T(0);
#pragma unroll
    for (int q1 = 0; q1 < N1; ++q1) {
      const int wi = ii[0] + N1 * (q1 + N1 * ii[2]);
      if (TRANS) {
        s0 += VBC(1, q1) * W[0][wi];
      } else {
        s0 += VBC(1, q1) * W[0][wi] + VDC(1, q1) * W[1][wi];
        s1 += VBC(1, q1) * W[2][wi];
      }
    }
    S[0][t] = s0;
    if (!TRANS) S[1][t] = s1;
  }
  __syncthreads();
  // ---- phase 6: along z, plus the faces whose layer holds the DoF
  if (vl) {
    const int ij = ii[0] + N1 * ii[1];
    T acc = T(0);
#pragma unroll
    for (int q2 = 0; q2 < N1; ++q2) {
      acc += VBC(2, q2) * S[0][ij + L * q2];
      if (!TRANS) acc += VDC(2, q2) * S[1][ij + L * q2];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int d = f / 2;
      if (ii[d] != ((f % 2) ? N1 - 1 : 0)) continue;
      acc += rf[f][ii[d == 0 ? 1 : 0] + N1 * ii[d == 2 ? 1 : 2]];
    }
    y[c * NL + t] += acc;
  }
}
#undef VBR
#undef VDR
#undef VBC
#undef VDC
#undef FR1
#undef FR2
#undef FC1
#undef FC2

// Entry (i, j) (row i: test function, column j: trial function) of the cell block of the
// convection terms of cell c: the volume term and the faces' inside-trace parts (the
// neighbour-trace parts couple to other cells). The point diagonal is entry (i, i).
__device__ inline double conv_block_entry(const VelArgs<double>& a, int n1, const double* __restrict__ bv,
                                          const double* __restrict__ bf, long long c, const int (&coord)[3], int i,
                                          int j) {
  const int m = a.m, m2 = m * m, m3 = m2 * m, l2 = n1 * n1;
  const int ii[3] = {i % n1, (i / n1) % n1, i / l2}, jj[3] = {j % n1, (j / n1) % n1, j / l2};
  const double* bc = bv + c * 3 * m3;
  double acc = 0.0;
  for (int q2 = 0; q2 < m; ++q2)
    for (int q1 = 0; q1 < m; ++q1)
      for (int q0 = 0; q0 < m; ++q0) {
        const int q = q0 + m * (q1 + m * q2);
        const double b0 = a.B[q0][ii[0]], b1 = a.B[q1][ii[1]], b2 = a.B[q2][ii[2]];
        const double phij = a.B[q0][jj[0]] * a.B[q1][jj[1]] * a.B[q2][jj[2]];
        const double w = a.w[q0] * a.w[q1] * a.w[q2];
        const double g = a.vs[0] * bc[q] * a.Dq[q0][ii[0]] * b1 * b2 +
                         a.vs[1] * bc[m3 + q] * b0 * a.Dq[q1][ii[1]] * b2 +
                         a.vs[2] * bc[2 * m3 + q] * b0 * b1 * a.Dq[q2][ii[2]];
        acc += w * phij * g;
      }
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, up = f % 2, layer = up ? n1 - 1 : 0;
    if (ii[d] != layer || jj[d] != layer) continue;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const bool interior = up ? coord[d] < a.n[d] - 1 : coord[d] > 0;
    const int bct = interior ? 0 : a.bc[f];
    if (!interior && bct != BC_DIRICHLET && bct != BC_OUTFLOW) continue;
    const double* bfc = bf + (c * 6 + f) * m2;
    const double ns = up ? 1.0 : -1.0;
    for (int q2 = 0; q2 < m; ++q2)
      for (int q1 = 0; q1 < m; ++q1) {
        const double psi2 = a.B[q1][ii[t1]] * a.B[q2][ii[t2]] * a.B[q1][jj[t1]] * a.B[q2][jj[t2]];
        const double cf = a.w[q1] * a.w[q2] * a.area[d] * (ns * bfc[q1 + m * q2]);
        if (bct == BC_OUTFLOW || cf >= 0.0) acc += cf * psi2;
      }
  }
  return acc;
}

// Entry (i, j) of the block coupling cell c to its neighbour n across face f (row i: c's test
// function, column j: n's trial function): the upwind term [cf < 0] cf psi_i psi_j at each face
// point, cf = w area (b . n_c). With trans, the entry of J^T's block (c, n), i.e. of
// J's block (n, c) transposed: [cf > 0] (-cf) psi_i psi_j.
__device__ inline double conv_coupling_entry(const VelArgs<double>& a, int n1, const double* __restrict__ bf,
                                             long long c, int f, int i, int j, bool trans) {
  const int d = f / 2, up = f % 2, l2 = n1 * n1, m = a.m, m2 = m * m;
  const int ii[3] = {i % n1, (i / n1) % n1, i / l2}, jj[3] = {j % n1, (j / n1) % n1, j / l2};
  if (ii[d] != (up ? n1 - 1 : 0) || jj[d] != (up ? 0 : n1 - 1)) return 0.0;
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  const double* bfc = bf + (c * 6 + f) * m2;
  const double ns = up ? 1.0 : -1.0;
  double acc = 0.0;
  for (int q2 = 0; q2 < m; ++q2)
    for (int q1 = 0; q1 < m; ++q1) {
      const double psi2 = a.B[q1][ii[t1]] * a.B[q2][ii[t2]] * a.B[q1][jj[t1]] * a.B[q2][jj[t2]];
      const double cf = a.w[q1] * a.w[q2] * a.area[d] * (ns * bfc[q1 + m * q2]);
      if (trans ? cf > 0.0 : cf < 0.0) acc += (trans ? -cf : cf) * psi2;
    }
  return acc;
}

// Adds a velocity field's convection terms to the block row of cell c of the assembled
// Jacobian (seven n_loc x n_loc blocks, row-major; slot 0 the cell, slot 1 + f the neighbour
// across face f, cols < 0 where there is none): C_c to slot 0, the upwind coupling to the
// others (TRANS: the blocks of J^T). One workgroup per cell.
template <bool TRANS>
__global__ __launch_bounds__(256) void conv_var_jac_kernel(const VelArgs<double>* __restrict__ va, int n1,
                                                           const double* __restrict__ bv, const double* __restrict__ bf,
                                                           double* __restrict__ blocks, const long long* __restrict__ cols) {
  const VelArgs<double>& a = *va;
  const long long c = blockIdx.x;
  const int nl = n1 * n1 * n1, nl2 = nl * nl;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  double* B = blocks + c * 7 * nl2;
  for (int k = threadIdx.x; k < 7 * nl2; k += blockDim.x) {
    const int slot = k / nl2, ij = k % nl2, i = ij / nl, j = ij % nl;
    if (slot == 0) {
      B[k] += TRANS ? conv_block_entry(a, n1, bv, bf, c, coord, j, i) : conv_block_entry(a, n1, bv, bf, c, coord, i, j);
    } else if (cols[c * 7 + slot] >= 0) {
      B[k] += conv_coupling_entry(a, n1, bf, c, slot - 1, i, j, TRANS);
    }
  }
}

// Block SOR with a velocity field: the convection terms' couplings of the cells of one
// hyperplane (order[first .. first + count)) to their neighbours, moved to the right-hand side:
// dtmp_c = d_c - sum_f sum_q [cf < 0] cf psi_c(q) v_n(q), over the lower neighbours (from
// v = 0: the upper ones are still zero) or, INPLACE, over all of them (the upper ones hold the
// old values). The sweep kernel then runs on dtmp. One thread per DoF of the hyperplane.
template <int N1, int M, typename T, bool INPLACE>
__global__ __launch_bounds__(256) void conv_var_sor_kernel(const T* __restrict__ d, const T* __restrict__ v,
                                                           T* __restrict__ dtmp, const T* __restrict__ bf,
                                                           const VelArgs<T> a, const int* __restrict__ order,
                                                           int first, int count) {
  constexpr int L = N1 * N1, NL = L * N1, M2 = M * M;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)count * NL) return;
  const long long c = order[first + gid / NL];
  const int li = (int)(gid % NL);
  const int ii[3] = {li % N1, (li / N1) % N1, li / L};
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  T acc = T(0);
  for (int f = 0; f < 6; ++f) {
    const int dd = f / 2, up = f % 2;
    if (!INPLACE && up) continue;
    if (ii[dd] != (up ? N1 - 1 : 0)) continue;
    if (up ? coord[dd] >= a.n[dd] - 1 : coord[dd] == 0) continue;  // no neighbour
    const long long nb = c + (up ? a.stride[dd] : -a.stride[dd]);
    const int t1 = dd == 0 ? 1 : 0, t2 = dd == 2 ? 1 : 2;
    const T* bfc = bf + (c * 6 + f) * M2;
    const T* vn = v + nb * NL;
    const T ns = up ? T(1) : T(-1);
    for (int q2 = 0; q2 < M; ++q2)
      for (int q1 = 0; q1 < M; ++q1) {
        const T cf = a.w[q1] * a.w[q2] * a.area[dd] * (ns * bfc[q1 + M * q2]);
        if (!(cf < T(0))) continue;
        T un = T(0);
        for (int p2 = 0; p2 < N1; ++p2)
          for (int p1 = 0; p1 < N1; ++p1) un += a.B[q1][p1] * a.B[q2][p2] * vn[face_dof<N1>(f ^ 1, p1, p2)];
        acc += cf * a.B[q1][ii[t1]] * a.B[q2][ii[t2]] * un;
      }
  }
  dtmp[c * NL + li] = d[c * NL + li] - acc;
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
