// block_sor.hh -- one forward block-SOR sweep of the DG operator, matrix-free.
//
// The sweep visits the cells in their index order c = x + nx (y + ny z) from v = 0:
//     v_c = omega D_c^{-1} (d_c - sum_{j < c} A_cj v_j),
// the matrix-free BlockSORPreconditionerLocalOperator
// (R:dune/pdelab/backend/istl/matrixfree/blocksorpreconditioner.hh:38-296, the cell update
// at :221-250) as GridOperatorPreconditioner::apply runs it (gridoperatorpreconditioner.hh:78-84)
// on the zeroed v that dune-istl's BiCGSTAB passes, with the exact inverse of the diagonal
// block (the block solve of AssembledBlockJacobiPreconditionerLocalOperator). The upper
// neighbours' v is still zero when a cell is updated, so only the lower ones contribute.
// Only the face neighbours couple, and the lower-index ones of cell (x, y, z) are
// (x-1, y, z), (x, y-1, z) and (x, y, z-1), all on the hyperplane x + y + z - 1. Every cell
// of hyperplane s therefore depends on hyperplane s - 1 only, and one launch per
// hyperplane, s = 0 .. nx + ny + nz - 3, computes exactly the values of the sequential
// sweep (the same operations on the same operands per cell).
//
// Per cell: the coupling sum_{j<c} A_cj v_j is the operator's line terms of the lower
// neighbours' lines with the cell's own values and its upper neighbours set to zero
// (line_op is linear in (u, ul, ur), and v is still zero there), followed by the three
// mass sweeps of the tensor-product operator (M (x) M (x) M) sum_d (I (x) I (x) M^{-1} A_d);
// then r = d_c - that, and the dense inverse block (per boundary type, or per cell for
// cell-wise kappa) gives v_c = omega D_c^{-1} r.
#ifndef PDELAB_BLOCK_SOR_HH
#define PDELAB_BLOCK_SOR_HH

#include "block_jacobi.hh"

namespace pdelab_amd {

// One wave per workgroup; LPC lanes per cell (the N1^2 line lanes for the line terms, the
// NL DoF lanes for the block solve), CPW cells per wave: 8 (Q1), 2 (Q2), 1 (Q3).
template <int N1>
struct SorShape {
  static constexpr int L = N1 * N1;
  static constexpr int NL = L * N1;
  static constexpr int LPC = NL <= 8 ? 8 : (NL <= 32 ? 32 : 64);
  static constexpr int CPW = 64 / LPC;
};

// Hyperplane s of the sweep: order[first .. first + count) holds its cells (int32 indices,
// ascending). dinv: the transposed dense inverse blocks D_e^{-1} (element [j][i] of entry e
// at e NL^2 + j NL + i), e the cell (per_cell) or its boundary type. Each lane issues its
// row of the block together with the neighbour loads, and the stop flag is read in the same
// batch, so a wave waits for memory twice (the cell index, then everything else).
// INPLACE: one SOR step from the v given (GridOperatorPreconditioner::apply on any v): the upper
// neighbours' lines (not yet updated: they lie on the next hyperplane) enter the coupling too,
// and v_c = (1 - omega) v_c + omega D_c^{-1} r; else the sweep from v = 0 that the solvers use.
template <int N1, typename T, bool CONV, bool VARK, bool INPLACE = false>
__global__ __launch_bounds__(64) void block_sor_wave_kernel(const T* __restrict__ d, T* __restrict__ v,
                                                            const T* __restrict__ fc, const DGArgs<T> a,
                                                            const LineMats<T>* __restrict__ lm,
                                                            const T* __restrict__ dinv, const BJGeom g,
                                                            const int* __restrict__ order, int first,
                                                            int count, T omega, const int* __restrict__ stop) {
  using S = SorShape<N1>;
  constexpr int L = S::L, NL = S::NL, LPC = S::LPC, CPW = S::CPW;
  __shared__ T B0[CPW][NL], B1[CPW][NL], B2[CPW][NL];
  const int t = threadIdx.x;
  const int cw = t / LPC, l = t % LPC;
  const int k = blockIdx.x * CPW + cw;
  const bool cell_ok = k < count;
  const long long c = cell_ok ? order[first + k] : 0;
  const int x = (int)(c % a.n[0]), y = (int)((c / a.n[0]) % a.n[1]), z = (int)(c / ((long long)a.n[0] * a.n[1]));
  const int coord[3] = {x, y, z};
  const bool line_lane = cell_ok && l < L;
  const bool dof_lane = cell_ok && l < NL;
  const int p = l % N1, q = l / N1;
  T* b0 = B0[cw];
  T* b1 = B1[cw];
  T* b2 = B2[cw];
  bool lo[3];
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) lo[dd] = coord[dd] > 0;
  // ---- all global loads
  T dz[N1], ux[N1], uy[N1], uz[N1], drow[NL];
  T wx[N1], wy[N1], wz[N1];  // INPLACE: the upper neighbours' lines
  bool hi[3];
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) hi[dd] = INPLACE && coord[dd] < a.n[dd] - 1;
  T kc[3] = {a.kappa0, a.kappa0, a.kappa0};
  FaceCoef<T> flo[3], fhi[3];
  const FaceCoef<T> fnone{a.wI_lo, a.wI_hi, a.PI};
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) flo[dd] = fhi[dd] = fnone;
  T vold = T(0);
  if (line_lane) {
    const T* vc = v + c * NL;
    const T* dc = d + c * NL;
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      dz[e] = dc[line_index<N1, 2>(p, q, e)];
      ux[e] = lo[0] ? vc[line_index<N1, 0>(p, q, e) - a.stride[0] * NL] : T(0);
      uy[e] = lo[1] ? vc[line_index<N1, 1>(p, q, e) - a.stride[1] * NL] : T(0);
      uz[e] = lo[2] ? vc[line_index<N1, 2>(p, q, e) - a.stride[2] * NL] : T(0);
      wx[e] = hi[0] ? vc[line_index<N1, 0>(p, q, e) + a.stride[0] * NL] : T(0);
      wy[e] = hi[1] ? vc[line_index<N1, 1>(p, q, e) + a.stride[1] * NL] : T(0);
      wz[e] = hi[2] ? vc[line_index<N1, 2>(p, q, e) + a.stride[2] * NL] : T(0);
    }
    if (VARK) {
      const T* fb = fc + c * FC;
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) {
        kc[dd] = fb[dd];
        if (lo[dd]) flo[dd] = FaceCoef<T>{fb[FCF + 3 * dd], fb[FCF + 1 + 3 * dd], fb[FCF + 2 + 3 * dd]};
        if (hi[dd]) {
          const T* fn = fb + a.stride[dd] * FC;
          fhi[dd] = FaceCoef<T>{fn[FCF + 3 * dd], fn[FCF + 1 + 3 * dd], fn[FCF + 2 + 3 * dd]};
        }
      }
    }
  }
  if (dof_lane) {
    long long e;
    if (g.per_cell) e = c;
    else e = bj_type(x, g.n[0]) + 4 * bj_type(y, g.n[1]) + 16 * bj_type(z, g.n[2]);
    const T* DT = dinv + e * NL * NL + l;
#pragma unroll
    for (int j = 0; j < NL; ++j) drow[j] = DT[j * NL];
    if (INPLACE) vold = v[c * NL + l];
  }
  const bool stopped = stop && *stop;  // a device-side Krylov loop has already stopped
  // ---- line terms of the neighbours (own values zero; upper values zero unless INPLACE)
  if (line_lane) {
    T zero[N1], sx[N1], sy[N1], sz[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) zero[e] = sx[e] = sy[e] = sz[e] = T(0);
    if (lo[0] || hi[0])
      line_apply<N1, 0, T, false, CONV, false, VARK>(a, lm, zero, ux, wx, lo[0], hi[0], kc[0], flo[0], fhi[0], sx);
    if (lo[1] || hi[1])
      line_apply<N1, 1, T, false, CONV, false, VARK>(a, lm, zero, uy, wy, lo[1], hi[1], kc[1], flo[1], fhi[1], sy);
    if (lo[2] || hi[2])
      line_apply<N1, 2, T, false, CONV, false, VARK>(a, lm, zero, uz, wz, lo[2], hi[2], kc[2], flo[2], fhi[2], sz);
    // the three directions summed in the y layout through the rotated cell slots
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      b0[lds_index<N1, 2>(p, q, e)] = sz[e];
      b1[lds_index<N1, 0>(p, q, e)] = sx[e];
      uy[e] = sy[e];
    }
  }
  __syncthreads();
  const MassVals<N1, T, false> Mv(a);
  if (line_lane) {
    T w[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int sl = lds_index<N1, 1>(p, q, e);
      w[e] = uy[e] + b0[sl] + b1[sl];
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * w[j];
      b2[lds_index<N1, 1>(p, q, i)] = acc;
    }
  }
  __syncthreads();
  if (line_lane) {
    T w[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) w[e] = b2[lds_index<N1, 0>(p, q, e)];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * w[j];
      b0[lds_index<N1, 0>(p, q, i)] = acc;
    }
  }
  __syncthreads();
  if (line_lane) {
    T w[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) w[e] = b0[lds_index<N1, 2>(p, q, e)];
    // r = d - coupling, in natural DoF order for the block solve
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * w[j];
      b1[line_index<N1, 2>(p, q, i)] = dz[i] - acc;
    }
  }
  __syncthreads();
  if (dof_lane && !stopped) {
    T res = T(0);
#pragma unroll
    for (int j = 0; j < NL; ++j) res += drow[j] * b1[j];
    v[c * NL + l] = INPLACE ? (T(1) - omega) * vold + omega * res : omega * res;
  }
}

}  // namespace pdelab_amd

#endif
