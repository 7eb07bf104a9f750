// jacobian.hh -- the assembled Jacobian in the 7-point block pattern and its apply.
//
// Layout (the BCRS matrix of GridOperator::jacobian with Blocking::fixed,
// R:dune/pdelab/gridoperator/gridoperator.hh:184-189): per cell c seven dense n_loc x n_loc
// blocks, row-major, blocks[(7 c + s) n_loc^2 + i n_loc + j] = dR_i(c) / dx_j(cols[7 c + s]);
// slot 0 is the cell itself, slot 1 + f its neighbour across face f = 2 d + side (x-, x+,
// y-, y+, z-, z+), with cols = -1 and a zero block where face f is on the boundary.
//
// Every block is a Kronecker product of 1D factors. With L_d the cell's 1D self operator of
// direction d (line_op on a unit vector with zero neighbour traces) and N_f the 1D coupling
// to the neighbour across face f (line_op on a unit neighbour trace):
//   diagonal block      sum_d  (M L_d) in direction d, M in the two others;
//   neighbour block f   (M N_f) in direction d(f), M in the two others.
#ifndef PDELAB_JACOBIAN_HH
#define PDELAB_JACOBIAN_HH

#include "block_jacobi.hh"

namespace pdelab_amd {

// 1D coupling of direction D to the neighbour across the lower (side 0) or upper (side 1)
// face: N[i][j] = (line_op with the neighbour's line e_j and a zero own line)[i].
template <int N1, int D, bool TRANS, bool CONV, bool VARK>
__device__ inline void jac_coupling_d(const DGArgs<double>& a, int side, int t, double kc, const FaceCoef<double>& flo,
                                      const FaceCoef<double>& fhi, double (&N)[N1][N1]) {
  double zero[N1], e[N1], s[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) zero[i] = 0.0;
#pragma unroll
  for (int j = 0; j < N1; ++j) {
#pragma unroll
    for (int i = 0; i < N1; ++i) e[i] = (i == j) ? 1.0 : 0.0;
    if (side == 0)
      line_op<N1, D, double, TRANS, CONV, true, VARK>(a, zero, e, zero, !(t & 1), !(t & 2), kc, flo, fhi, s);
    else
      line_op<N1, D, double, TRANS, CONV, true, VARK>(a, zero, zero, e, !(t & 1), !(t & 2), kc, flo, fhi, s);
#pragma unroll
    for (int i = 0; i < N1; ++i) N[i][j] = s[i];
  }
}

// One workgroup per cell: the seven blocks of its block row, and its column indices.
template <int N1, bool TRANS, bool CONV, bool VARK>
__global__ __launch_bounds__(256) void jacobian_kernel(const DGArgs<double> a, const double* __restrict__ fc,
                                                       const BJGeom g, double* __restrict__ blocks,
                                                       long long* __restrict__ cols) {
  constexpr int NL = N1 * N1 * N1;
  __shared__ double F[9][N1][N1];  // M L_0..2 (self), then M N_f for the six faces
  __shared__ int has[6];
  const long long c = blockIdx.x;
  const int tid = threadIdx.x;
  if (tid < 9) {
    const int d = tid < 3 ? tid : (tid - 3) / 2;
    const int co = bj_coord(g, c, d);
    const int t = bj_type(co, g.n[d]);
    double kc;
    FaceCoef<double> flo, fhi;
    bj_cell_coeffs(a, fc, g, c, d, co, kc, flo, fhi);
    double L[N1][N1];
    bool present = true;
    if (tid < 3) {
      bj_line<N1, TRANS, CONV, VARK>(a, d, t, kc, flo, fhi, L);
    } else {
      const int side = (tid - 3) & 1;
      present = side == 0 ? !(t & 1) : !(t & 2);
      if (present) {
        if (d == 0) jac_coupling_d<N1, 0, TRANS, CONV, VARK>(a, side, t, kc, flo, fhi, L);
        else if (d == 1) jac_coupling_d<N1, 1, TRANS, CONV, VARK>(a, side, t, kc, flo, fhi, L);
        else jac_coupling_d<N1, 2, TRANS, CONV, VARK>(a, side, t, kc, flo, fhi, L);
      }
      has[tid - 3] = present ? 1 : 0;
      cols[c * 7 + 1 + (tid - 3)] = present ? c + (side ? g.stride[d] : -g.stride[d]) : -1;
    }
    for (int i = 0; i < N1; ++i)
      for (int j = 0; j < N1; ++j) {
        double s = 0.0;
        if (present)
          for (int l = 0; l < N1; ++l) s += a.M[i][l] * L[l][j];
        F[tid][i][j] = s;
      }
    if (tid == 0) cols[c * 7] = c;
  }
  __syncthreads();
  double* B = blocks + c * 7 * NL * NL;
  for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
    const int i = ij / NL, j = ij - i * NL;
    const int ii[3] = {i % N1, (i / N1) % N1, i / (N1 * N1)};
    const int jj[3] = {j % N1, (j / N1) % N1, j / (N1 * N1)};
    double m[3], diag = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = a.M[ii[d]][jj[d]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double other = m[(d + 1) % 3] * m[(d + 2) % 3];
      diag += F[d][ii[d]][jj[d]] * other;
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int f = 2 * d + side;
        B[(1 + f) * NL * NL + ij] = has[f] ? F[3 + f][ii[d]][jj[d]] * other : 0.0;
      }
    }
    B[ij] = diag;
  }
}

// y = A z for a block matrix in the layout above (the matrix-based operator apply,
// MatrixAdapter::apply over the assembled BCRS matrix). One workgroup per cell; a group of
// NLP lanes (NL rounded up to a power of two) reads one row of a block coalesced, the
// seven blocks of the row are summed per lane and the group reduces in a fixed order.
template <int N1>
__global__ __launch_bounds__(256) void block_matrix_apply_kernel(const double* __restrict__ blocks,
                                                                 const long long* __restrict__ cols,
                                                                 const double* __restrict__ z, double* __restrict__ y,
                                                                 long long ncells) {
  constexpr int NL = N1 * N1 * N1;
  constexpr int NLP = NL <= 8 ? 8 : (NL <= 32 ? 32 : 64);
  constexpr int RPW = 64 / NLP;               // rows per wave and step
  constexpr int RPS = RPW * 4;                // rows per step of the workgroup (4 waves)
  __shared__ double zs[7][NLP];
  __shared__ long long cs[7];
  const long long c = blockIdx.x;
  const int tid = threadIdx.x;
  if (tid < 7) cs[tid] = cols[c * 7 + tid];
  __syncthreads();
  for (int k = tid; k < 7 * NLP; k += blockDim.x) {
    const int s = k / NLP, j = k - s * NLP;
    const long long cn = cs[s];
    zs[s][j] = (j < NL && cn >= 0 && cn < ncells) ? z[cn * NL + j] : 0.0;
  }
  __syncthreads();
  const int lane = tid & 63, w = tid >> 6;
  const int rg = lane / NLP, jl = lane - rg * NLP;
  for (int i0 = 0; i0 < NL; i0 += RPS) {
    const int i = i0 + w * RPW + rg;
    double acc = 0.0;
    if (i < NL && jl < NL)
      for (int s = 0; s < 7; ++s)
        if (cs[s] >= 0) acc += blocks[((c * 7 + s) * NL + i) * NL + jl] * zs[s][jl];
#pragma unroll
    for (int off = NLP / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, NLP);
    if (jl == 0 && i < NL) y[c * NL + i] = acc;
  }
}

}  // namespace pdelab_amd

#endif
