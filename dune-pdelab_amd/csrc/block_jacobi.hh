// block_jacobi.hh -- per-cell block Jacobi preconditioner of the DG operator.
//
// The diagonal block of cell c is D_c = (M (x) M (x) M) K_c with the Kronecker sum
// K_c = I(x)I(x)L_x + I(x)L_y(x)I + L_z(x)I(x)I of the cell's 1D self operators
// L_d = M^{-1} A_d, the operator of one line with both neighbour traces set to zero (what
// line_op computes for a unit vector). The reference inverts D_c by LU with partial
// pivoting (AssembledBlockJacobiPreconditionerLocalOperator,
// R:dune/pdelab/backend/istl/matrixfree/assembledblockjacobipreconditioner.hh:96-242;
// the matrix-based SeqJac over Blocking::fixed, R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:240-243).
// Two exact representations of D_c^{-1} are used here:
//
//  * fast diagonalisation (symmetric A_d: SIPG without convection): with the generalised
//    eigenvectors A_d V_d = M V_d diag(lam_d), V_d^T M V_d = I,
//        D_c^{-1} = (V_z (x) V_y (x) V_x) diag(1 / (lam_x + lam_y + lam_z)) (V_z^T (x) V_y^T (x) V_x^T),
//    stored as (V_d, lam_d) per direction: 3 (n^2 + n) values per cell, or per boundary type
//    for constant coefficients;
//  * dense inverse blocks (any operator): D_c^{-1} = K_c^{-1} (M^{-1} (x) M^{-1} (x) M^{-1}),
//    stored transposed, one block per cell or one per boundary type.
//
// Boundary type of a cell in direction d: 0 interior, 1 lower face on the boundary, 2 upper
// face on the boundary, 3 both. With constant coefficients D_c depends only on these.
#ifndef PDELAB_BLOCK_JACOBI_HH
#define PDELAB_BLOCK_JACOBI_HH

#include "block_reduce.hh"
#include "stream_io.hh"
#include "dg_kernels.hh"
#include "dg_velocity.hh"

namespace pdelab_amd {

enum { BJ_PREC = 0, BJ_ADD = 1, BJ_CGR = 2 };

__host__ __device__ inline int bj_type(int i, int n) { return (i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0); }

struct BJGeom {
  int n[3];
  long long stride[3];
  long long ncells;
  int per_cell;  // data per cell instead of per boundary type
  int vark;      // cell-wise kappa: coefficients from the per-cell face coefficients
};

// 1D mass matrix data: Ci = C^{-1} for the Cholesky factor M = C C^T, and M^{-1}.
struct BJMass {
  double Ci[4][4];
  double Minv[4][4];
};

// Dense application: one lane per DoF, NLP lanes per cell (a power of two so that a cell
// never straddles two waves).
template <int N1>
struct BJShape {
  static constexpr int NL = N1 * N1 * N1;
  static constexpr int NLP = NL <= 8 ? 8 : (NL <= 32 ? 32 : 64);
  static constexpr int CPW = 64 / NLP;       // cells per wave: 8 (Q1), 2 (Q2), 1 (Q3)
  static constexpr int WAVES = 16;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int CPG = CPW * WAVES;    // cells per workgroup: 128 (Q1), 32 (Q2), 16 (Q3)
  static constexpr int ES = N1 * N1 + N1;    // fast-diagonalisation entry: V (row-major), lam
};

// Fast-diagonalisation application: one lane per z-line (N1 DoFs in registers), N1^2
// lanes per cell, as many whole cells per wave as fit.
template <int N1>
struct BJFdmShape {
  static constexpr int NL = N1 * N1 * N1;
  static constexpr int LPC = N1 * N1;        // lanes per cell
  static constexpr int CPW = 64 / LPC;       // cells per wave: 16 (Q1), 7 (Q2), 4 (Q3)
  static constexpr int WAVES = 8;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int CPG = CPW * WAVES;    // cells per workgroup: 128 (Q1), 56 (Q2), 32 (Q3)
  static constexpr int ES = N1 * N1 + N1;
  static constexpr int EW = CPW * 3 * ES;    // per-cell entry values staged per wave
};

// ---------------------------------------------------------------------------
// 1D self operator of direction D for a cell of boundary type t:
// L[i][j] = (line_op e_j)[i] with zero neighbour traces.
template <int N1, int D, bool TRANS, bool CONV, bool VARK>
__device__ inline void bj_line_d(const DGArgs<double>& a, int t, double kc, const FaceCoef<double>& flo,
                                 const FaceCoef<double>& fhi, double (&L)[N1][N1]) {
  double zero[N1], e[N1], s[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) zero[i] = 0.0;
#pragma unroll
  for (int j = 0; j < N1; ++j) {
#pragma unroll
    for (int i = 0; i < N1; ++i) e[i] = (i == j) ? 1.0 : 0.0;
    line_op<N1, D, double, TRANS, CONV, true, VARK>(a, e, zero, zero, !(t & 1), !(t & 2), kc, flo, fhi, s);
#pragma unroll
    for (int i = 0; i < N1; ++i) L[i][j] = s[i];
  }
}

template <int N1, bool TRANS, bool CONV, bool VARK>
__device__ inline void bj_line(const DGArgs<double>& a, int d, int t, double kc, const FaceCoef<double>& flo,
                               const FaceCoef<double>& fhi, double (&L)[N1][N1]) {
  if (d == 0) bj_line_d<N1, 0, TRANS, CONV, VARK>(a, t, kc, flo, fhi, L);
  else if (d == 1) bj_line_d<N1, 1, TRANS, CONV, VARK>(a, t, kc, flo, fhi, L);
  else bj_line_d<N1, 2, TRANS, CONV, VARK>(a, t, kc, flo, fhi, L);
}

// Cell coefficients of direction d for line_op: the cell's coefficient in direction d and its
// two faces normal to d.
__device__ inline void bj_cell_coeffs(const DGArgs<double>& a, const double* __restrict__ fc, const BJGeom& g,
                                      long long c, int d, int coord, double& kc, FaceCoef<double>& flo,
                                      FaceCoef<double>& fhi) {
  if (!g.vark) {
    kc = a.kappa0;
    flo = FaceCoef<double>{a.wI_lo, a.wI_hi, a.PI};
    fhi = flo;
    return;
  }
  const double* f = fc + c * FC;
  kc = f[d];
  flo = FaceCoef<double>{f[FCF + 3 * d], f[FCF + 1 + 3 * d], f[FCF + 2 + 3 * d]};
  fhi = FaceCoef<double>{0.0, 0.0, 0.0};
  if (coord < g.n[d] - 1) {
    const double* fu = fc + (c + g.stride[d]) * FC;
    fhi = FaceCoef<double>{fu[FCF + 3 * d], fu[FCF + 1 + 3 * d], fu[FCF + 2 + 3 * d]};
  }
}

// Coordinate d of cell c with 32-bit divisions (the host keeps ncells below 2^31).
__device__ inline int bj_coord(const BJGeom& g, long long c, int d) {
  const unsigned cu = (unsigned)c, n0 = (unsigned)g.n[0], n1 = (unsigned)g.n[1];
  const unsigned q = cu / n0;
  return d == 0 ? (int)(cu - q * n0) : (d == 1 ? (int)(q % n1) : (int)(q / n1));
}

// 1/x: the hardware reciprocal refined by two Newton steps (FP64), or the hardware
// reciprocal (FP32, 1 ulp).
__device__ __forceinline__ double bj_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float bj_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Symmetric eigen-decomposition S = Q diag(w) Q^T by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, 4th ed., algorithm 8.5.2).
template <int N>
__device__ inline void sym_eig_jacobi(double (&S)[N][N], double (&Q)[N][N], double (&w)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) Q[i][j] = (i == j) ? 1.0 : 0.0;
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) fro += S[i][j] * S[i][j];
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += S[p][q] * S[p][q];
    if (off <= 1e-34 * fro) break;
#pragma unroll
    for (int p = 0; p < N; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = S[p][q];
        if (apq == 0.0) continue;
        // rotation angle that annihilates S[p][q]
        const double tau = (S[q][q] - S[p][p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
#pragma unroll
        for (int k = 0; k < N; ++k) {  // S <- S J (columns p, q)
          const double skp = S[k][p], skq = S[k][q];
          S[k][p] = cs * skp - sn * skq;
          S[k][q] = sn * skp + cs * skq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {  // S <- J^T S (rows p, q)
          const double spk = S[p][k], sqk = S[q][k];
          S[p][k] = cs * spk - sn * sqk;
          S[q][k] = sn * spk + cs * sqk;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {  // Q <- Q J
          const double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = cs * qkp - sn * qkq;
          Q[k][q] = sn * qkp + cs * qkq;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = S[i][i];
}

// Fast-diagonalisation data, one thread per (entry, direction): entries d*4 + t for
// boundary types (constant coefficients) or c*3 + d per cell.
template <int N1, bool TRANS, bool VARK>
__global__ __launch_bounds__(256) void bj_fdm_setup_kernel(const DGArgs<double> a, const double* __restrict__ fc,
                                                           const BJGeom g, const BJMass ch, long long nentries,
                                                           double* __restrict__ out64, float* __restrict__ out32) {
  using S = BJShape<N1>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nentries) return;
  long long c = 0;
  int d, t;
  if (g.per_cell) {
    c = e / 3;
    d = (int)(e - 3 * c);
    t = bj_type(bj_coord(g, c, d), g.n[d]);
  } else {
    d = (int)(e / 4);
    t = (int)(e % 4);
  }
  double kc;
  FaceCoef<double> flo, fhi;
  bj_cell_coeffs(a, fc, g, c, d, g.per_cell ? bj_coord(g, c, d) : 0, kc, flo, fhi);
  double L[N1][N1], A[N1][N1], T1[N1][N1], Sm[N1][N1], Q[N1][N1], w[N1];
  bj_line<N1, TRANS, false, VARK>(a, d, t, kc, flo, fhi, L);
  // A = M L (symmetric), then Sm = Ci A Ci^T
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < N1; ++l) s += a.M[i][l] * L[l][j];
      A[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < N1; ++l) s += ch.Ci[i][l] * A[l][j];
      T1[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < N1; ++l) s += T1[i][l] * ch.Ci[j][l];
      Sm[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int j = i + 1; j < N1; ++j) {
      const double m = 0.5 * (Sm[i][j] + Sm[j][i]);
      Sm[i][j] = m;
      Sm[j][i] = m;
    }
  sym_eig_jacobi<N1>(Sm, Q, w);
  // V = Ci^T Q
  const long long base = e * S::ES;
#pragma unroll
  for (int l = 0; l < N1; ++l)
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N1; ++m) s += ch.Ci[m][l] * Q[m][i];
      if (out64) out64[base + l * N1 + i] = s;
      if (out32) out32[base + l * N1 + i] = (float)s;
    }
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    if (out64) out64[base + N1 * N1 + i] = w[i];
    if (out32) out32[base + N1 * N1 + i] = (float)w[i];
  }
}

// Dense blocks, one workgroup per entry (boundary-type triple tx + 4 ty + 16 tz, or cell).
// out_mode 0: D^{-1} transposed (DT[j][i] = D^{-1}[i][j]) into out64/out32;
// out_mode 1: the block D itself, row-major, into out64 (per cell only).
// VARB: a velocity field b(x) at the quadrature points (va, bv, bf; per cell): the block is
// D = (M (x) M (x) M) K + C_c with C_c the cell block of its convection terms
// (conv_block_entry, transposed for J^T), formed in place of K before the inversion.
template <int N1, bool TRANS, bool CONV, bool VARK, bool VARB = false>
__global__ __launch_bounds__(256) void bj_dense_setup_kernel(const DGArgs<double> a, const double* __restrict__ fc,
                                                             const BJGeom g, const BJMass ch, int out_mode,
                                                             double* __restrict__ out64, float* __restrict__ out32,
                                                             const VelArgs<double>* __restrict__ va,
                                                             const double* __restrict__ bv,
                                                             const double* __restrict__ bf) {
  constexpr int NL = N1 * N1 * N1;
  __shared__ double Ls[3][N1][N1];
  __shared__ double K[NL][NL + 1];
  __shared__ double fcol[NL];
  __shared__ int perm[NL];
  __shared__ int pivrow;
  const long long e = blockIdx.x;
  const int tid = threadIdx.x;
  long long c = 0;
  int t[3];
  if (g.per_cell) {
    c = e;
    for (int d = 0; d < 3; ++d) t[d] = bj_type(bj_coord(g, c, d), g.n[d]);
  } else {
    t[0] = (int)(e % 4);
    t[1] = (int)((e / 4) % 4);
    t[2] = (int)(e / 16);
  }
  if (tid < 3) {
    double kc;
    FaceCoef<double> flo, fhi;
    bj_cell_coeffs(a, fc, g, c, tid, g.per_cell ? bj_coord(g, c, tid) : 0, kc, flo, fhi);
    double L[N1][N1];
    bj_line<N1, TRANS, CONV, VARK>(a, tid, t[tid], kc, flo, fhi, L);
    for (int i = 0; i < N1; ++i)
      for (int j = 0; j < N1; ++j) Ls[tid][i][j] = L[i][j];
  }
  __syncthreads();
  // Kronecker sum, index i = ix + N1 iy + N1^2 iz
  for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
    const int i = ij / NL, j = ij - i * NL;
    const int ix = i % N1, iy = (i / N1) % N1, iz = i / (N1 * N1);
    const int jx = j % N1, jy = (j / N1) % N1, jz = j / (N1 * N1);
    double s = 0.0;
    if (iy == jy && iz == jz) s += Ls[0][ix][jx];
    if (ix == jx && iz == jz) s += Ls[1][iy][jy];
    if (ix == jx && iy == jy) s += Ls[2][iz][jz];
    K[i][j] = s;
  }
  __syncthreads();
  if constexpr (VARB) {
    // K <- D = (M (x) M (x) M) K + C_c, one column at a time through fcol
    const int coord[3] = {bj_coord(g, c, 0), bj_coord(g, c, 1), bj_coord(g, c, 2)};
    for (int j = 0; j < NL; ++j) {
      for (int i = tid; i < NL; i += blockDim.x) {
        const int ix = i % N1, iy = (i / N1) % N1, iz = i / (N1 * N1);
        double s = 0.0;
        for (int l = 0; l < NL; ++l) {
          const int lx = l % N1, ly = (l / N1) % N1, lz = l / (N1 * N1);
          s += a.M[ix][lx] * a.M[iy][ly] * a.M[iz][lz] * K[l][j];
        }
        fcol[i] = s + (TRANS ? conv_block_entry(*va, N1, bv, bf, c, coord, j, i)
                             : conv_block_entry(*va, N1, bv, bf, c, coord, i, j));
      }
      __syncthreads();
      for (int i = tid; i < NL; i += blockDim.x) K[i][j] = fcol[i];
      __syncthreads();
    }
    if (out_mode == 1) {
      for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
        const int i = ij / NL, j = ij - i * NL;
        if (out64) out64[e * NL * NL + ij] = K[i][j];
        if (out32) out32[e * NL * NL + ij] = (float)K[i][j];
      }
      return;
    }
  } else if (out_mode == 1) {
    // D = (M (x) M (x) M) K
    for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
      const int i = ij / NL, j = ij - i * NL;
      const int ix = i % N1, iy = (i / N1) % N1, iz = i / (N1 * N1);
      double s = 0.0;
      for (int l = 0; l < NL; ++l) {
        const int lx = l % N1, ly = (l / N1) % N1, lz = l / (N1 * N1);
        s += a.M[ix][lx] * a.M[iy][ly] * a.M[iz][lz] * K[l][j];
      }
      if (out64) out64[e * NL * NL + ij] = s;
      if (out32) out32[e * NL * NL + ij] = (float)s;
    }
    return;
  }
  // in-place Gauss-Jordan inversion of K with partial (row) pivoting
  for (int k = 0; k < NL; ++k) {
    if (tid < 64) {  // pivot search by the first wave
      double best = -1.0;
      int br = k;
      for (int i = k + tid; i < NL; i += 64) {
        const double v = fabs(K[i][k]);
        if (v > best) { best = v; br = i; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_down(best, off, 64);
        const int orow = __shfl_down(br, off, 64);
        if (ob > best || (ob == best && orow < br)) { best = ob; br = orow; }
      }
      if (tid == 0) { pivrow = br; perm[k] = br; }
    }
    __syncthreads();
    const int p = pivrow;
    if (p != k)
      for (int j = tid; j < NL; j += blockDim.x) {
        const double tmp = K[k][j];
        K[k][j] = K[p][j];
        K[p][j] = tmp;
      }
    __syncthreads();
    const double piv = K[k][k];
    for (int i = tid; i < NL; i += blockDim.x) fcol[i] = K[i][k];
    __syncthreads();
    // row k: K[k][j] /= piv (j != k), K[k][k] = 1 / piv
    for (int j = tid; j < NL; j += blockDim.x) K[k][j] = (j == k) ? 1.0 / piv : K[k][j] / piv;
    __syncthreads();
    for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
      const int i = ij / NL, j = ij - i * NL;
      if (i == k) continue;
      if (j == k) K[i][j] = -fcol[i] * K[k][k];
      else K[i][j] -= fcol[i] * K[k][j];
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = NL - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int i = tid; i < NL; i += blockDim.x) {
        const double tmp = K[i][k];
        K[i][k] = K[i][p];
        K[i][p] = tmp;
      }
    __syncthreads();
  }
  // D^{-1} = K^{-1} (M^{-1} (x) M^{-1} (x) M^{-1}) (VARB: K^{-1} is D^{-1}), written transposed
  for (int ij = tid; ij < NL * NL; ij += blockDim.x) {
    const int j = ij / NL, i = ij - j * NL;
    const int jx = j % N1, jy = (j / N1) % N1, jz = j / (N1 * N1);
    double s = 0.0;
    if (VARB) {
      s = K[i][j];
    } else {
      for (int l = 0; l < NL; ++l) {
        const int lx = l % N1, ly = (l / N1) % N1, lz = l / (N1 * N1);
        s += K[i][l] * ch.Minv[lx][jx] * ch.Minv[ly][jy] * ch.Minv[lz][jz];
      }
    }
    if (out64) out64[e * NL * NL + ij] = s;
    if (out32) out32[e * NL * NL + ij] = (float)s;
  }
}

// ---------------------------------------------------------------------------
// Application. A workgroup takes CPG consecutive cells; every cell lies inside one wave
// (NLP lanes, NL of them own a DoF), so the 1D sweeps exchange values through the wave's
// own LDS region with no workgroup barrier. Vectors are loaded and stored coalesced.
//   BJ_PREC: out = W^{-1} in, one partial of out.in per workgroup
//   BJ_ADD : out += W^{-1} in
//   BJ_CGR : r -= lambda q (lambda = *lam), out = W^{-1} r, partials
//            of (r.r, out.r) per workgroup
// Every mode returns at once when a state is given and its stop flag state[0] is set (the
// device-side Krylov loops).
template <typename T, typename A>
struct BJVec {
  const T* in;
  T* out;
  T* r;
  const T* q;
  const A* lam;
  A* partial;
  const int* state;
};

// Orders the LDS accesses of one wave (the wave executes them in program order; the
// fences keep the compiler from moving them across), as rocPRIM's wave_barrier does.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N1, int D, typename T>
__device__ __forceinline__ T bj_sweep(const T* __restrict__ src, const T (&w)[N1], int ix, int iy, int iz) {
  T s = T(0);
#pragma unroll
  for (int l = 0; l < N1; ++l) {
    const int idx = D == 0 ? l + N1 * iy + N1 * N1 * iz : (D == 1 ? ix + N1 * l + N1 * N1 * iz : ix + N1 * iy + N1 * N1 * l);
    s += w[l] * src[idx];
  }
  return s;
}

// Dense inverse blocks: lane l of a cell computes row l of D^{-1} r_c.
template <int N1, typename T, typename A, int MODE>
__global__ __launch_bounds__(BJShape<N1>::THREADS) void bj_dense_kernel(const BJVec<T, A> v, const T* __restrict__ data,
                                                                       const BJGeom g) {
  using S = BJShape<N1>;
  constexpr int NL = S::NL, NLP = S::NLP, CPW = S::CPW;
  if (v.state && v.state[0]) return;  // a device-side Krylov loop has already stopped
  __shared__ T buf[S::WAVES][64];
  const int tid = threadIdx.x;
  const int w = tid >> 6, lw = tid & 63;
  const int cw = lw / NLP, l = lw - cw * NLP;
  const long long c = (long long)blockIdx.x * S::CPG + (long long)w * CPW + cw;
  const bool act = l < NL && c < g.ncells;
  const long long cc = c < g.ncells ? c : g.ncells - 1;
  const long long idx = cc * NL + (l < NL ? l : 0);
  T lambda = T(0);
  if (MODE == BJ_CGR) lambda = (T)*v.lam;
  A acc0 = A(0), acc1 = A(0);
  T u = T(0);
  if (act) {
    u = v.in[idx];
    if (MODE == BJ_CGR) {
      u -= lambda * v.q[idx];
      v.r[idx] = u;
      acc0 += (A)u * (A)u;
    }
  }
  long long e;
  if (g.per_cell) e = cc;
  else
    e = bj_type(bj_coord(g, cc, 0), g.n[0]) + 4 * bj_type(bj_coord(g, cc, 1), g.n[1]) +
        16 * bj_type(bj_coord(g, cc, 2), g.n[2]);
  const T* DT = data + e * NL * NL + (l < NL ? l : 0);
  T* b0 = buf[w] + cw * NLP;
  b0[l] = u;
  wave_lds_sync();
  T res = T(0);
#pragma unroll 9
  for (int j = 0; j < NL; ++j) res += DT[j * NL] * b0[j];
  if (act) {
    if (MODE == BJ_ADD) v.out[idx] += res;
    else v.out[idx] = res;
    if (MODE == BJ_PREC) acc0 += (A)res * (A)u;
    if (MODE == BJ_CGR) acc1 += (A)res * (A)u;
  }
  // one partial (BJ_PREC) or pair (BJ_CGR) per workgroup, reduced through buf
  if (MODE == BJ_PREC) {
    A r[1] = {acc0};
    block_sum_reuse<S::WAVES, A, 1>(r, &buf[0][0]);
    if (tid == 0) v.partial[blockIdx.x] = r[0];
  } else if (MODE == BJ_CGR) {
    A r[2] = {acc0, acc1};
    block_sum_reuse<S::WAVES, A, 2>(r, &buf[0][0]);
    if (tid == 0) {
      v.partial[2 * blockIdx.x] = r[0];
      v.partial[2 * blockIdx.x + 1] = r[1];
    }
  }
}

// Dense inverse blocks shared by boundary type (constant coefficients), Q1 and Q2: the cells of
// one type form a box (interior or boundary range in each direction), and every wave of this
// kernel takes 64 consecutive cells of one type, one cell per lane, so the block is
// wave-uniform: its rows are read by scalar loads from the row-major table and enter the
// FMAs as scalar operands. The wave stages its cells' vectors through LDS (coalesced loads
// and stores, each lane reading its own cell's column at a stride of NL values, which is odd
// and so spreads the 32 lanes of a pass over distinct banks), keeps its cell's values in
// registers and forms row i of D^{-1} u per step.
struct BJTypes {
  int first[65];       // first wave of boundary type t (t = tx + 4 ty + 16 tz); first[64] = all waves
  int lo[3][4];        // first cell index of type t_d in direction d
  int cnt[3][4];       // cells of type t_d in direction d
};

template <int N1, typename T, typename A, int MODE>
__global__ __launch_bounds__(64) void bj_typed_kernel(const BJVec<T, A> v, const T* __restrict__ rm, const BJTypes ty,
                                                      const BJGeom g) {
  constexpr int NL = N1 * N1 * N1;
  if (v.state && v.state[0]) return;  // a device-side Krylov loop has already stopped
  __shared__ T buf[64 * NL];
  const int lane = threadIdx.x;
  const int wv = blockIdx.x;
  int t = 0;
  for (int s = 0; s < 64; ++s)
    if (wv >= ty.first[s] && wv < ty.first[s + 1]) t = s;
  const int tx = t & 3, tyy = (t >> 2) & 3, tz = t >> 4;
  const int cx = ty.cnt[0][tx], cy = ty.cnt[1][tyy];
  const long long cnt = (long long)cx * cy * ty.cnt[2][tz];
  const long long base = (long long)(wv - ty.first[t]) * 64;
  // global index of the lane's cell (local index base + lane of the type's box; fewer than 2^31
  // cells, which bj_prepare checks)
  int gc = -1;
  {
    const long long li = base + lane;
    if (li < cnt) {
      const unsigned lu = (unsigned)li;
      const unsigned q = lu / (unsigned)cx, iz = q / (unsigned)cy;
      const unsigned ix = lu - q * (unsigned)cx, iy = q - iz * (unsigned)cy;
      gc = (int)((long long)(ty.lo[0][tx] + (int)ix) + g.stride[1] * (ty.lo[1][tyy] + (int)iy) +
                 g.stride[2] * (ty.lo[2][tz] + (int)iz));
    }
  }
  T lambda = T(0);
  if (MODE == BJ_CGR) lambda = (T)*v.lam;
  A acc0 = A(0), acc1 = A(0);
  // coalesced staging: element k of local cell cl at buf[cl NL + k]. All loads of a phase are
  // issued before the first use; lanes past the type's last cell load the first cell of the
  // wave (which exists) and discard it.
  const int cfirst = __shfl(gc, 0);
  auto where = [&](int r, long long& idx) -> bool {
    const int flat = r * 64 + lane, cl = flat / NL, k = flat - cl * NL;
    const int c = __shfl(gc, cl);
    idx = (long long)(c >= 0 ? c : cfirst) * NL + k;
    return c >= 0;
  };
  {
    T st[NL], sq[MODE == BJ_CGR ? NL : 1];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      long long idx;
      where(r, idx);
      st[r] = v.in[idx];
      if (MODE == BJ_CGR) sq[r] = v.q[idx];
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      long long idx;
      const bool ok = where(r, idx);
      T u = ok ? st[r] : T(0);
      if (MODE == BJ_CGR) {
        u = ok ? u - lambda * sq[r] : T(0);
        if (ok) v.r[idx] = u;
        acc0 += (A)u * (A)u;
      }
      buf[r * 64 + lane] = u;
    }
  }
  wave_lds_sync();
  T u[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) u[k] = buf[lane * NL + k];
  wave_lds_sync();  // every lane holds its column before buf is reused for the results
  const T* B = rm + (size_t)t * NL * NL;
#pragma unroll 1
  for (int i = 0; i < NL; ++i) {
    T y = T(0);
#pragma unroll
    for (int k = 0; k < NL; ++k) y += B[i * NL + k] * u[k];
    // slot i of the lane's column still holds u_i (read there: a register array indexed by
    // the loop counter would live in scratch memory)
    const T ui = buf[lane * NL + i];
    buf[lane * NL + i] = y;
    if (MODE == BJ_PREC) acc0 += (A)y * (A)ui;
    if (MODE == BJ_CGR) acc1 += (A)y * (A)ui;
  }
  wave_lds_sync();
  {
    T sy[MODE == BJ_ADD ? NL : 1];
    if (MODE == BJ_ADD) {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        long long idx;
        where(r, idx);
        sy[r] = v.out[idx];
      }
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      long long idx;
      if (where(r, idx)) {
        const T res = buf[r * 64 + lane];
        if (MODE == BJ_ADD) v.out[idx] = sy[r] + res;
        else v.out[idx] = res;
      }
    }
  }
  if (MODE == BJ_PREC) {
    A s[1] = {acc0};
    block_sum_n<64, A, 1>(s);
    if (lane == 0) v.partial[wv] = s[0];
  } else if (MODE == BJ_CGR) {
    A s[2] = {acc0, acc1};
    block_sum_n<64, A, 2>(s);
    if (lane == 0) {
      v.partial[2 * wv] = s[0];
      v.partial[2 * wv + 1] = s[1];
    }
  }
}

// transposed dense blocks (storage j NL + i holds entry (i, j)) -> row-major copies
template <typename T>
__global__ __launch_bounds__(256) void bj_rowmajor_kernel(const T* __restrict__ src, T* __restrict__ dst, int nl,
                                                          long long total) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  const long long e = x / ((long long)nl * nl);
  const int ij = (int)(x - e * nl * nl), i = ij / nl, j = ij - i * nl;
  dst[x] = src[e * nl * nl + (long long)j * nl + i];
}

// Fast diagonalisation: lane (ix, iy) of a cell holds the z-line u(ix, iy, 0..N1-1). The
// x and y sweeps exchange lines through the wave's LDS region; the z sweeps stay in
// registers. D^{-1} = (Vz (x) Vy (x) Vx) diag(1/(lx + ly + lz)) (Vz^T (x) Vy^T (x) Vx^T).
// ZU (boundary-type table only, nx ny a multiple of CPW): every cell of a wave lies in one
// z-plane, so the z entries (V_z, lam_z) are wave-uniform and are read from the table in
// global memory by scalar loads instead of being held per lane (40 fewer VGPRs in FP64).
template <int N1, typename T, typename A, int MODE, bool PER_CELL, bool ZU = false>
__global__ __launch_bounds__(BJFdmShape<N1>::THREADS) void bj_fdm_kernel(const BJVec<T, A> v,
                                                                        const T* __restrict__ data, const BJGeom g) {
  using S = BJFdmShape<N1>;
  constexpr int NL = S::NL, LPC = S::LPC, CPW = S::CPW, ES = S::ES;
  if (v.state && v.state[0]) return;  // a device-side Krylov loop has already stopped
  // lines of the wave's cells, value k of lane l at X[w][k][l]: the sweep reads of a wave
  // touch every LDS bank at most twice (64-bit) or once (32-bit)
  __shared__ T X[S::WAVES][N1][64];
  // per-cell entries of each wave's cells, or the whole table of the 12 boundary-type entries
  __shared__ T ent[PER_CELL ? S::WAVES : 1][PER_CELL ? S::EW : 12 * S::ES];
  const int tid = threadIdx.x;
  const int w = tid >> 6, lw = tid & 63;
  const int cw = lw / LPC, l2 = lw - cw * LPC;
  const int ix = l2 % N1, iy = l2 / N1;
  const long long cbase = (long long)blockIdx.x * S::CPG + (long long)w * CPW;
  const long long c = cbase + cw;
  const bool act = cw < CPW && c < g.ncells;
  const long long cc = c < g.ncells ? c : g.ncells - 1;
  const int cs = cw < CPW ? cw : 0;  // LDS cell slot (idle lanes alias cell 0 and write nothing)
  const long long base = cc * NL + l2;
  // stage the coefficients in LDS: per-cell entries (V_d, lam_d) of the wave's cells, value
  // j = (cell, direction, offset), or the boundary-type table for the workgroup
  if constexpr (PER_CELL) {
#pragma unroll
    for (int j0 = 0; j0 < S::EW; j0 += 64) {
      const int j = j0 + lw;
      if (j < S::EW) {
        long long c2 = cbase + j / (3 * ES);
        if (c2 >= g.ncells) c2 = g.ncells - 1;
        ent[w][j] = data[c2 * 3 * ES + j % (3 * ES)];
      }
    }
  } else {
    for (int j = tid; j < 12 * ES; j += S::THREADS) ent[0][j] = data[j];
    __syncthreads();
  }
  T lambda = T(0);
  if (MODE == BJ_CGR) lambda = (T)*v.lam;
  A acc0 = A(0);
  T u[N1];
#pragma unroll
  for (int k = 0; k < N1; ++k) {
    u[k] = T(0);
    if (act) {
      const long long idx = base + LPC * k;
      if (MODE == BJ_CGR) {
        // r (reread next iteration) and q (dead after this step) stream past the caches
        const T x = ld_stream<T>(v.in + idx) - lambda * ld_stream<T>(v.q + idx);
        st_stream<T>(x, v.r + idx);
        acc0 += (A)x * (A)x;
        u[k] = x;
      } else {
        u[k] = v.in[idx];
      }
    }
  }
  // this lane's coefficients (LDS): the cell's staged entries or its boundary types' entries
  const T* E[3];
  if constexpr (PER_CELL) {
    wave_lds_sync();
#pragma unroll
    for (int d = 0; d < 3; ++d) E[d] = ent[w] + (cs * 3 + d) * ES;
  } else {
    const unsigned cu = (unsigned)cc, n0 = (unsigned)g.n[0], n1 = (unsigned)g.n[1];
    const unsigned q = cu / n0, qy = q / n1;
    const int co[3] = {(int)(cu - q * n0), (int)(q - qy * n1), (int)qy};
#pragma unroll
    for (int d = 0; d < 3; ++d) E[d] = ent[0] + (d * 4 + bj_type(co[d], g.n[d])) * ES;
  }
  // each coefficient set is read after the wave fence that precedes its sweep (a fence
  // keeps the reads from being hoisted), so only one set is live at a time
  T cf[N1], Vz[N1][N1], lz[N1];
  const T* Ez = E[2];
  if constexpr (ZU && !PER_CELL) {
    const unsigned w0 = (unsigned)__builtin_amdgcn_readfirstlane(tid >> 6);
    unsigned c0 = (unsigned)blockIdx.x * S::CPG + w0 * CPW;
    if (c0 >= (unsigned)g.ncells) c0 = (unsigned)g.ncells - 1;
    const unsigned kz = c0 / ((unsigned)g.n[0] * (unsigned)g.n[1]);
    Ez = data + (2 * 4 + bj_type((int)kz, g.n[2])) * ES;
  }
  const int xo = cs * LPC;  // the cell's lines: value k of line (ix, iy) at X[w][k][xo + ix + N1 iy]
  T t[N1];
  // forward x: t = sum_m Vx[m][ix] u(m, iy, .)
#pragma unroll
  for (int k = 0; k < N1; ++k) X[w][k][lw] = u[k];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < N1; ++m) cf[m] = E[0][m * N1 + ix];
#pragma unroll
  for (int k = 0; k < N1; ++k) t[k] = T(0);
#pragma unroll
  for (int m = 0; m < N1; ++m)
#pragma unroll
    for (int k = 0; k < N1; ++k) t[k] += cf[m] * X[w][k][xo + m + N1 * iy];
  wave_lds_sync();
  // forward y
#pragma unroll
  for (int k = 0; k < N1; ++k) X[w][k][lw] = t[k];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < N1; ++m) cf[m] = E[1][m * N1 + iy];
#pragma unroll
  for (int k = 0; k < N1; ++k) u[k] = T(0);
#pragma unroll
  for (int m = 0; m < N1; ++m)
#pragma unroll
    for (int k = 0; k < N1; ++k) u[k] += cf[m] * X[w][k][xo + ix + N1 * m];
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    lz[m] = Ez[N1 * N1 + m];
#pragma unroll
    for (int k = 0; k < N1; ++k) Vz[m][k] = Ez[m * N1 + k];
  }
  const T lxy = E[0][N1 * N1 + ix] + E[1][N1 * N1 + iy];
  // forward z, scaling, backward z (registers); with w = V^T r, r.(D^{-1} r) is the
  // sum of w^2 / lam over the transformed values
  A dot = A(0);
#pragma unroll
  for (int k = 0; k < N1; ++k) {
    T s = T(0);
#pragma unroll
    for (int m = 0; m < N1; ++m) s += Vz[m][k] * u[m];
    t[k] = s * bj_rcp(lxy + lz[k]);
    dot += (A)s * (A)t[k];
  }
#pragma unroll
  for (int k = 0; k < N1; ++k) {
    T s = T(0);
#pragma unroll
    for (int m = 0; m < N1; ++m) s += Vz[k][m] * t[m];
    u[k] = s;
  }
  wave_lds_sync();
  // backward y: t = sum_m Vy[iy][m] u(ix, m, .)
#pragma unroll
  for (int k = 0; k < N1; ++k) X[w][k][lw] = u[k];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < N1; ++m) cf[m] = E[1][iy * N1 + m];
#pragma unroll
  for (int k = 0; k < N1; ++k) t[k] = T(0);
#pragma unroll
  for (int m = 0; m < N1; ++m)
#pragma unroll
    for (int k = 0; k < N1; ++k) t[k] += cf[m] * X[w][k][xo + ix + N1 * m];
  wave_lds_sync();
  // backward x
#pragma unroll
  for (int k = 0; k < N1; ++k) X[w][k][lw] = t[k];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < N1; ++m) cf[m] = E[0][ix * N1 + m];
#pragma unroll
  for (int k = 0; k < N1; ++k) u[k] = T(0);
#pragma unroll
  for (int m = 0; m < N1; ++m)
#pragma unroll
    for (int k = 0; k < N1; ++k) u[k] += cf[m] * X[w][k][xo + m + N1 * iy];
  if (act) {
#pragma unroll
    for (int k = 0; k < N1; ++k) {
      const long long idx = base + LPC * k;
      if (MODE == BJ_ADD) v.out[idx] += u[k];
      else v.out[idx] = u[k];
    }
  }
  if (MODE == BJ_PREC || MODE == BJ_CGR) {
    // out.in (for BJ_CGR out.r with the updated r) from the active lanes only: a lane
    // beyond the wave's cells (lane 63 at Q2) aliases cell 0 in the sweeps and holds its
    // transformed values, which must not be counted twice
    const A d = act ? dot : A(0);
    // one partial (BJ_PREC) or pair (BJ_CGR) per workgroup, reduced through X
    if (MODE == BJ_PREC) {
      A r[1] = {d};
      block_sum_reuse<S::WAVES, A, 1>(r, &X[0][0][0]);
      if (tid == 0) v.partial[blockIdx.x] = r[0];
    } else {
      A r[2] = {acc0, d};
      block_sum_reuse<S::WAVES, A, 2>(r, &X[0][0][0]);
      if (tid == 0) {
        v.partial[2 * blockIdx.x] = r[0];
        v.partial[2 * blockIdx.x + 1] = r[1];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Iterative block Jacobi (IterativeBlockJacobiPreconditionerLocalOperator,
// R:dune/pdelab/backend/istl/matrixfree/iterativeblockjacobipreconditioner.hh:267-450):
// out_c = y with D_c y = in_c solved by dune-istl's BiCGSTAB from y = 0 (half steps counted,
// stop at def < reduction def0 or def < 1e-30, breakdown checks with EPSILON 1e-80),
// preconditioned by weight * diag(D_c)^{-1} (LocalPointJacobiPreconditioner) or not at all.
// One wave per cell, lane l holding DoF l and row l of D_c in registers; the dot products are
// wave sums broadcast from lane 0, so every lane takes the same branches.

enum { IBJ_OK = 0, IBJ_RHO = 1, IBJ_OMEGA = 2, IBJ_H = 3, IBJ_NAN = 4 };

struct IBJOptions {
  double reduction;
  int maxit;
  int precondition;
  double weight;
};

template <typename T>
__device__ __forceinline__ T ibj_wave_sum(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}

// D: the diagonal blocks, row-major (entry e at e NL^2), e the cell or its boundary type.
// err: the first failing block's code is kept (atomicCAS from 0); iters: sum of the block
// solvers' iteration counts (half steps rounded up), when non-null.
// ADD: out += y (GridOperatorPreconditioner::apply accumulates); else out = y.
template <int N1, typename T, bool ADD = false>
__global__ __launch_bounds__(64) void bj_iter_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                     const T* __restrict__ D, const BJGeom g, const IBJOptions opt,
                                                     int* __restrict__ err, unsigned long long* __restrict__ iters,
                                                     const int* __restrict__ stop) {
  constexpr int NL = N1 * N1 * N1;
  if (stop && *stop) return;  // a device-side Krylov loop has already stopped
  __shared__ T xs[64];
  const long long c = blockIdx.x;
  const int l = threadIdx.x;
  const bool act = l < NL;
  long long e;
  if (g.per_cell) e = c;
  else
    e = bj_type(bj_coord(g, c, 0), g.n[0]) + 4 * bj_type(bj_coord(g, c, 1), g.n[1]) +
        16 * bj_type(bj_coord(g, c, 2), g.n[2]);
  T drow[NL];
  const T* Dr = D + e * NL * NL + (long long)(act ? l : 0) * NL;
#pragma unroll
  for (int j = 0; j < NL; ++j) drow[j] = act ? Dr[j] : T(0);
  const T dinv = act ? T(opt.weight) / Dr[l] : T(0);
  // y = D x for the wave's cell (x staged through LDS)
  auto matvec = [&](T xl) -> T {
    __syncthreads();
    xs[l] = xl;
    __syncthreads();
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < NL; ++j) acc += drow[j] * xs[j];
    return act ? acc : T(0);
  };
  auto prec = [&](T vl) -> T { return opt.precondition ? dinv * vl : vl; };
  const T EPS = T(1e-80);
  T x = T(0);
  T r = act ? in[c * NL + l] : T(0);  // r = b - A 0
  const T rt = r;
  T p = T(0), v = T(0), t = T(0), y = T(0);
  T norm = sqrt(ibj_wave_sum(r * r));
  const T def0 = norm;
  double it_last = 0.0;
  int code = IBJ_OK;
  auto converged = [&](T d) { return d < def0 * T(opt.reduction) || d < T(1e-30); };
  if (!(norm == norm) || isinf((double)norm)) code = IBJ_NAN;
  else if (!converged(norm)) {
    T rho = T(1), alpha = T(1), omega = T(1), rho_new, beta, h;
    for (double it = 0.5; it < opt.maxit; it += 0.5) {
      rho_new = ibj_wave_sum(rt * r);
      if (fabs(rho) <= EPS) { code = IBJ_RHO; break; }
      if (fabs(omega) <= EPS) { code = IBJ_OMEGA; break; }
      if (it < 1) p = r;
      else {
        beta = (norm == T(0)) ? T(0) : (rho_new / rho) * (alpha / omega);
        p = beta * (p - omega * v) + r;
      }
      y = prec(p);
      v = matvec(y);
      h = ibj_wave_sum(rt * v);
      if (fabs(h) < EPS) { code = IBJ_H; break; }
      alpha = (norm == T(0)) ? T(0) : rho_new / h;
      x += alpha * y;
      r -= alpha * v;
      norm = sqrt(ibj_wave_sum(r * r));
      it_last = it;
      if (!(norm == norm) || isinf((double)norm)) { code = IBJ_NAN; break; }
      if (converged(norm)) break;
      it += 0.5;
      y = prec(r);
      t = matvec(y);
      const T tr = ibj_wave_sum(t * r), tt = ibj_wave_sum(t * t);
      omega = (norm == T(0)) ? T(0) : tr / tt;
      x += omega * y;
      r -= omega * t;
      rho = rho_new;
      norm = sqrt(ibj_wave_sum(r * r));
      it_last = it;
      if (!(norm == norm) || isinf((double)norm)) { code = IBJ_NAN; break; }
      if (converged(norm)) break;
    }
  }
  if (act) out[c * NL + l] = ADD ? out[c * NL + l] + x : x;
  if (l == 0) {
    if (code != IBJ_OK) atomicCAS(err, 0, code);
    if (iters) atomicAdd(iters, (unsigned long long)ceil(it_last));
  }
}

}  // namespace pdelab_amd

#endif
