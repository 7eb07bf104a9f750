// dg_kernels.hh -- gfx950 kernels of the SIPG/NIPG/IIPG DG operator (ConvectionDiffusionDG)
// on a structured hexahedral grid, in tensor-product ("Kronecker") form.
//
// Operator structure (exact for affine axis-aligned cells, cell-constant kappa,
// constant b and c; see dg_tables.hh for the 1D matrices):
//
//   J = sum_d  (M (x) M)_{tangential to d} (x) A_d  +  V c M (x) M (x) M
//
// where A_d is the 1D DG operator along grid lines in direction d: per cell the
// volume part S_d*kappa*K - Sc_d*b_d*G and the face parts of
// R:dune/pdelab/localoperator/convectiondiffusiondg.hh:271-471 (interior) and
// :684-879 (boundary), S_d = V/h_d^2, Sc_d = V/h_d, V = hx*hy*hz. Every face term is
// a rank-one combination of the traces e0/e1 and normal derivatives d0/d1, so a
// line only needs its own values plus the trace value and trace derivative of the
// two neighbouring lines. The kernel evaluates
//
//   y (+)= (M (x) M (x) M) * sum_d [ M^{-1} A_d z ]_d
//
// with the M^{-1}-premultiplied 1D tables, so each direction is a register-local
// line operation and the tangential mass is applied once at the end.
//
// Work decomposition: a workgroup handles CPB cells (64 for Q1, 32 for Q2, 16 for Q3);
// within a cell, thread (p,q) owns one grid line in each of the three layouts (z-,
// y- and x-lines). The cell's values and the per-cell partial sums live in LDS.
#ifndef PDELAB_DG_KERNELS_HH
#define PDELAB_DG_KERNELS_HH

#include <hip/hip_runtime.h>

#include "block_reduce.hh"

namespace pdelab_amd {

enum { BC_DIRICHLET = 1, BC_NEUMANN = -1, BC_OUTFLOW = -2, BC_NONE = -3 };

template <typename T>
struct DGArgs {
  T MiK[4][4], MiG[4][4], MiGt[4][4], M[4][4];
  T Mie0[4], Mie1[4], Mid0[4], Mid1[4], d0[4], d1[4];
  T Mdiag[4], Kdiag[4], Gdiag[4], e0d0[4], e1d1[4];
  T S[3], Sc[3], b[3];
  T V, c, theta, alphak, kappa0;
  // constant-kappa face coefficients (interior: w of lower cell, w of upper cell,
  // penalty; boundary penalty)
  T wI_lo, wI_hi, PI, PB;
  // constant kappa without convection: the line operator of direction d is S[d] times
  // one centro-symmetric matrix Ah with neighbour couplings Ph/Qh ([0] J, [1] J^T);
  // vc = V c is the reaction term of the x lines
  T Ah[2][4][4], Ph[2][4], Qh[2][4];
  T vc;
  int n[3];
  long long stride[3];
  long long ncells;
  int weightsOn;
  int bc[6];
  int aniso;  // cell-wise diagonal tensor A = diag(a_x, a_y, a_z): kappa holds 3 values per cell
};

// Index of the diffusion coefficient of cell c in direction d: kappa holds one value per
// cell (A = kappa I) or three (A = diag(a_x, a_y, a_z)).
__host__ __device__ __forceinline__ long long kappa_index(int aniso, long long c, int d) {
  return aniso ? 3 * c + d : c;
}

// Maps the hardware block index to a logical block so that each of the 8 XCDs
// (blocks are dispatched round-robin over XCDs) works on one contiguous range of
// cells: neighbouring lines in y and z then hit the same XCD's L2.
__device__ __forceinline__ long long xcd_logical_block(long long bid, long long nb) {
  const long long q = nb >> 3, r = nb & 7;
  const long long x = bid & 7, i = bid >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Interior face coefficients, seen from the inside cell (upper, larger index) as the
// reference assembler does (R:dune/pdelab/gridoperator/default/assembler.hh:181):
// omega and harmonic average convectiondiffusiondg.hh:326-338, penalty :346.
template <typename T>
__host__ __device__ __forceinline__ void face_coeffs(const DGArgs<T>& a, T klo, T khi, T& wlo, T& whi, T& P) {
  if (a.weightsOn) {
    const T den = khi + klo + T(1e-20);
    const T omega_s = klo / den, omega_n = khi / den;
    whi = omega_s * khi;
    wlo = omega_n * klo;
    P = a.alphak * (T(2.0) * khi * klo / den);
  } else {
    whi = T(0.5) * khi;
    wlo = T(0.5) * klo;
    P = a.alphak;
  }
}

// Coefficients of one interior face as face_coeffs returns them: weights of the lower
// and upper cell's normal derivative and the penalty.
template <typename T>
struct FaceCoef {
  T wlo, whi, P;
};

// Per-cell face coefficients for cell-wise kappa, FC values per cell: the cell's diffusion
// coefficient in x, y and z (a_d = n.A n for the faces normal to direction d; all three equal
// for A = kappa I), then (wlo, whi, P) of its lower face in x, y and z (zero where that face is
// a boundary face); the upper face of a cell is the lower face of its upper neighbour.
constexpr int FC = 12;
constexpr int FCF = 3;  // offset of the face coefficients

template <typename T>
__global__ __launch_bounds__(256) void face_coeff_kernel(const double* __restrict__ kappa, T* __restrict__ fc,
                                                         const DGArgs<double> a) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  T* f = fc + c * FC;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double kc = kappa[kappa_index(a.aniso, c, d)];
    f[d] = (T)kc;
    double wlo = 0, whi = 0, P = 0;
    if (coord[d] > 0) face_coeffs<double>(a, kappa[kappa_index(a.aniso, c - a.stride[d], d)], kc, wlo, whi, P);
    f[FCF + 3 * d] = (T)wlo;
    f[FCF + 1 + 3 * d] = (T)whi;
    f[FCF + 2 + 3 * d] = (T)P;
  }
}

// The 1D mass matrix is symmetric and centro-symmetric (M_ij = M_ji = M_{n-1-i,n-1-j}: the
// equidistant Lagrange nodes are symmetric about 1/2). The apply kernels read entry (i, j)
// at the lexicographically first of those four positions, so a sweep references only the
// distinct entries (6 of 16 at Q3): fewer kernel-argument values live in scalar registers.
template <int N1>
__host__ __device__ constexpr int mass_canon(int i, int j) {
  const int a = i * N1 + j, b = j * N1 + i, c = (N1 - 1 - i) * N1 + (N1 - 1 - j), d = (N1 - 1 - j) * N1 + (N1 - 1 - i);
  const int m1 = a < b ? a : b, m2 = c < d ? c : d;
  return m1 < m2 ? m1 : m2;
}
// The distinct mass entries of an apply kernel; with VREG (FP64 constant symmetric
// operators) copied once into vector registers. The FP64 kernels' uniform constants (line
// matrices, traces, mass) exceed the scalar register file, and the compiler then parks
// scalars in vector-register lanes and reloads them with v_readlane, one VALU instruction
// per 32 bits per use; the mass entries, used in the last phase only, are cheaper held in
// vector registers (2 per entry at Q3: 12 in all, within the 128 that keep 4 waves per
// SIMD). The cell-wise-kappa and convection kernels hold more vector registers already and
// keep the entries in scalar registers.
template <int N1, typename T, bool VREG>
struct MassVals {
  T m[N1 * N1];
  __device__ __forceinline__ explicit MassVals(const DGArgs<T>& a) {
#pragma unroll
    for (int k = 0; k < N1 * N1; ++k) {
      if (mass_canon<N1>(k / N1, k % N1) != k) continue;  // only the distinct entries are loaded
      T v = a.M[k / N1][k % N1];
      if constexpr (VREG && sizeof(T) == 8) asm("v_mov_b64 %0, %1" : "=v"(v) : "s"(v));
      m[k] = v;
    }
  }
  __device__ __forceinline__ T operator()(int i, int j) const { return m[mass_canon<N1>(i, j)]; }
};

template <int N1, int D>
__device__ __forceinline__ int line_index(int p, int q, int e) {
  constexpr int L = N1 * N1;
  return D == 0 ? e + N1 * p + L * q : (D == 1 ? p + N1 * e + L * q : p + N1 * q + L * e);
}

// Per-cell apply geometry: CPB cells per workgroup, (k+1)^2 threads per cell.
// LDS keeps each cell in a rotated layout, slot(x,y,z) = ((x+z)%n) + n*((y+z)%n) + n^2*z:
// in each of the three line layouts the n^2 lanes of a cell then touch n^2 distinct
// slots modulo n^2 (conflict-free banks), and the padded cell stride NLP puts the
// cells of one wave on disjoint bank ranges.
template <int N1>
struct ApplyShape {
  static constexpr int L = N1 * N1;
  static constexpr int NL = L * N1;
  static constexpr int NLP = NL + (N1 == 4 ? 16 : 1);
  static constexpr int CPB = N1 == 2 ? 64 : (N1 == 3 ? 32 : 16);
  static constexpr int THREADS = CPB * L;  // 256 (Q1, Q3) or 288 (Q2)
  static constexpr int WAVES = (THREADS + 63) / 64;
};

template <int N1, int D>
__host__ __device__ constexpr int lds_index(int p, int q, int e) {
  // (x, y, z) of element e of the lane's line in layout D
  const int x = D == 0 ? e : p;
  const int yy = D == 0 ? p : (D == 1 ? e : q);
  const int zz = D == 2 ? e : q;
  int xs = x + zz, ys = yy + zz;
  if (xs >= N1) xs -= N1;
  if (ys >= N1) ys -= N1;
  return xs + N1 * ys + N1 * N1 * zz;
}

// lds_index of every lane of a cell: [layout D][lane l = p + N1 q][element e], read by the
// row kernel instead of evaluating the rotation per access.
template <int N1>
struct SlotTable {
  int s[3][N1 * N1][4];
};
template <int N1>
constexpr SlotTable<N1> make_slot_table() {
  SlotTable<N1> t{};
  for (int l = 0; l < N1 * N1; ++l)
    for (int e = 0; e < 4; ++e) {
      const int p = l % N1, q = l / N1, ee = e < N1 ? e : 0;
      t.s[0][l][e] = lds_index<N1, 0>(p, q, ee);
      t.s[1][l][e] = lds_index<N1, 1>(p, q, ee);
      t.s[2][l][e] = lds_index<N1, 2>(p, q, ee);
    }
  return t;
}
static __constant__ const SlotTable<2> kSlots2 = make_slot_table<2>();
static __constant__ const SlotTable<3> kSlots3 = make_slot_table<3>();
static __constant__ const SlotTable<4> kSlots4 = make_slot_table<4>();
template <int N1>
__device__ __forceinline__ const SlotTable<N1>& slot_table() {
  if constexpr (N1 == 2) return kSlots2;
  else if constexpr (N1 == 3) return kSlots3;
  else return kSlots4;
}

// s = M^{-1} (A_d u) on one grid line of a cell (or M^{-1} A_d^T u for TRANS), from the
// line values u, the neighbouring lines ul/ur (valid when has_lo/has_hi), the kappa of
// the cell and the coefficients of its lower and upper face in direction D.
template <int N1, int D, typename T, bool TRANS, bool CONV, bool REACT, bool VARK>
__host__ __device__ __forceinline__ void line_op(const DGArgs<T>& a, const T (&u)[N1], const T (&ul)[N1], const T (&ur)[N1],
                                        bool has_lo, bool has_hi, T kc, const FaceCoef<T>& flo,
                                        const FaceCoef<T>& fhi, T (&s)[N1]) {
  const T val0 = u[0], val1 = u[N1 - 1];
  T der0 = T(0), der1 = T(0);
#pragma unroll
  for (int e = 0; e < N1; ++e) { der0 += a.d0[e] * u[e]; der1 += a.d1[e] * u[e]; }
  const T Sd = a.S[D], Sc = a.Sc[D], bd = a.b[D], theta = a.theta;
  T ce0 = T(0), cd0 = T(0), ce1 = T(0), cd1 = T(0);
  // ---- lower face
  if (has_lo) {
    const T v1l = ul[N1 - 1];
    T d1l = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) d1l += a.d1[e] * ul[e];
    const T wlo = flo.wlo, whi = flo.whi, P = flo.P;
    const T J = v1l - val0;              // jump u_lo - u_hi
    const T G = wlo * d1l + whi * der0;  // weighted average of the normal derivative
    if (!TRANS) { ce0 += Sd * (G - P * J); cd0 += Sd * theta * whi * J; }
    else { ce0 += Sd * (-P * J - theta * G); cd0 += -Sd * whi * J; }
    if (CONV) {
      if (!TRANS) ce0 += Sc * (-bd) * (bd > T(0) ? v1l : val0);
      else if (bd <= T(0)) ce0 += Sc * (-bd) * (val0 - v1l);
    }
  } else {
    const int bct = a.bc[2 * D];
    if (bct == BC_DIRICHLET) {
      const T Pb = VARK ? a.alphak * (a.weightsOn ? kc : T(1)) : a.PB;
      if (!TRANS) { ce0 += Sd * (kc * der0 + Pb * val0); cd0 += -Sd * theta * kc * val0; }
      else { ce0 += Sd * (Pb * val0 - theta * kc * der0); cd0 += Sd * kc * val0; }
      if (CONV && bd <= T(0)) ce0 += Sc * (-bd) * val0;
    } else if (bct == BC_OUTFLOW) {
      if (CONV) ce0 += Sc * (-bd) * val0;
    }
  }
  // ---- upper face
  if (has_hi) {
    const T v0r = ur[0];
    T d0r = T(0);
#pragma unroll
    for (int e = 0; e < N1; ++e) d0r += a.d0[e] * ur[e];
    const T wlo = fhi.wlo, whi = fhi.whi, P = fhi.P;
    const T J = val1 - v0r;
    const T G = wlo * der1 + whi * d0r;
    if (!TRANS) { ce1 += Sd * (P * J - G); cd1 += Sd * theta * wlo * J; }
    else { ce1 += Sd * (P * J + theta * G); cd1 += -Sd * wlo * J; }
    if (CONV) {
      if (!TRANS) ce1 += Sc * bd * (bd > T(0) ? val1 : v0r);
      else if (bd > T(0)) ce1 += Sc * bd * (val1 - v0r);
    }
  } else {
    const int bct = a.bc[2 * D + 1];
    if (bct == BC_DIRICHLET) {
      const T Pb = VARK ? a.alphak * (a.weightsOn ? kc : T(1)) : a.PB;
      if (!TRANS) { ce1 += Sd * (Pb * val1 - kc * der1); cd1 += Sd * theta * kc * val1; }
      else { ce1 += Sd * (Pb * val1 + theta * kc * der1); cd1 += -Sd * kc * val1; }
      if (CONV && bd >= T(0)) ce1 += Sc * bd * val1;
    } else if (bct == BC_OUTFLOW) {
      if (CONV) ce1 += Sc * bd * val1;
    }
  }
  // ---- volume + assembly of the face terms
  const T sk = Sd * kc;
  const T sb = -Sc * bd;
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    T acc = ce0 * a.Mie0[i] + cd0 * a.Mid0[i] + ce1 * a.Mie1[i] + cd1 * a.Mid1[i];
    T kv = T(0);
#pragma unroll
    for (int j = 0; j < N1; ++j) kv += a.MiK[i][j] * u[j];
    acc += sk * kv;
    if (CONV) {
      T gv = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) gv += (TRANS ? a.MiGt[i][j] : a.MiG[i][j]) * u[j];
      acc += sb * gv;
    }
    if (REACT && D == 0) acc += a.V * a.c * u[i];
    s[i] = acc;
  }
}


// Constant-coefficient line matrices of direction d, extracted on the host from line_op
// (build_line_mats): s = A u + Pl ul[n-1] + Ql (d1 . ul) + Pr ur[0] + Qr (d0 . ur) with
// interior faces; a boundary face replaces its neighbour term by Dlo u (Dhi u).
template <typename T>
struct LineMats {
  T A[3][4][4], Dlo[3][4][4], Dhi[3][4][4];
  T Pl[3][4], Ql[3][4], Pr[3][4], Qr[3][4];
};

// Constant kappa without convection (line_op with VARK = CONV = false): the operator of
// direction d is S[d] (Ah u + Ph ul[n-1] + Qh (d1 . ul) + mirrored upper-neighbour
// terms) with Ah centro-symmetric, Pr[i] = Ph[n-1-i], Qr[i] = -Qh[n-1-i] and
// d0[j] = -d1[n-1-j]; a boundary face adds the correction Dlo u (Dhi u) of the line
// matrices instead of the neighbour term. The host extracts all of it from line_op and
// checks the structure (build_line_mats), so both forms are the same operator; this one
// needs ~44 multiply-adds per line and few scalar registers.
template <int N1, int D, typename T, bool TRANS, bool REACT>
__device__ __forceinline__ void line_op_sym(const DGArgs<T>& a, const LineMats<T>* __restrict__ lm, const T (&u)[N1],
                                            const T (&ul)[N1], const T (&ur)[N1], bool has_lo, bool has_hi,
                                            T (&s)[N1]) {
  constexpr int tr = TRANS ? 1 : 0;
  T acc[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    T v = T(0);
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      const bool keep = i * N1 + j <= (N1 - 1 - i) * N1 + (N1 - 1 - j);
      v += a.Ah[tr][keep ? i : N1 - 1 - i][keep ? j : N1 - 1 - j] * u[j];
    }
    acc[i] = v;
  }
  if (has_lo) {
    const T vl = ul[N1 - 1];
    T dl = T(0);
#pragma unroll
    for (int j = 0; j < N1; ++j) dl += a.d1[j] * ul[j];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      acc[i] += a.Ph[tr][i] * vl;
      acc[i] += a.Qh[tr][i] * dl;
    }
  }
  if (has_hi) {
    const T vr = ur[0];
    T dr = T(0);  // -(d0 . ur)
#pragma unroll
    for (int j = 0; j < N1; ++j) dr += a.d1[N1 - 1 - j] * ur[j];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      acc[i] += a.Ph[tr][N1 - 1 - i] * vr;
      acc[i] += a.Qh[tr][N1 - 1 - i] * dr;
    }
  }
  const T Sd = a.S[D];
#pragma unroll
  for (int i = 0; i < N1; ++i) s[i] = Sd * acc[i];
  if (!has_lo) {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T v = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) v += lm->Dlo[D][i][j] * u[j];
      s[i] += v;
    }
  }
  if (!has_hi) {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T v = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) v += lm->Dhi[D][i][j] * u[j];
      s[i] += v;
    }
  }
  if (REACT && D == 0) {
#pragma unroll
    for (int i = 0; i < N1; ++i) s[i] += a.vc * u[i];
  }
}

// The line operator used by the apply kernels: the general form for cell-wise kappa or
// convection (CONV also marks a constant-coefficient operator without the symmetric
// structure), the symmetric constant-coefficient form otherwise.
template <int N1, int D, typename T, bool TRANS, bool CONV, bool REACT, bool VARK>
__device__ __forceinline__ void line_apply(const DGArgs<T>& a, const LineMats<T>* __restrict__ lm, const T (&u)[N1],
                                           const T (&ul)[N1], const T (&ur)[N1], bool has_lo, bool has_hi, T kc,
                                           const FaceCoef<T>& flo, const FaceCoef<T>& fhi, T (&s)[N1]) {
  if constexpr (VARK || CONV)
    line_op<N1, D, T, TRANS, CONV, REACT, VARK>(a, u, ul, ur, has_lo, has_hi, kc, flo, fhi, s);
  else
    line_op_sym<N1, D, T, TRANS, REACT>(a, lm, u, ul, ur, has_lo, has_hi, s);
}

template <int N1, int D, typename T>
__device__ __forceinline__ void load_line(const T* __restrict__ base, int off, int p, int q, T (&u)[N1]) {
#pragma unroll
  for (int e = 0; e < N1; ++e) u[e] = base[off + line_index<N1, D>(p, q, e)];
}

// y (+)= J z   (or J^T z), general grids.  ACCUM: add into y (GridOperator::jacobian_apply
// semantics); otherwise y = J z. A workgroup takes CPB consecutive cells in lexicographic
// order (the chunk may span x-rows).
//
// Every global access is a coalesced z-line access (n^2 lanes of a cell read or write
// n^2 consecutive values per instruction). Lines in x and y are served from LDS:
// the own cells, the cells of the rows y-1 and y+1 (same x positions) and the two cells
// beyond the ends of the chunk are staged there; x-neighbours inside the chunk are the
// staged own cells. All global loads are issued before the first wait.
template <int N1, typename T, bool TRANS, bool CONV, bool REACT, bool VARK, bool ACCUM, bool DOT>
__global__ __launch_bounds__(ApplyShape<N1>::THREADS) void dg_apply_kernel(const T* __restrict__ z,
                                                                           T* __restrict__ y,
                                                                           const T* __restrict__ fc,
                                                                           const DGArgs<T> a,
                                                                           const LineMats<T>* __restrict__ lm,
                                                                           long long nblocks,
                                                                           T* __restrict__ dotp,
                                                                           const int* __restrict__ stop) {
  if (stop && *stop) return;  // a device-side Krylov loop has already stopped
  using SH = ApplyShape<N1>;
  constexpr int L = SH::L, NL = SH::NL, NLP = SH::NLP, CPB = SH::CPB;
  __shared__ T U[(CPB + 2) * NLP];  // [0] cell before the chunk, [1..CPB] own cells, [CPB+1] cell after
  __shared__ T YM[CPB * NLP];       // row y-1, later the z partial sums and the x mass sweep
  __shared__ T YP[CPB * NLP];       // row y+1, later the x partial sums
  const MassVals<N1, T, !VARK && !CONV> Mv(a);
  const int t = threadIdx.x;
  const int lc = t / L, l = t % L, p = l % N1, q = l / N1;
  const long long blk = xcd_logical_block(blockIdx.x, nblocks);
  const long long c0 = blk * CPB;
  int coord[3];
  const long long c = c0 + lc;
  const bool active = c < a.ncells;
  {
    const long long cc = active ? c : 0;
    coord[0] = (int)(cc % a.n[0]);
    coord[1] = (int)((cc / a.n[0]) % a.n[1]);
    coord[2] = (int)(cc / ((long long)a.n[0] * a.n[1]));
  }
  const int off = lc * NL;  // lane's cell block relative to c0
  const T* zc = z + c0 * NL;
  bool lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = active && coord[d] > 0;
    hi[d] = active && coord[d] < a.n[d] - 1;
  }
  T uz[N1], zl[N1], zr[N1], yl[N1], yr[N1], xe[N1];
  T kc[3] = {a.kappa0, a.kappa0, a.kappa0};
  FaceCoef<T> flo[3], fhi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) flo[d] = fhi[d] = FaceCoef<T>{a.wI_lo, a.wI_hi, a.PI};
  // the end-cell loads: lanes of cell 0 load the cell before the chunk, lanes of the
  // last cell the cell after it (only where that cell exists and is an x-neighbour)
  const bool first = lc == 0, last = lc == CPB - 1;
  const bool ld_xe = active && ((first && lo[0]) || (last && hi[0]));
  if (active) {
    load_line<N1, 2, T>(zc, off, p, q, uz);
    if (lo[2]) load_line<N1, 2, T>(zc - a.stride[2] * NL, off, p, q, zl);
    if (hi[2]) load_line<N1, 2, T>(zc + a.stride[2] * NL, off, p, q, zr);
    if (lo[1]) load_line<N1, 2, T>(zc - a.stride[1] * NL, off, p, q, yl);
    if (hi[1]) load_line<N1, 2, T>(zc + a.stride[1] * NL, off, p, q, yr);
    if (ld_xe) load_line<N1, 2, T>(zc + (first && lo[0] ? -NL : NL), off, p, q, xe);
    if (VARK) {
      const T* fb = fc + (c0 + lc) * FC;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        kc[d] = fb[d];
        if (lo[d]) flo[d] = FaceCoef<T>{fb[FCF + 3 * d], fb[FCF + 1 + 3 * d], fb[FCF + 2 + 3 * d]};
        if (hi[d]) {
          const T* fn = fb + a.stride[d] * FC;
          fhi[d] = FaceCoef<T>{fn[FCF + 3 * d], fn[FCF + 1 + 3 * d], fn[FCF + 2 + 3 * d]};
        }
      }
    }
  }
  T* Ul = U + (lc + 1) * NLP;
  T* YMl = YM + lc * NLP;
  T* YPl = YP + lc * NLP;
  if (active) {
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int sl = lds_index<N1, 2>(p, q, e);
      Ul[sl] = uz[e];
      if (lo[1]) YMl[sl] = yl[e];
      if (hi[1]) YPl[sl] = yr[e];
      if (ld_xe) (first && lo[0] ? U : U + (CPB + 1) * NLP)[sl] = xe[e];
    }
  }
  __syncthreads();
  T sx[N1], sy[N1], sz[N1];
  if (active) {
    line_apply<N1, 2, T, TRANS, CONV, REACT, VARK>(a, lm, uz, zl, zr, lo[2], hi[2], kc[2], flo[2], fhi[2], sz);
    T u[N1], ul[N1], ur[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int sl = lds_index<N1, 1>(p, q, e);
      u[e] = Ul[sl];
      ul[e] = lo[1] ? YMl[sl] : T(0);
      ur[e] = hi[1] ? YPl[sl] : T(0);
    }
    line_apply<N1, 1, T, TRANS, CONV, REACT, VARK>(a, lm, u, ul, ur, lo[1], hi[1], kc[1], flo[1], fhi[1], sy);
    // x-neighbours: staged cell lc-1 / lc+1 (U[0] and U[CPB+1] hold the end cells)
    const T* xlo = Ul - NLP;
    const T* xhi = Ul + NLP;
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int sl = lds_index<N1, 0>(p, q, e);
      u[e] = Ul[sl];
      ul[e] = lo[0] ? xlo[sl] : T(0);
      ur[e] = hi[0] ? xhi[sl] : T(0);
    }
    line_apply<N1, 0, T, TRANS, CONV, REACT, VARK>(a, lm, u, ul, ur, lo[0], hi[0], kc[0], flo[0], fhi[0], sx);
  }
  __syncthreads();  // all reads of the staged cells done
  T pq = T(0);
  const T* PZ = uz;
  if (active) {
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      YMl[lds_index<N1, 2>(p, q, e)] = sz[e];
      YPl[lds_index<N1, 0>(p, q, e)] = sx[e];
    }
  }
  __syncthreads();
  if (active) {
    // y-layout: s = sy + sz + sx, then the mass along y
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      const int sl = lds_index<N1, 1>(p, q, e);
      v[e] = sy[e] + YMl[sl] + YPl[sl];
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
      Ul[lds_index<N1, 1>(p, q, i)] = acc;
    }
  }
  __syncthreads();
  if (active) {
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) v[e] = Ul[lds_index<N1, 0>(p, q, e)];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
      YMl[lds_index<N1, 0>(p, q, i)] = acc;
    }
  }
  __syncthreads();
  if (active) {
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) v[e] = YMl[lds_index<N1, 2>(p, q, e)];
    T* yc = y + c0 * NL;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
      const int idx = off + line_index<N1, 2>(p, q, i);
      if (ACCUM) yc[idx] += acc;
      else yc[idx] = acc;
      if (DOT) pq += PZ[i] * acc;
    }
  }
  if (DOT) {
    // fused z . (J z) for the Krylov step length, one partial per workgroup, reduced
    // through the kernel's own LDS array (a separate array would cost a workgroup per CU)
    T v[1] = {pq};
    block_sum_reuse<SH::WAVES, T, 1>(v, U);
    if (threadIdx.x == 0) dotp[blockIdx.x] = v[0];
  }
}

// y (+)= J z (or J^T z), staged through registers: the CPB cells of a workgroup are one
// chunk of an x-row (so the y and z boundary flags are uniform over the workgroup; when
// CPB does not divide nx, the last chunk of a row is moved back to end at the row end),
// consecutive logical blocks walk z fastest, then y, then the x-chunk. The chunk and its
// two end cells are written to one LDS array in the rotated, conflict-free cell layout
// of lds_index; each lane reads its twelve slot offsets once from SlotTable, so every
// LDS access is a register plus an immediate offset. The z-neighbour lines (z-layout) and the y-neighbour lines (y-layout) are
// loaded straight into registers; the regions after the chunk hold the partial sums.
//
// FP64 Q3 stores the cells unpadded (64 doubles) and separates neighbouring cells' banks
// by their layout parity instead: cell lc (lc = -1 .. CPB, in every region) keeps slot s
// at s ^ (16 (lc & 1)), which moves odd cells by half of the 64 four-byte banks, as the
// padded stride of 80 doubles does. A 64-lane b64 access is served in two passes of 32
// lanes (two cells of opposite parity), so the banks of each pass are those of the padded
// layout, and the 25,600 B of LDS fit six workgroups per CU instead of five.
template <int N1, typename T, bool TRANS, bool CONV, bool REACT, bool VARK, bool ACCUM, bool DOT, int ZW = 1>
__global__ __launch_bounds__(ApplyShape<N1>::THREADS) void dg_apply_rowreg_kernel(
    const T* __restrict__ z, T* __restrict__ y, const T* __restrict__ fc, const DGArgs<T> a,
    const LineMats<T>* __restrict__ lm, long long nblocks, T* __restrict__ dotp, const int* __restrict__ stop) {
  if (stop && *stop) return;  // a device-side Krylov loop has already stopped
  using SH = ApplyShape<N1>;
  constexpr int L = SH::L, NL = SH::NL, CPB = SH::CPB;
  // parity-swizzled unpadded cells (FP64 Q3) or the padded stride
  constexpr bool XS = sizeof(T) == 8 && N1 == 4;
  constexpr int NLP = XS ? NL : SH::NLP;
  // cell before the chunk, the CPB own cells, cell after the chunk (from UO - NLP on);
  // two chunk-sized regions (YMO, YPO) for the partial sums and the mass sweeps
  constexpr int UO = NLP, YMO = (CPB + 2) * NLP, YPO = (2 * CPB + 2) * NLP;
  __shared__ T LS[(3 * CPB + 2) * NLP];
  const MassVals<N1, T, !VARK && !CONV> Mv(a);
  const int t = threadIdx.x;
  const int lc = t / L, l = t % L, p = l % N1, q = l / N1;
  const long long blk = xcd_logical_block(blockIdx.x, nblocks);
  const unsigned nzu = (unsigned)a.n[2], nyu = (unsigned)a.n[1];
  const unsigned ub = (unsigned)blk;
  // ZW > 1: the workgroup walks ZW consecutive z-chunks (nz % ZW == 0), carrying the
  // z-lines: a step's own lines are the next step's lower neighbours, its upper neighbours
  // the next step's own lines
  const unsigned nzw = nzu / ZW;
  const unsigned kzw = ub % nzw, rest = ub / nzw;
  const unsigned jy = rest % nyu, xc = rest / nyu;
  // x0: first cell of the chunk. A partial last chunk of a row (CPB does not divide nx;
  // the launch ensures nx > CPB) is moved back to end at the row end, so every lane has a
  // cell; its lanes over cells of the previous chunk store nothing. All of it is uniform
  // over the workgroup, and full chunks store unmasked.
  const int xs = (int)(xc * CPB);
  const bool full = xs + CPB <= a.n[0];
  const int x0 = full ? xs : a.n[0] - CPB;
  T uz[N1], zl[N1], zr[N1];
  T pq = T(0);
  for (int s = 0; s < ZW; ++s) {
  const unsigned kz = kzw * ZW + s;
  const long long c0 = (long long)x0 + a.stride[1] * jy + a.stride[2] * kz;
  const bool xlo_ok = x0 > 0, xhi_ok = x0 + CPB < a.n[0];
  const int gx = x0 + lc;
  bool lo[3], hi[3];
  lo[0] = gx > 0;
  hi[0] = gx < a.n[0] - 1;
  lo[1] = jy > 0;
  hi[1] = (int)jy < a.n[1] - 1;
  lo[2] = kz > 0;
  hi[2] = (int)kz < a.n[2] - 1;
  const int off = lc * NL;  // lane's cell block relative to c0
  const T* zc = z + c0 * NL;
  T yl[N1], yr[N1], xe[N1];
  T kc[3] = {a.kappa0, a.kappa0, a.kappa0};
  FaceCoef<T> flo[3], fhi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) flo[d] = fhi[d] = FaceCoef<T>{a.wI_lo, a.wI_hi, a.PI};
  const bool first = lc == 0, last = lc == CPB - 1;
  const bool ld_xe = (first && xlo_ok) || (last && xhi_ok);
  if (s == 0) {
    load_line<N1, 2, T>(zc, off, p, q, uz);
    if (lo[2]) load_line<N1, 2, T>(zc - a.stride[2] * NL, off, p, q, zl);
  } else {
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      zl[e] = uz[e];
      uz[e] = zr[e];
    }
    __syncthreads();  // the previous step's reads of the LDS array are done
  }
  if (hi[2]) load_line<N1, 2, T>(zc + a.stride[2] * NL, off, p, q, zr);
  // the y-neighbours' lines are loaded in y-layout (the lane's y-line at x = p, z = q), so
  // the y line operation takes them from registers: 16 fewer LDS accesses per lane than
  // staging them in z-layout, for 32-byte rather than 128-byte segments per cell and load
  if (lo[1]) load_line<N1, 1, T>(zc - a.stride[1] * NL, off, p, q, yl);
  if (hi[1]) load_line<N1, 1, T>(zc + a.stride[1] * NL, off, p, q, yr);
  if (ld_xe) load_line<N1, 2, T>(zc + (first ? -NL : NL), off, p, q, xe);
  // cell-wise kappa: the cell's kappa and the z faces now, the y and x faces after the
  // staging barrier (fewer registers live across the loads and the staging)
  const T* fb = fc + (c0 + lc) * FC;
  auto load_faces = [&](int d) {
    kc[d] = fb[d];
    if (lo[d]) flo[d] = FaceCoef<T>{fb[FCF + 3 * d], fb[FCF + 1 + 3 * d], fb[FCF + 2 + 3 * d]};
    if (hi[d]) {
      const T* fn = fb + a.stride[d] * FC;
      fhi[d] = FaceCoef<T>{fn[FCF + 3 * d], fn[FCF + 1 + 3 * d], fn[FCF + 2 + 3 * d]};
    }
  };
  if (VARK) load_faces(2);
  // slot of element e of the lane's x-, y- and z-line, relative to the cell region
  // (an x-neighbour, of the other parity, keeps the same element at the own slot ^ XP)
  constexpr int XP = XS ? 16 : 0;
  int ax[N1], ay[N1], az[N1];
  {
    const SlotTable<N1>& st = slot_table<N1>();
    const int cb = lc * NLP, px = (lc & 1) * XP;
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      ax[e] = cb + (st.s[0][l][e] ^ px);
      ay[e] = cb + (st.s[1][l][e] ^ px);
      az[e] = cb + (st.s[2][l][e] ^ px);
    }
  }
  {
    const int xo = first ? UO - NLP : UO + NLP;
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      LS[UO + az[e]] = uz[e];
      if (ld_xe) LS[xo + (az[e] ^ XP)] = xe[e];
    }
  }
  __syncthreads();
  if (VARK) {
    load_faces(1);
    load_faces(0);
  }
  // the partial sums of z and x go to the YMO / YPO regions as soon as they are formed:
  // those regions hold nothing of this step before (the step's first barrier ordered the
  // previous step's reads of them), so the writes need no barrier after the UO reads
  T sy[N1];
  {
    {
      T sz[N1];
      line_apply<N1, 2, T, TRANS, CONV, REACT, VARK>(a, lm, uz, zl, zr, lo[2], hi[2], kc[2], flo[2], fhi[2], sz);
#pragma unroll
      for (int e = 0; e < N1; ++e) LS[YMO + az[e]] = sz[e];
    }
    T u[N1], ul[N1], ur[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      u[e] = LS[UO + ay[e]];
      ul[e] = lo[1] ? yl[e] : T(0);
      ur[e] = hi[1] ? yr[e] : T(0);
    }
    line_apply<N1, 1, T, TRANS, CONV, REACT, VARK>(a, lm, u, ul, ur, lo[1], hi[1], kc[1], flo[1], fhi[1], sy);
    // x-neighbours: the staged cells lc-1 and lc+1 (the end cells for the chunk ends)
#pragma unroll
    for (int e = 0; e < N1; ++e) {
      u[e] = LS[UO + ax[e]];
      const int xn = ax[e] ^ XP;
      ul[e] = lo[0] ? LS[UO - NLP + xn] : T(0);
      ur[e] = hi[0] ? LS[UO + NLP + xn] : T(0);
    }
    T sx[N1];
    line_apply<N1, 0, T, TRANS, CONV, REACT, VARK>(a, lm, u, ul, ur, lo[0], hi[0], kc[0], flo[0], fhi[0], sx);
#pragma unroll
    for (int e = 0; e < N1; ++e) LS[YPO + ax[e]] = sx[e];
  }
  __syncthreads();
  {
    // y-layout: s = sy + sz + sx, then the mass along y
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) v[e] = sy[e] + LS[YMO + ay[e]] + LS[YPO + ay[e]];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
      LS[UO + ay[i]] = acc;
    }
  }
  __syncthreads();
  {
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) v[e] = LS[UO + ax[e]];
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
      LS[YMO + ax[i]] = acc;
    }
  }
  __syncthreads();
  {
    T v[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) v[e] = LS[YMO + az[e]];
    T* yc = y + c0 * NL;
    auto store = [&](bool own) {
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) acc += Mv(i, j) * v[j];
        const int idx = off + line_index<N1, 2>(p, q, i);
        if (own) {
          if (ACCUM) yc[idx] += acc;
          else yc[idx] = acc;
          if (DOT) pq += uz[i] * acc;
        }
      }
    };
    if (full) store(true);
    else store(gx >= xs);
  }
  }  // z-walk step
  if (DOT) {
    // fused z . (J z) for the Krylov step length, one partial per workgroup, reduced
    // through the kernel's own LDS array (a separate array would cost a workgroup per CU)
    T v[1] = {pq};
    block_sum_reuse<SH::WAVES, T, 1>(v, LS);
    if (threadIdx.x == 0) dotp[blockIdx.x] = v[0];
  }
}

// d += diag(J) (PointDiagonalLocalOperatorWrapper, R:dune/pdelab/localoperator/
// pointdiagonalwrapper.hh:51-131): per DoF, sum over directions of the own-cell 1D
// diagonal times the tangential mass diagonals. One thread per DoF.
template <int N1, typename T, bool VARK>
__global__ __launch_bounds__(256) void dg_diag_kernel(T* __restrict__ d, const T* __restrict__ kappa,
                                                      const DGArgs<T> a) {
  constexpr int L = N1 * N1, NL = L * N1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= a.ncells * NL) return;
  const long long c = gid / NL;
  const int li = (int)(gid % NL);
  const int ii[3] = {li % N1, (li / N1) % N1, li / L};
  const int coord[3] = {(int)(c % a.n[0]), (int)((c / a.n[0]) % a.n[1]), (int)(c / ((long long)a.n[0] * a.n[1]))};
  const T theta = a.theta;
  T acc = T(0);
  for (int D = 0; D < 3; ++D) {
    const T kc = VARK ? kappa[kappa_index(a.aniso, c, D)] : a.kappa0;
    const int i = ii[D];
    const T Sd = a.S[D], Sc = a.Sc[D], bd = a.b[D];
    const T e0 = (i == 0) ? T(1) : T(0), e1 = (i == N1 - 1) ? T(1) : T(0);
    const T e0d0 = a.e0d0[i], e1d1 = a.e1d1[i];
    T dd = Sd * kc * a.Kdiag[i] - Sc * bd * a.Gdiag[i];
    if (coord[D] > 0) {
      T wlo, whi, P;
      if (VARK) face_coeffs<T>(a, kappa[kappa_index(a.aniso, c - a.stride[D], D)], kc, wlo, whi, P);
      else { wlo = a.wI_lo; whi = a.wI_hi; P = a.PI; }
      dd += Sd * (whi * e0d0 + P * e0 - theta * whi * e0d0);
      if (bd <= T(0)) dd += Sc * (-bd) * e0;
    } else {
      const int bct = a.bc[2 * D];
      if (bct == BC_DIRICHLET) {
        const T Pb = VARK ? a.alphak * (a.weightsOn ? kc : T(1)) : a.PB;
        dd += Sd * (kc * e0d0 + Pb * e0 - theta * kc * e0d0);
        if (bd <= T(0)) dd += Sc * (-bd) * e0;
      } else if (bct == BC_OUTFLOW) {
        dd += Sc * (-bd) * e0;
      }
    }
    if (coord[D] < a.n[D] - 1) {
      T wlo, whi, P;
      if (VARK) face_coeffs<T>(a, kc, kappa[kappa_index(a.aniso, c + a.stride[D], D)], wlo, whi, P);
      else { wlo = a.wI_lo; whi = a.wI_hi; P = a.PI; }
      dd += Sd * (-wlo * e1d1 + P * e1 + theta * wlo * e1d1);
      if (bd > T(0)) dd += Sc * bd * e1;
    } else {
      const int bct = a.bc[2 * D + 1];
      if (bct == BC_DIRICHLET) {
        const T Pb = VARK ? a.alphak * (a.weightsOn ? kc : T(1)) : a.PB;
        dd += Sd * (-kc * e1d1 + Pb * e1 + theta * kc * e1d1);
        if (bd >= T(0)) dd += Sc * bd * e1;
      } else if (bct == BC_OUTFLOW) {
        dd += Sc * bd * e1;
      }
    }
    T mt = T(1);
    for (int e = 0; e < 3; ++e)
      if (e != D) mt *= a.Mdiag[ii[e]];
    acc += dd * mt;
  }
  acc += a.V * a.c * a.Mdiag[ii[0]] * a.Mdiag[ii[1]] * a.Mdiag[ii[2]];
  d[gid] += acc;
}

// y += V (M (x) M (x) M) z, the DG mass matrix (ScalarL2, R:dune/pdelab/localoperator/l2.hh:45-83).
template <int N1, typename T>
__global__ __launch_bounds__(256) void dg_mass_kernel(const T* __restrict__ z, T* __restrict__ y, const DGArgs<T> a) {
  constexpr int L = N1 * N1, NL = L * N1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= a.ncells * NL) return;
  const long long c = gid / NL;
  const int li = (int)(gid % NL);
  const int i0 = li % N1, i1 = (li / N1) % N1, i2 = li / L;
  const T* zc = z + c * NL;
  T acc = T(0);
  for (int j2 = 0; j2 < N1; ++j2)
    for (int j1 = 0; j1 < N1; ++j1) {
      T m12 = a.M[i1][j1] * a.M[i2][j2];
      T row = T(0);
      for (int j0 = 0; j0 < N1; ++j0) row += a.M[i0][j0] * zc[j0 + N1 * j1 + L * j2];
      acc += m12 * row;
    }
  y[gid] += a.V * acc;
}

}  // namespace pdelab_amd

#endif
