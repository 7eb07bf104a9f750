// blas_kernels.hh -- vector backend kernels (ISTL::BlockVector dot / two_norm / axpy,
// R:dune/pdelab/backend/istl/vector.hh:249-282) and the fused Krylov updates.
//
// Reductions are deterministic: a fixed grid of kRedBlocks blocks walks the vector
// with a fixed stride, each block reduces in a fixed tree (wave64 shuffles, then LDS)
// and one block sums the per-block partials in order. Results are therefore
// bitwise reproducible run to run.
#ifndef PDELAB_BLAS_KERNELS_HH
#define PDELAB_BLAS_KERNELS_HH

#include <hip/hip_runtime.h>

#include "block_reduce.hh"
#include "stream_io.hh"

namespace pdelab_amd {

// Block-wide sum of NR accumulators; result valid in thread 0.
template <typename A, int NR>
__device__ __forceinline__ void block_sum(A (&v)[NR]) {
  __shared__ A red[NR][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    v[j] = wave_sum(v[j]);
    if (lane == 0) red[j][wid] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      A s = red[j][0];
      for (int w = 1; w < kRedThreads / 64; ++w) s += red[j][w];
      v[j] = s;
    }
  }
}

// partial[b*NR + j] = block b's share of sum_i f_j(i)
template <typename T, typename A>
__global__ __launch_bounds__(kRedThreads) void dot_partial_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                                  long long n, A* __restrict__ partial) {
  A v[1] = {A(0)};
  for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)kRedBlocks * kRedThreads)
    v[0] += (A)x[i] * (A)y[i];
  block_sum<A, 1>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// Single-workgroup sums of per-workgroup partials (up to tens of thousands of them):
// kFinalThreads threads, each with four independent accumulators so that four loads
// per thread are in flight, then the workgroup tree. The order is fixed.
constexpr int kFinalThreads = 1024;

template <typename A, int NR>
__device__ __forceinline__ void strided_partial_sum(const A* __restrict__ partial, int nb, A (&v)[NR]) {
  A a[4][NR];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int j = 0; j < NR; ++j) a[u][j] = A(0);
  int b = threadIdx.x;
  for (; b + 3 * kFinalThreads < nb; b += 4 * kFinalThreads)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < NR; ++j) a[u][j] += partial[(b + u * kFinalThreads) * NR + j];
  for (; b < nb; b += kFinalThreads)
#pragma unroll
    for (int j = 0; j < NR; ++j) a[0][j] += partial[b * NR + j];
#pragma unroll
  for (int j = 0; j < NR; ++j) v[j] = (a[0][j] + a[1][j]) + (a[2][j] + a[3][j]);
  block_sum_n<kFinalThreads, A, NR>(v);
}

// out[j] = sum_b partial[b*NR + j]   (one workgroup of kFinalThreads threads)
template <typename A, int NR>
__global__ __launch_bounds__(kFinalThreads) void final_reduce_kernel(const A* __restrict__ partial, int nb,
                                                                     A* __restrict__ out,
                                                                     const int* __restrict__ stop) {
  if (stop && *stop) return;  // a device-side Krylov loop has already stopped
  A v[NR];
  strided_partial_sum<A, NR>(partial, nb, v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NR; ++j) out[j] = v[j];
}

template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(long long n, T a, const T* __restrict__ x, T* __restrict__ y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(long long n, T a, T* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] *= a;
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(long long n, T a, T* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = a;
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(long long n, const S* __restrict__ x, D* __restrict__ y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = (D)x[i];
}

// x -= (double) z32   (refinement update)
__global__ __launch_bounds__(256) void sub_f32_kernel(long long n, const float* __restrict__ z, double* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] -= (double)z[i];
}

// y32 = (float)(s x64): the refinement residual handed to the FP32 solve at unit norm
__global__ __launch_bounds__(256) void scale_to_f32_kernel(long long n, double s, const double* __restrict__ x,
                                                           float* __restrict__ y) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = (float)(s * x[i]);
}

// x -= s (double) z32   (refinement update of a correction computed at scale 1/s)
__global__ __launch_bounds__(256) void sub_scaled_f32_kernel(long long n, double s, const float* __restrict__ z,
                                                             double* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] -= s * (double)z[i];
}

// kappa = exp(log kappa) in FP64 and FP32
__global__ __launch_bounds__(256) void exp_kernel(long long n, const double* __restrict__ logk, double* __restrict__ k64,
                                                  float* __restrict__ k32) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = exp(logk[i]);
    k64[i] = v;
    k32[i] = (float)v;
  }
}

// dinv = 1 / d
template <typename T>
__global__ __launch_bounds__(256) void invert_kernel(long long n, const double* __restrict__ d, T* __restrict__ dinv) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dinv[i] = (T)(1.0 / d[i]);
}

// y = dinv .* x  (SeqJac with one step and relaxation 1 from a zero start, or the
// identity when dinv == nullptr); also returns partials of y . x
template <typename T, typename A>
__global__ __launch_bounds__(kRedThreads) void prec_dot_kernel(long long n, const T* __restrict__ dinv,
                                                               const T* __restrict__ x, T* __restrict__ y,
                                                               A* __restrict__ partial) {
  A v[1] = {A(0)};
  for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)kRedBlocks * kRedThreads) {
    const T xi = x[i];
    const T yi = dinv ? dinv[i] * xi : xi;
    y[i] = yi;
    v[0] += (A)yi * (A)xi;
  }
  block_sum<A, 1>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// Device-side CG iteration (dune-istl CGSolver::apply, R:dune/istl/solvers.hh), five
// launches per iteration: the apply with the fused p.q partials, their reduction into
// scal[0] (cg_pq_kernel), cg_rupdate_kernel, cg_check_kernel and cg_xpupdate_kernel.
// state = {stopped, aborted, stop iteration}; dstate = {defect, def0, reduction}. Every
// kernel of an iteration after the one that stopped the loop returns at once.
constexpr int kVecUnroll = 4;
// Scalar slots of the step lengths: the single-workgroup kernels divide once, and the
// vector kernels read the quotients (an FP64 division per wave cost ~10 us per launch).
constexpr int kSlotLambda = 8, kSlotBeta = 9;

// scal[0] = p.q from the apply's partials, scal[kSlotLambda] = lambda = scal[slot_rho] / p.q
template <typename A>
__global__ __launch_bounds__(kFinalThreads) void cg_pq_kernel(const A* __restrict__ partial, int nb,
                                                            A* __restrict__ scal, int slot_rho,
                                                            const int* __restrict__ stop) {
  if (*stop) return;
  A v[1];
  strided_partial_sum<A, 1>(partial, nb, v);
  if (threadIdx.x == 0) {
    scal[0] = v[0];
    scal[kSlotLambda] = scal[slot_rho] / v[0];
  }
}

// r -= lambda q with lambda = scal[kSlotLambda]; partials of r.r and (dinv r).r
template <typename T, typename A>
__global__ __launch_bounds__(kRedThreads) void cg_rupdate_kernel(long long n, T* __restrict__ r, const T* __restrict__ q,
                                                                 const T* __restrict__ dinv, const A* __restrict__ scal,
                                                                 A* __restrict__ partial,
                                                                 const int* __restrict__ state) {
  if (state[0]) return;
  const T lambda = (T)scal[kSlotLambda];
  A v[2] = {A(0), A(0)};
  const long long S = (long long)gridDim.x * kRedThreads;
  for (long long i0 = (long long)blockIdx.x * kRedThreads + threadIdx.x; i0 < n; i0 += kVecUnroll * S) {
    T rv[kVecUnroll], qv[kVecUnroll], dv[kVecUnroll];
#pragma unroll
    for (int u = 0; u < kVecUnroll; ++u) {
      const long long i = i0 + u * S;
      rv[u] = qv[u] = T(0);
      dv[u] = T(1);
      if (i < n) {  // r and q stream in; r is stored cached for the x/p update that follows
        rv[u] = ld_stream<T>(r + i);
        qv[u] = ld_stream<T>(q + i);
        if (dinv) dv[u] = dinv[i];
      }
    }
#pragma unroll
    for (int u = 0; u < kVecUnroll; ++u) {
      const long long i = i0 + u * S;
      if (i < n) {
        const T ri = rv[u] - lambda * qv[u];
        r[i] = ri;
        const T zi = dv[u] * ri;
        v[0] += (A)ri * (A)ri;
        v[1] += (A)zi * (A)ri;
      }
    }
  }
  block_sum_n<kRedThreads, A, 2>(v);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = v[0];
    partial[2 * blockIdx.x + 1] = v[1];
  }
}

// IterativeSolver::Iteration::step on the device (single workgroup): reduces the
// partials of (r.r, z.r) into out[0..1], writes beta = z.r / rho_old (when beta is not
// null), records the defect and the iteration number of a running loop and raises the
// stop flag on convergence (def < reduction*def0 or def < 1e-30) or on a non-finite defect.
template <typename A>
__global__ __launch_bounds__(kFinalThreads) void cg_check_kernel(const A* __restrict__ partial, int nb,
                                                               A* __restrict__ out, int* __restrict__ state,
                                                               double* __restrict__ dstate, int it,
                                                               const A* __restrict__ rho_old = nullptr,
                                                               A* __restrict__ beta = nullptr) {
  if (state[0]) return;
  A v[2];
  strided_partial_sum<A, 2>(partial, nb, v);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
    if (beta) *beta = v[1] / *rho_old;
    const double def = sqrt((double)v[0]);
    state[2] = it;
    dstate[0] = def;
    if (!isfinite(def)) {
      state[1] = 1;
      state[0] = 1;
    } else if (def < dstate[1] * dstate[2] || def < 1e-30) {
      state[0] = 1;
    }
  }
}

// x += lambda p (lambda = scal[kSlotLambda]), then p = z + beta p with
// beta = scal[kSlotBeta] and z = W^{-1} r: the given z (block Jacobi), else
// dinv r or r; on the iteration that stopped the loop only x is updated (dune-istl leaves
// the loop after the defect check).
// Elements per thread of the x/p update: one double, or two contiguous floats (8 B per
// lane and stream either way).
template <typename T>
constexpr int xp_ept() { return sizeof(T) == 4 ? 2 : 1; }

// Source of z: the given vector (block Jacobi), dinv r (point Jacobi) or r (none), fixed at
// compile time (alone at 2^24 doubles the run-time choice cost 104.5 against 96.8 us; inside
// the CG loop the two are within noise).
enum { XP_Z = 0, XP_DINV_R = 1, XP_R = 2 };

template <typename T, typename A, int ZSRC>
__global__ __launch_bounds__(256) void cg_xpupdate_kernel(long long n, T* __restrict__ x, T* __restrict__ p,
                                                          const T* __restrict__ r, const T* __restrict__ dinv,
                                                          const T* __restrict__ z, const A* __restrict__ scal,
                                                          const int* __restrict__ state, int it) {
  // xp_ept<T>() contiguous elements per thread over an exact grid: the fastest of the
  // access patterns measured for this five-stream update
  constexpr int E = xp_ept<T>();
  if (state[0] && state[2] < it) return;
  const long long i0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * E;
  if (i0 >= n) return;
  const bool pupd = state[0] == 0;
  const T lambda = (T)scal[kSlotLambda];
  const T beta = pupd ? (T)scal[kSlotBeta] : T(0);
  // x and z (or r) stream (non-temporal); p, reread by the next apply, stays cached
  if constexpr (E == 2) {
    if (i0 + 1 < n) {  // two contiguous floats as one 8-byte access per stream
      typedef T v2 __attribute__((ext_vector_type(2)));
      const v2 pv = *reinterpret_cast<const v2*>(p + i0);
      v2* xv = reinterpret_cast<v2*>(x + i0);
      st_stream<v2>(ld_stream<v2>(xv) + lambda * pv, xv);
      if (pupd) {
        v2 zv;
        if constexpr (ZSRC == XP_Z) zv = ld_stream<v2>(reinterpret_cast<const v2*>(z + i0));
        else if constexpr (ZSRC == XP_DINV_R)
          zv = *reinterpret_cast<const v2*>(dinv + i0) * ld_stream<v2>(reinterpret_cast<const v2*>(r + i0));
        else zv = ld_stream<v2>(reinterpret_cast<const v2*>(r + i0));
        *reinterpret_cast<v2*>(p + i0) = zv + beta * pv;
      }
      return;
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const long long i = i0 + e;
    if (i < n) {
      const T pv = p[i];
      st_stream<T>(ld_stream<T>(x + i) + lambda * pv, x + i);
      if (pupd) {
        T zv;
        if constexpr (ZSRC == XP_Z) zv = ld_stream<T>(z + i);
        else if constexpr (ZSRC == XP_DINV_R) zv = dinv[i] * ld_stream<T>(r + i);
        else zv = ld_stream<T>(r + i);
        p[i] = zv + beta * pv;
      }
    }
  }
}

// CG direction update: p = dinv r + beta p
template <typename T>
__global__ __launch_bounds__(256) void cg_pupdate_kernel(long long n, T* __restrict__ p, const T* __restrict__ r,
                                                         const T* __restrict__ dinv, T beta) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const T zi = dinv ? dinv[i] * r[i] : r[i];
    p[i] = zi + beta * p[i];
  }
}

// BiCGSTAB: p = r + beta (p - omega v)   (or p = r when first)
template <typename T>
__global__ __launch_bounds__(256) void bicg_p_kernel(long long n, T* __restrict__ p, const T* __restrict__ r,
                                                     const T* __restrict__ v, T beta, T omega, int first) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = first ? r[i] : beta * (p[i] - omega * v[i]) + r[i];
}

// x += a y; r -= a v; partials of r.r
template <typename T, typename A>
__global__ __launch_bounds__(kRedThreads) void xr_update_kernel(long long n, T* __restrict__ x, const T* __restrict__ y,
                                                                T* __restrict__ r, const T* __restrict__ v, T a,
                                                                A* __restrict__ partial) {
  A s[1] = {A(0)};
  for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)kRedBlocks * kRedThreads) {
    x[i] += a * y[i];
    const T ri = r[i] - a * v[i];
    r[i] = ri;
    s[0] += (A)ri * (A)ri;
  }
  block_sum<A, 1>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// partials of t.r and t.t
template <typename T, typename A>
__global__ __launch_bounds__(kRedThreads) void dot2_kernel(long long n, const T* __restrict__ t, const T* __restrict__ r,
                                                           A* __restrict__ partial) {
  A s[2] = {A(0), A(0)};
  for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)kRedBlocks * kRedThreads) {
    const A ti = (A)t[i];
    s[0] += ti * (A)r[i];
    s[1] += ti * ti;
  }
  block_sum<A, 2>(s);
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = s[0]; partial[2 * blockIdx.x + 1] = s[1]; }
}

// ---------------------------------------------------------------------------
// Device-side BiCGSTAB (dune-istl BiCGSTABSolver::apply with its half-step Iteration
// counting). Per iteration: p update, W^{-1} p, the apply, rt.v, x/r update with the r.r
// partials, the half-step check, W^{-1} r, the apply, (t.r, t.t), x/r update with the
// (r.r, rt.r) partials and the full-step check; single-workgroup kernels keep the solver's
// scalars in FP64 as the host loop kept them:
//   bs = {rho, rho_new, alpha, omega, beta}; state = {stopped, abort code, half steps};
//   dstate = {defect, def0, reduction}.
// Once state[0] is set every later kernel that writes x, r or the scalars returns at once,
// so x and r are those of the half step at which the loop stopped.
enum { BICG_RUNNING = 0, BICG_NAN = 1, BICG_RHO = 2, BICG_OMEGA = 3, BICG_H = 4 };
constexpr int kBsRho = 0, kBsRhoNew = 1, kBsAlpha = 2, kBsOmega = 3, kBsBeta = 4;
constexpr double kBicgEps = 1e-80;  // dune-istl's EPSILON of BiCGSTABSolver

// Iteration::step(half / 2, norm) on the device; true when the loop stops here.
__device__ __forceinline__ bool bicg_step(int* state, double* dstate, int half, double norm) {
  state[2] = half;
  dstate[0] = norm;
  if (!isfinite(norm)) {
    state[1] = BICG_NAN;
    state[0] = 1;
    return true;
  }
  if (norm < dstate[1] * dstate[2] || norm < 1e-30) {
    state[0] = 1;
    return true;
  }
  return false;
}

// bs at the start: rho = alpha = omega = 1, rho_new = rt.r (= r.r, rt being r), beta = 0
__global__ __launch_bounds__(64) void bicg_init_kernel(double* __restrict__ bs, double rr) {
  if (threadIdx.x == 0) {
    bs[kBsRho] = 1.0;
    bs[kBsRhoNew] = rr;
    bs[kBsAlpha] = 1.0;
    bs[kBsOmega] = 1.0;
    bs[kBsBeta] = 0.0;
  }
}

// p = r (first iteration) or p = r + beta (p - omega v)
template <typename T>
__global__ __launch_bounds__(256) void bicg_p_dev_kernel(long long n, T* __restrict__ p, const T* __restrict__ r,
                                                         const T* __restrict__ v, const double* __restrict__ bs,
                                                         const int* __restrict__ state, int first) {
  if (state[0]) return;
  const T beta = (T)bs[kBsBeta], omega = (T)bs[kBsOmega];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = first ? r[i] : beta * (p[i] - omega * v[i]) + r[i];
}

// y = dinv .* x (or y = x), skipped once the loop has stopped
template <typename T>
__global__ __launch_bounds__(256) void prec_stop_kernel(long long n, const T* __restrict__ dinv, const T* __restrict__ x,
                                                        T* __restrict__ y, const int* __restrict__ state) {
  if (state[0]) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = dinv ? dinv[i] * x[i] : x[i];
}

// x += a y; r -= a v with a = bs[slot]; partials of r.r (NR = 1) or of (r.r, rt.r) (NR = 2),
// in the order of xr_update_kernel and dot_partial_kernel
template <typename T, typename A, int NR>
__global__ __launch_bounds__(kRedThreads) void bicg_xr_kernel(long long n, T* __restrict__ x, const T* __restrict__ y,
                                                              T* __restrict__ r, const T* __restrict__ v,
                                                              const T* __restrict__ rt, const double* __restrict__ bs,
                                                              int slot, A* __restrict__ partial,
                                                              const int* __restrict__ state) {
  if (state[0]) return;
  const T a = (T)bs[slot];
  A s[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) s[j] = A(0);
  for (long long i = (long long)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (long long)kRedBlocks * kRedThreads) {
    x[i] += a * y[i];
    const T ri = r[i] - a * v[i];
    r[i] = ri;
    s[0] += (A)ri * (A)ri;
    if constexpr (NR == 2) s[1] += (A)rt[i] * (A)ri;
  }
  block_sum<A, NR>(s);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NR; ++j) partial[NR * blockIdx.x + j] = s[j];
}

// The scalar steps, one workgroup reducing the partials first:
//   PHASE 0: h = rt.v; |h| < EPSILON aborts; alpha = rho_new / h (0 when the defect is 0)
//   PHASE 1: defect of the half step, Iteration::step
//   PHASE 2: omega = t.r / t.t (0 when the defect is 0)
//   PHASE 3: defect of the full step, rho = rho_new, rho_new = rt.r, Iteration::step; then
//            the breakdown checks at the top of dune-istl's next iteration (|rho|, |omega|
//            <= EPSILON) and beta = (rho_new / rho)(alpha / omega)
template <typename A, int PHASE>
__global__ __launch_bounds__(kFinalThreads) void bicg_scalar_kernel(const A* __restrict__ partial, int nb,
                                                                    double* __restrict__ bs, int* __restrict__ state,
                                                                    double* __restrict__ dstate, int half) {
  if (state[0]) return;
  constexpr int NR = (PHASE == 2 || PHASE == 3) ? 2 : 1;
  A v[NR];
  strided_partial_sum<A, NR>(partial, nb, v);
  if (threadIdx.x != 0) return;
  const double norm = dstate[0];
  if constexpr (PHASE == 0) {
    const double hh = (double)v[0];
    if (fabs(hh) < kBicgEps) {
      state[1] = BICG_H;
      state[0] = 1;
      return;
    }
    bs[kBsAlpha] = norm == 0.0 ? 0.0 : bs[kBsRhoNew] / hh;
  } else if constexpr (PHASE == 1) {
    bicg_step(state, dstate, half, sqrt((double)v[0]));
  } else if constexpr (PHASE == 2) {
    bs[kBsOmega] = norm == 0.0 ? 0.0 : (double)v[0] / (double)v[1];
  } else {
    const double nrm = sqrt((double)v[0]);
    bs[kBsRho] = bs[kBsRhoNew];
    bs[kBsRhoNew] = (double)v[1];
    if (bicg_step(state, dstate, half, nrm)) return;
    if (fabs(bs[kBsRho]) <= kBicgEps) {
      state[1] = BICG_RHO;
      state[0] = 1;
      return;
    }
    if (fabs(bs[kBsOmega]) <= kBicgEps) {
      state[1] = BICG_OMEGA;
      state[0] = 1;
      return;
    }
    bs[kBsBeta] = nrm == 0.0 ? 0.0 : (bs[kBsRhoNew] / bs[kBsRho]) * (bs[kBsAlpha] / bs[kBsOmega]);
  }
}

}  // namespace pdelab_amd

#endif
