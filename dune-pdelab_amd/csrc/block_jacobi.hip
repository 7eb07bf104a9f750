// block_jacobi.hip -- set-up and launches of the per-cell block Jacobi preconditioner, and
// the assembled Jacobian in the 7-point block pattern.
#include <cmath>
#include <cstring>

#include "handle.hh"
#include "block_sor.hh"
#include "jacobian.hh"

namespace {

// Cholesky factor of the 1D mass matrix and its inverse, and M^{-1}, from the long
// double tables.
BJMass mass_data(const H* h) {
  using LD = long double;
  const int n = h->n1;
  LD C[4][4] = {}, Ci[4][4] = {};
  for (int j = 0; j < n; ++j) {
    LD s = h->tab.M[j][j];
    for (int k = 0; k < j; ++k) s -= C[j][k] * C[j][k];
    C[j][j] = std::sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      LD t = h->tab.M[i][j];
      for (int k = 0; k < j; ++k) t -= C[i][k] * C[j][k];
      C[i][j] = t / C[j][j];
    }
  }
  for (int j = 0; j < n; ++j) {  // forward substitution C Ci = I, column by column
    for (int i = 0; i < n; ++i) {
      LD s = (i == j) ? 1 : 0;
      for (int k = 0; k < i; ++k) s -= C[i][k] * Ci[k][j];
      Ci[i][j] = s / C[i][i];
    }
  }
  BJMass m;
  std::memset(&m, 0, sizeof(m));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      m.Ci[i][j] = (double)Ci[i][j];
      m.Minv[i][j] = (double)h->tab.Minv[i][j];
    }
  return m;
}

// workgroups of one application (CPG cells each)
template <template <int> class SH>
long long groups_of(const H* h) {
  const int cpg = h->n1 == 2 ? SH<2>::CPG : (h->n1 == 3 ? SH<3>::CPG : SH<4>::CPG);
  return (h->ncells + cpg - 1) / cpg;
}
long long bj_groups(const H* h, bool fdm) { return fdm ? groups_of<BJFdmShape>(h) : groups_of<BJShape>(h); }

// Dense blocks shared by boundary type at Q1 and Q2 run bj_typed_kernel: 64 cells of one type per
// wave. The cells of type t_d in direction d: interior 1 .. n-2, lower {0} and upper {n-1}
// (n >= 2), both {0} (n = 1), as bj_type classifies them.
bool typed_dense(const H* h, const BlockJacobiData& b) { return !b.fdm && !b.per_cell && h->n1 <= 3; }
BJTypes typed_waves(const H* h) {
  BJTypes ty;
  std::memset(&ty, 0, sizeof(ty));
  for (int d = 0; d < 3; ++d) {
    const int n = h->a64.n[d];
    ty.lo[d][0] = 1;
    ty.cnt[d][0] = n >= 3 ? n - 2 : 0;
    ty.lo[d][1] = 0;
    ty.cnt[d][1] = n >= 2 ? 1 : 0;
    ty.lo[d][2] = n - 1;
    ty.cnt[d][2] = n >= 2 ? 1 : 0;
    ty.lo[d][3] = 0;
    ty.cnt[d][3] = n == 1 ? 1 : 0;
  }
  int w = 0;
  for (int t = 0; t < 64; ++t) {
    ty.first[t] = w;
    const long long c = (long long)ty.cnt[0][t & 3] * ty.cnt[1][(t >> 2) & 3] * ty.cnt[2][t >> 4];
    w += (int)((c + 63) / 64);
  }
  ty.first[64] = w;
  return ty;
}

// partials of the reducing modes: one (or one pair) per workgroup
long long bj_partials(const H* h, const BlockJacobiData& b) {
  return typed_dense(h, b) ? typed_waves(h).first[64] : bj_groups(h, b.fdm);
}

BJGeom geom(const H* h, bool per_cell) {
  BJGeom g;
  g.vark = h->vark ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = h->a64.n[d];
    g.stride[d] = h->a64.stride[d];
  }
  g.ncells = h->ncells;
  g.per_cell = per_cell ? 1 : 0;
  return g;
}

template <int N1, bool TRANS, bool VARK>
int fdm_setup(H* h, const BJGeom& g, long long entries, double* o64, float* o32) {
  const int nb = (int)((entries + 255) / 256);
  hipLaunchKernelGGL((bj_fdm_setup_kernel<N1, TRANS, VARK>), dim3(nb), dim3(256), 0, h->stream, h->a64,
                     (const double*)h->d_fc64, g, mass_data(h), entries, o64, o32);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <int N1, bool TRANS, bool CONV, bool VARK>
int dense_setup(H* h, const BJGeom& g, long long entries, int out_mode, double* o64, float* o32) {
  // one workgroup per block; with a velocity field (constant b zero, blocks per cell) the
  // cell blocks add its convection terms
  if constexpr (!CONV) {
    if (h->varb) {
      hipLaunchKernelGGL((bj_dense_setup_kernel<N1, TRANS, false, VARK, true>), dim3((unsigned)entries), dim3(256), 0,
                         h->stream, h->a64, (const double*)h->d_fc64, g, mass_data(h), out_mode, o64, o32,
                         (const VelArgs<double>*)h->d_va64, (const double*)h->d_bv64, (const double*)h->d_bf64);
      HIPCHK(h, hipGetLastError());
      return PDELAB_OK;
    }
  }
  hipLaunchKernelGGL((bj_dense_setup_kernel<N1, TRANS, CONV, VARK>), dim3((unsigned)entries), dim3(256), 0, h->stream,
                     h->a64, (const double*)h->d_fc64, g, mass_data(h), out_mode, o64, o32,
                     (const VelArgs<double>*)nullptr, (const double*)nullptr, (const double*)nullptr);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <int N1, bool TRANS>
int setup_n(H* h, const BlockJacobiData& b, const BJGeom& g, double* o64, float* o32) {
  if (b.fdm) {
    return h->vark ? fdm_setup<N1, TRANS, true>(h, g, b.entries, o64, o32)
                   : fdm_setup<N1, TRANS, false>(h, g, b.entries, o64, o32);
  }
  if (h->conv)
    return h->vark ? dense_setup<N1, TRANS, true, true>(h, g, b.entries, 0, o64, o32)
                   : dense_setup<N1, TRANS, true, false>(h, g, b.entries, 0, o64, o32);
  return h->vark ? dense_setup<N1, TRANS, false, true>(h, g, b.entries, 0, o64, o32)
                 : dense_setup<N1, TRANS, false, false>(h, g, b.entries, 0, o64, o32);
}

template <bool TRANS>
int setup_t(H* h, const BlockJacobiData& b, const BJGeom& g, double* o64, float* o32) {
  switch (h->n1) {
    case 2: return setup_n<2, TRANS>(h, b, g, o64, o32);
    case 3: return setup_n<3, TRANS>(h, b, g, o64, o32);
    default: return setup_n<4, TRANS>(h, b, g, o64, o32);
  }
}

template <typename T>
const T* bj_rowmajor(const BlockJacobiData& b);
template <>
const double* bj_rowmajor<double>(const BlockJacobiData& b) { return b.rm64; }
template <>
const float* bj_rowmajor<float>(const BlockJacobiData& b) { return b.rm32; }

template <int N1, typename T, typename A, int MODE>
int launch_n(H* h, const BlockJacobiData& b, const T* data, const BJVec<T, A>& v) {
  const BJGeom g = geom(h, b.per_cell);
  const long long nb = bj_groups(h, b.fdm);
  if (b.fdm && b.per_cell)
    hipLaunchKernelGGL((bj_fdm_kernel<N1, T, A, MODE, true>), dim3(nb), dim3(BJFdmShape<N1>::THREADS), 0, h->stream, v,
                       data, g);
  else if (b.fdm && ((long long)g.n[0] * g.n[1]) % BJFdmShape<N1>::CPW == 0)
    hipLaunchKernelGGL((bj_fdm_kernel<N1, T, A, MODE, false, true>), dim3(nb), dim3(BJFdmShape<N1>::THREADS), 0,
                       h->stream, v, data, g);
  else if (b.fdm)
    hipLaunchKernelGGL((bj_fdm_kernel<N1, T, A, MODE, false>), dim3(nb), dim3(BJFdmShape<N1>::THREADS), 0, h->stream, v,
                       data, g);
  else if (typed_dense(h, b)) {
    if constexpr (N1 <= 3) {
      const BJTypes ty = typed_waves(h);
      hipLaunchKernelGGL((bj_typed_kernel<N1, T, A, MODE>), dim3((unsigned)ty.first[64]), dim3(64), 0, h->stream, v,
                         bj_rowmajor<T>(b), ty, g);
    }
  } else
    hipLaunchKernelGGL((bj_dense_kernel<N1, T, A, MODE>), dim3(nb), dim3(BJShape<N1>::THREADS), 0, h->stream, v, data,
                       g);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <typename T, typename A, int MODE>
int launch_m(H* h, const BlockJacobiData& b, const T* data, const BJVec<T, A>& v) {
  switch (h->n1) {
    case 2: return launch_n<2, T, A, MODE>(h, b, data, v);
    case 3: return launch_n<3, T, A, MODE>(h, b, data, v);
    default: return launch_n<4, T, A, MODE>(h, b, data, v);
  }
}

template <typename T>
const T* bj_data(const BlockJacobiData& b);
template <>
const double* bj_data<double>(const BlockJacobiData& b) { return b.d64; }
template <>
const float* bj_data<float>(const BlockJacobiData& b) { return b.d32; }

}  // namespace

void bj_invalidate(H* h) {
  for (auto& b : h->bj) b.valid64 = b.valid32 = false;
}

void bj_free(H* h) {
  for (auto& b : h->bj) {
    if (b.d64) (void)hipFree(b.d64);
    if (b.d32) (void)hipFree(b.d32);
    if (b.rm64) (void)hipFree(b.rm64);
    if (b.rm32) (void)hipFree(b.rm32);
    b = BlockJacobiData();
  }
  if (h->d_bjpart) (void)hipFree(h->d_bjpart);
  h->d_bjpart = nullptr;
  if (h->d_ibj_stat) (void)hipFree(h->d_ibj_stat);
  h->d_ibj_stat = nullptr;
  if (h->d_sor_order) (void)hipFree(h->d_sor_order);
  h->d_sor_order = nullptr;
  h->sor_first.clear();
}

// Block data of slot 0 (J) or 1 (J^T), or of slot 2 (J, dense blocks only: block SOR).
static int bj_prepare_slot(H* h, int slot, int precision) {
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "block Jacobi: more than 2^31 - 1 cells");
  BlockJacobiData& b = h->bj[slot];
  const bool trans = slot == 1;
  const bool fdm = slot != 2 && !h->conv && !h->varb && h->cfg.method == PDELAB_SIPG;
  const bool per_cell = h->vark || h->varb;
  const long long nl = h->nloc;
  const long long entries = fdm ? (per_cell ? 3 * h->ncells : 12) : (per_cell ? h->ncells : 64);
  const long long esize = fdm ? (long long)(h->n1 * h->n1 + h->n1) : nl * nl;
  if (b.fdm != fdm || b.per_cell != per_cell || b.entries != entries || b.esize != esize) {
    if (b.d64) (void)hipFree(b.d64);
    if (b.d32) (void)hipFree(b.d32);
    if (b.rm64) (void)hipFree(b.rm64);
    if (b.rm32) (void)hipFree(b.rm32);
    b = BlockJacobiData();
    b.fdm = fdm;
    b.per_cell = per_cell;
    b.entries = entries;
    b.esize = esize;
  }
  if (!h->d_bjpart) {
    // sized for every kernel choice (set_kappa may change the operator later)
    const size_t np = (size_t)std::max(std::max(bj_groups(h, true), bj_groups(h, false)),
                                       (long long)typed_waves(h).first[64]);
    HIPCHK(h, hipMalloc(&h->d_bjpart, sizeof(double) * 2 * np));
  }
  const bool need64 = precision == 64 && !b.valid64;
  const bool need32 = precision == 32 && !b.valid32;
  if (!need64 && !need32) return PDELAB_OK;
  const size_t count = (size_t)(entries * esize);
  if (need64 && !b.d64) HIPCHK(h, hipMalloc(&b.d64, sizeof(double) * count));
  if (need32 && !b.d32) HIPCHK(h, hipMalloc(&b.d32, sizeof(float) * count));
  const BJGeom g = geom(h, per_cell);
  double* o64 = need64 ? b.d64 : nullptr;
  float* o32 = need32 ? b.d32 : nullptr;
  CHK(trans ? setup_t<true>(h, b, g, o64, o32) : setup_t<false>(h, b, g, o64, o32));
  if (typed_dense(h, b)) {
    const int nb = (int)((count + 255) / 256);
    if (need64) {
      if (!b.rm64) HIPCHK(h, hipMalloc(&b.rm64, sizeof(double) * count));
      hipLaunchKernelGGL(bj_rowmajor_kernel<double>, dim3(nb), dim3(256), 0, h->stream, (const double*)b.d64, b.rm64,
                         (int)nl, (long long)count);
    }
    if (need32) {
      if (!b.rm32) HIPCHK(h, hipMalloc(&b.rm32, sizeof(float) * count));
      hipLaunchKernelGGL(bj_rowmajor_kernel<float>, dim3(nb), dim3(256), 0, h->stream, (const float*)b.d32, b.rm32,
                         (int)nl, (long long)count);
    }
    HIPCHK(h, hipGetLastError());
  }
  if (need64) b.valid64 = true;
  if (need32) b.valid32 = true;
  return PDELAB_OK;
}

int bj_prepare(H* h, bool trans, int precision) { return bj_prepare_slot(h, trans ? 1 : 0, precision); }

template <typename T, typename A>
int bj_launch(H* h, bool trans, int mode, BJVec<T, A> v, int* nparts) {
  CHK(bj_prepare(h, trans, sizeof(T) == 8 ? 64 : 32));
  const BlockJacobiData& b = h->bj[trans ? 1 : 0];
  const T* data = bj_data<T>(b);
  v.partial = static_cast<A*>(h->d_bjpart);
  if (nparts) *nparts = (int)bj_partials(h, b);
  if (mode == BJ_PREC) return launch_m<T, A, BJ_PREC>(h, b, data, v);
  if (mode == BJ_ADD) return launch_m<T, A, BJ_ADD>(h, b, data, v);
  return launch_m<T, A, BJ_CGR>(h, b, data, v);
}

template int bj_launch<double, double>(H*, bool, int, BJVec<double, double>, int*);
template int bj_launch<float, float>(H*, bool, int, BJVec<float, float>, int*);

template <int N1, bool TRANS>
static int jac_launch_t(H* h, const BJGeom& g, double* blocks, long long* cols) {
  const dim3 grid((unsigned)h->ncells), block(256);
  if (h->conv) {
    if (h->vark)
      hipLaunchKernelGGL((jacobian_kernel<N1, TRANS, true, true>), grid, block, 0, h->stream, h->a64,
                         (const double*)h->d_fc64, g, blocks, cols);
    else
      hipLaunchKernelGGL((jacobian_kernel<N1, TRANS, true, false>), grid, block, 0, h->stream, h->a64,
                         (const double*)h->d_fc64, g, blocks, cols);
  } else if (h->vark) {
    hipLaunchKernelGGL((jacobian_kernel<N1, TRANS, false, true>), grid, block, 0, h->stream, h->a64,
                       (const double*)h->d_fc64, g, blocks, cols);
  } else {
    hipLaunchKernelGGL((jacobian_kernel<N1, TRANS, false, false>), grid, block, 0, h->stream, h->a64,
                       (const double*)h->d_fc64, g, blocks, cols);
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int jac_assemble(H* h, bool trans, double* blocks, long long* cols) {
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "jacobian: more than 2^31 - 1 cells");
  const BJGeom g = geom(h, true);
  switch (h->n1) {
    case 2: CHK(trans ? jac_launch_t<2, true>(h, g, blocks, cols) : jac_launch_t<2, false>(h, g, blocks, cols)); break;
    case 3: CHK(trans ? jac_launch_t<3, true>(h, g, blocks, cols) : jac_launch_t<3, false>(h, g, blocks, cols)); break;
    default: CHK(trans ? jac_launch_t<4, true>(h, g, blocks, cols) : jac_launch_t<4, false>(h, g, blocks, cols)); break;
  }
  if (h->varb) {  // the velocity field's convection terms on top of the b = 0 blocks
    const VelArgs<double>* va = h->d_va64;
    if (trans)
      hipLaunchKernelGGL(conv_var_jac_kernel<true>, dim3((unsigned)h->ncells), dim3(256), 0, h->stream, va, h->n1,
                         (const double*)h->d_bv64, (const double*)h->d_bf64, blocks, (const long long*)cols);
    else
      hipLaunchKernelGGL(conv_var_jac_kernel<false>, dim3((unsigned)h->ncells), dim3(256), 0, h->stream, va, h->n1,
                         (const double*)h->d_bv64, (const double*)h->d_bf64, blocks, (const long long*)cols);
    HIPCHK(h, hipGetLastError());
  }
  return PDELAB_OK;
}

int jac_block_apply(H* h, const double* blocks, const long long* cols, const double* z, double* y) {
  const dim3 grid((unsigned)h->ncells), block(256);
  switch (h->n1) {
    case 2:
      hipLaunchKernelGGL(block_matrix_apply_kernel<2>, grid, block, 0, h->stream, blocks, cols, z, y, h->ncells);
      break;
    case 3:
      hipLaunchKernelGGL(block_matrix_apply_kernel<3>, grid, block, 0, h->stream, blocks, cols, z, y, h->ncells);
      break;
    default:
      hipLaunchKernelGGL(block_matrix_apply_kernel<4>, grid, block, 0, h->stream, blocks, cols, z, y, h->ncells);
      break;
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

int bj_block_diagonal(H* h, bool trans, double* blocks) {
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "block diagonal: more than 2^31 - 1 cells");
  const BJGeom g = geom(h, true);
  BlockJacobiData b;
  b.entries = h->ncells;
  auto run = [&](auto n1tag) -> int {
    constexpr int N1 = decltype(n1tag)::value;
    if (trans) {
      if (h->conv)
        return h->vark ? dense_setup<N1, true, true, true>(h, g, b.entries, 1, blocks, nullptr)
                       : dense_setup<N1, true, true, false>(h, g, b.entries, 1, blocks, nullptr);
      return h->vark ? dense_setup<N1, true, false, true>(h, g, b.entries, 1, blocks, nullptr)
                     : dense_setup<N1, true, false, false>(h, g, b.entries, 1, blocks, nullptr);
    }
    if (h->conv)
      return h->vark ? dense_setup<N1, false, true, true>(h, g, b.entries, 1, blocks, nullptr)
                     : dense_setup<N1, false, true, false>(h, g, b.entries, 1, blocks, nullptr);
    return h->vark ? dense_setup<N1, false, false, true>(h, g, b.entries, 1, blocks, nullptr)
                   : dense_setup<N1, false, false, false>(h, g, b.entries, 1, blocks, nullptr);
  };
  switch (h->n1) {
    case 2: return run(std::integral_constant<int, 2>());
    case 3: return run(std::integral_constant<int, 3>());
    default: return run(std::integral_constant<int, 4>());
  }
}

// ---------------------------------------------------------------------------
// Block SOR (block_sor.hh): one launch per hyperplane x + y + z = s.

namespace {

template <int N1, typename T, bool CONV, bool VARK, bool INPLACE>
int sor_waves(H* h, const T* d, T* v, const T* dinv, const int* stop) {
  using S = SorShape<N1>;
  const BJGeom g = geom(h, h->bj[2].per_cell);
  const T* fc = VARK ? fc_for<T>(h) : nullptr;
  const DGArgs<T>& a = args_for<T>(h);
  const LineMats<T>* lm = line_mats<T>(h, false);
  const int ns = (int)h->sor_first.size() - 1;
  for (int s = 0; s < ns; ++s) {
    const int first = h->sor_first[s], count = h->sor_first[s + 1] - first;
    const unsigned nb = (unsigned)((count + S::CPW - 1) / S::CPW);
    const T* ds = d;
    if (h->varb) CHK(vel_sor_rhs<T>(h, d, v, first, count, INPLACE, &ds));  // the field's couplings
    hipLaunchKernelGGL((block_sor_wave_kernel<N1, T, CONV, VARK, INPLACE>), dim3(nb), dim3(64), 0, h->stream, ds, v, fc, a, lm,
                       dinv, g, (const int*)h->d_sor_order, first, count, (T)h->sor_omega, stop);
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <int N1, typename T, bool INPLACE>
int sor_i(H* h, const T* d, T* v, const T* dinv, const int* stop) {
  if (h->vark)
    return h->conv ? sor_waves<N1, T, true, true, INPLACE>(h, d, v, dinv, stop)
                   : sor_waves<N1, T, false, true, INPLACE>(h, d, v, dinv, stop);
  return h->conv ? sor_waves<N1, T, true, false, INPLACE>(h, d, v, dinv, stop)
                 : sor_waves<N1, T, false, false, INPLACE>(h, d, v, dinv, stop);
}

template <int N1, typename T>
int sor_n(H* h, const T* d, T* v, const T* dinv, const int* stop, bool inplace) {
  return inplace ? sor_i<N1, T, true>(h, d, v, dinv, stop) : sor_i<N1, T, false>(h, d, v, dinv, stop);
}

}  // namespace

int sor_prepare(H* h, int precision) {
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "block SOR: more than 2^31 - 1 cells");
  CHK(bj_prepare_slot(h, 2, precision));
  if (!h->d_sor_order) {
    // hyperplane s: z ascending, then y ascending (x = s - z - y), i.e. ascending cell index
    const int nx = h->a64.n[0], ny = h->a64.n[1], nz = h->a64.n[2];
    std::vector<int> order;
    order.reserve((size_t)h->ncells);
    h->sor_first.assign(1, 0);
    for (int s = 0; s <= nx + ny + nz - 3; ++s) {
      for (int z = std::max(0, s - (nx - 1) - (ny - 1)); z <= std::min(nz - 1, s); ++z)
        for (int y = std::max(0, s - z - (nx - 1)); y <= std::min(ny - 1, s - z); ++y)
          order.push_back((s - z - y) + nx * (y + ny * z));
      h->sor_first.push_back((int)order.size());
    }
    if ((long long)order.size() != h->ncells) return fail(h, PDELAB_ERR_INVALID_ARG, "block SOR: hyperplane count");
    HIPCHK(h, hipMalloc(&h->d_sor_order, sizeof(int) * order.size()));
    HIPCHK(h, hipMemcpy(h->d_sor_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  }
  return PDELAB_OK;
}

template <typename T>
int sor_launch(H* h, const T* d, T* v, const int* stop, bool inplace) {
  CHK(sor_prepare(h, sizeof(T) == 8 ? 64 : 32));
  const T* dinv = bj_data<T>(h->bj[2]);
  switch (h->n1) {
    case 2: return sor_n<2, T>(h, d, v, dinv, stop, inplace);
    case 3: return sor_n<3, T>(h, d, v, dinv, stop, inplace);
    default: return sor_n<4, T>(h, d, v, dinv, stop, inplace);
  }
}

template int sor_launch<double>(H*, const double*, double*, const int*, bool);
template int sor_launch<float>(H*, const float*, float*, const int*, bool);

// ---------------------------------------------------------------------------
// Iterative block Jacobi (block_jacobi.hh, bj_iter_kernel): the diagonal blocks D_c of J,
// row-major, per boundary type (constant coefficients) or per cell (cell-wise kappa).

int ibj_prepare(H* h, int precision) {
  if (h->ncells >= (1LL << 31)) return fail(h, PDELAB_ERR_UNSUPPORTED, "iterative block Jacobi: more than 2^31 - 1 cells");
  BlockJacobiData& b = h->bj[3];
  const bool per_cell = h->vark || h->varb;
  const long long nl = h->nloc, entries = per_cell ? h->ncells : 64;
  if (b.per_cell != per_cell || b.entries != entries || b.esize != nl * nl) {
    if (b.d64) (void)hipFree(b.d64);
    if (b.d32) (void)hipFree(b.d32);
    b = BlockJacobiData();
    b.per_cell = per_cell;
    b.entries = entries;
    b.esize = nl * nl;
  }
  if (!h->d_ibj_stat) HIPCHK(h, hipMalloc(&h->d_ibj_stat, 16));
  const bool need64 = precision == 64 && !b.valid64;
  const bool need32 = precision == 32 && !b.valid32;
  if (!need64 && !need32) return PDELAB_OK;
  const size_t count = (size_t)(entries * nl * nl);
  if (need64 && !b.d64) HIPCHK(h, hipMalloc(&b.d64, sizeof(double) * count));
  if (need32 && !b.d32) HIPCHK(h, hipMalloc(&b.d32, sizeof(float) * count));
  const BJGeom g = geom(h, per_cell);
  double* o64 = need64 ? b.d64 : nullptr;
  float* o32 = need32 ? b.d32 : nullptr;
  auto run = [&](auto n1tag) -> int {
    constexpr int N1 = decltype(n1tag)::value;
    if (h->conv)
      return h->vark ? dense_setup<N1, false, true, true>(h, g, entries, 1, o64, o32)
                     : dense_setup<N1, false, true, false>(h, g, entries, 1, o64, o32);
    return h->vark ? dense_setup<N1, false, false, true>(h, g, entries, 1, o64, o32)
                   : dense_setup<N1, false, false, false>(h, g, entries, 1, o64, o32);
  };
  switch (h->n1) {
    case 2: CHK(run(std::integral_constant<int, 2>())); break;
    case 3: CHK(run(std::integral_constant<int, 3>())); break;
    default: CHK(run(std::integral_constant<int, 4>())); break;
  }
  if (need64) b.valid64 = true;
  if (need32) b.valid32 = true;
  return PDELAB_OK;
}

template <int N1, typename T>
static void ibj_kernel(H* h, const T* d, T* v, const T* D, const BJGeom& g, int* err, unsigned long long* iters,
                       const int* stop, bool add) {
  const dim3 grid((unsigned)h->ncells), block(64);
  if (add)
    hipLaunchKernelGGL((bj_iter_kernel<N1, T, true>), grid, block, 0, h->stream, d, v, D, g, h->ibj_opt, err, iters, stop);
  else
    hipLaunchKernelGGL((bj_iter_kernel<N1, T, false>), grid, block, 0, h->stream, d, v, D, g, h->ibj_opt, err, iters, stop);
}

template <typename T>
int ibj_launch(H* h, const T* d, T* v, int* err, unsigned long long* iters, const int* stop, bool add) {
  CHK(ibj_prepare(h, sizeof(T) == 8 ? 64 : 32));
  const BlockJacobiData& b = h->bj[3];
  const BJGeom g = geom(h, b.per_cell);
  const T* D = bj_data<T>(b);
  switch (h->n1) {
    case 2: ibj_kernel<2, T>(h, d, v, D, g, err, iters, stop, add); break;
    case 3: ibj_kernel<3, T>(h, d, v, D, g, err, iters, stop, add); break;
    default: ibj_kernel<4, T>(h, d, v, D, g, err, iters, stop, add); break;
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template int ibj_launch<double>(H*, const double*, double*, int*, unsigned long long*, const int*, bool);
template int ibj_launch<float>(H*, const float*, float*, int*, unsigned long long*, const int*, bool);
