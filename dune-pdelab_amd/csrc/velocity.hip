// velocity.hip -- set-up and launches of the convection terms of a velocity field b(x)
// tabulated at the quadrature points (kernels in dg_velocity.hh).
#include <cstring>
#include <vector>

#include "handle.hh"

template <typename T>
static void fill_vel_args(const H* h, VelArgs<T>& a) {
  std::memset(&a, 0, sizeof(a));
  const Tables1D& t = h->tab;
  const pdelab_config& c = h->cfg;
  for (int d = 0; d < 3; ++d) {
    a.n[d] = h->a64.n[d];
    a.stride[d] = h->a64.stride[d];
  }
  a.ncells = h->ncells;
  for (int q = 0; q < t.m; ++q) {
    a.w[q] = (T)t.qw[q];
    for (int i = 0; i < t.n1; ++i) {
      a.B[q][i] = (T)t.P[q][i];
      a.Dq[q][i] = (T)lagrange_dp(t.k, i, t.qx[q]);
    }
  }
  const double hh[3] = {c.lx / c.nx, c.ly / c.ny, c.lz / c.nz};
  const double V = hh[0] * hh[1] * hh[2];
  for (int d = 0; d < 3; ++d) {
    a.vs[d] = (T)(-V / hh[d]);
    a.area[d] = (T)(V / hh[d]);
  }
  for (int s = 0; s < 6; ++s) a.bc[s] = h->a64.bc[s];
  a.m = t.m;
}

void vel_free(H* h) {
  void* p[] = {h->d_bv64, h->d_bf64, h->d_bv32, h->d_bf32, h->d_va64, h->d_sortmp64, h->d_sortmp32};
  for (void* q : p)
    if (q) (void)hipFree(q);
  h->d_bv64 = h->d_bf64 = nullptr;
  h->d_bv32 = h->d_bf32 = nullptr;
  h->d_va64 = nullptr;
  h->d_sortmp64 = nullptr;
  h->d_sortmp32 = nullptr;
  h->varb = false;
}

// The quadrature of the kernels: M = N1 (intorderadd 0, 1) or N1 + 1 (intorderadd 2, 3).
static bool vel_supported(const H* h) { return h->tab.m == h->n1 || h->tab.m == h->n1 + 1; }

int vel_set(H* h, const double* bvol, const double* bface) {
  HIPCHK(h, hipStreamSynchronize(h->stream));  // no launch in flight reads the old field
  if (!bvol) {
    vel_free(h);
    h->R0_valid = false;
    h->diag_valid = false;
    bj_invalidate(h);
    return PDELAB_OK;
  }
  if (!bface) return fail(h, PDELAB_ERR_INVALID_ARG, "velocity field: b_face is required with b_vol");
  const pdelab_config& c = h->cfg;
  if (c.b[0] != 0.0 || c.b[1] != 0.0 || c.b[2] != 0.0)
    return fail(h, PDELAB_ERR_INVALID_ARG, "velocity field: the configuration's constant b must be zero");
  if (!vel_supported(h))
    return fail(h, PDELAB_ERR_UNSUPPORTED, "velocity field: intorderadd 0..3 only");
  const long long m = h->tab.m, m2 = m * m, m3 = m2 * m;
  const int n[3] = {c.nx, c.ny, c.nz};
  // faces normal to d: positions (n + e_d), lexicographic, x fastest
  long long fdir[4] = {0, 0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    long long nf = 1;
    for (int e = 0; e < 3; ++e) nf *= n[e] + (e == d ? 1 : 0);
    fdir[d + 1] = fdir[d] + nf;
  }
  auto face_of = [&](const int pos[3], int d) {
    const long long e0 = n[0] + (d == 0), e1 = n[1] + (d == 1);
    return fdir[d] + pos[0] + e0 * (pos[1] + e1 * (long long)pos[2]);
  };
  // outflow sides take no inflow (convectiondiffusiondg.hh:800-805)
  for (int s = 0; s < 6; ++s) {
    if (h->a64.bc[s] != BC_OUTFLOW) continue;
    const int d = s / 2, up = s % 2, t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const double ns = up ? 1.0 : -1.0;
    for (int j2 = 0; j2 < n[t2]; ++j2)
      for (int j1 = 0; j1 < n[t1]; ++j1) {
        int pos[3];
        pos[d] = up ? n[d] : 0;
        pos[t1] = j1;
        pos[t2] = j2;
        const double* v = bface + face_of(pos, d) * m2;
        for (long long q = 0; q < m2; ++q)
          if (ns * v[q] < -1e-30) return fail(h, PDELAB_ERR_OUTFLOW_ON_INFLOW, "Outflow boundary condition on inflow!");
      }
  }
  const long long nc = h->ncells;
  std::vector<double> bv((size_t)(nc * 3 * m3)), bf((size_t)(nc * 6 * m2));
  for (long long cc = 0; cc < nc; ++cc) {
    for (int d = 0; d < 3; ++d)
      for (long long q = 0; q < m3; ++q) bv[(size_t)(cc * 3 * m3 + d * m3 + q)] = bvol[(size_t)((cc * m3 + q) * 3 + d)];
    const int coord[3] = {(int)(cc % n[0]), (int)((cc / n[0]) % n[1]), (int)(cc / ((long long)n[0] * n[1]))};
    for (int f = 0; f < 6; ++f) {
      const int d = f / 2;
      int pos[3] = {coord[0], coord[1], coord[2]};
      pos[d] += f % 2;
      std::memcpy(&bf[(size_t)((cc * 6 + f) * m2)], bface + face_of(pos, d) * m2, sizeof(double) * m2);
    }
  }
  std::vector<float> bv32(bv.begin(), bv.end()), bf32(bf.begin(), bf.end());
  vel_free(h);
  HIPCHK(h, hipMalloc(&h->d_bv64, sizeof(double) * bv.size()));
  HIPCHK(h, hipMalloc(&h->d_bf64, sizeof(double) * bf.size()));
  HIPCHK(h, hipMalloc(&h->d_bv32, sizeof(float) * bv.size()));
  HIPCHK(h, hipMalloc(&h->d_bf32, sizeof(float) * bf.size()));
  HIPCHK(h, hipMemcpy(h->d_bv64, bv.data(), sizeof(double) * bv.size(), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_bf64, bf.data(), sizeof(double) * bf.size(), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_bv32, bv32.data(), sizeof(float) * bv.size(), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_bf32, bf32.data(), sizeof(float) * bf.size(), hipMemcpyHostToDevice));
  fill_vel_args<double>(h, h->va64);
  fill_vel_args<float>(h, h->va32);
  HIPCHK(h, hipMalloc(&h->d_va64, sizeof(VelArgs<double>)));
  HIPCHK(h, hipMemcpy(h->d_va64, &h->va64, sizeof(VelArgs<double>), hipMemcpyHostToDevice));
  h->varb = true;
  h->R0_valid = false;
  h->diag_valid = false;
  bj_invalidate(h);  // the cell blocks take the field
  return PDELAB_OK;
}

template <typename T>
static const T* bv_for(const H* h);
template <>
const double* bv_for<double>(const H* h) { return h->d_bv64; }
template <>
const float* bv_for<float>(const H* h) { return h->d_bv32; }
template <typename T>
static const T* bf_for(const H* h);
template <>
const double* bf_for<double>(const H* h) { return h->d_bf64; }
template <>
const float* bf_for<float>(const H* h) { return h->d_bf32; }
template <typename T>
static const VelArgs<T>& va_for(const H* h);
template <>
const VelArgs<double>& va_for<double>(const H* h) { return h->va64; }
template <>
const VelArgs<float>& va_for<float>(const H* h) { return h->va32; }

// Block SOR with a field: dtmp = d less the convection couplings of hyperplane
// order[first .. first + count) (conv_var_sor_kernel); returns dtmp for the sweep kernel.
template <int N1, int M, typename T, bool INPLACE>
static void vel_sor_nm(H* h, const T* d, const T* v, T* dtmp, int first, int count) {
  const long long n = (long long)count * N1 * N1 * N1;
  hipLaunchKernelGGL((conv_var_sor_kernel<N1, M, T, INPLACE>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     h->stream, d, v, dtmp, bf_for<T>(h), va_for<T>(h), (const int*)h->d_sor_order, first, count);
}

template <typename T>
static T*& sortmp_for(H* h);
template <>
double*& sortmp_for<double>(H* h) { return h->d_sortmp64; }
template <>
float*& sortmp_for<float>(H* h) { return h->d_sortmp32; }

template <typename T>
int vel_sor_rhs(H* h, const T* d, const T* v, int first, int count, bool inplace, const T** out) {
  T*& tmp = sortmp_for<T>(h);
  if (!tmp) HIPCHK(h, hipMalloc(&tmp, sizeof(T) * h->N));
  const bool m1 = h->tab.m == h->n1;
#define VSOR(N1_, M_)                                                                  \
  (inplace ? vel_sor_nm<N1_, M_, T, true>(h, d, v, tmp, first, count)                  \
           : vel_sor_nm<N1_, M_, T, false>(h, d, v, tmp, first, count))
  switch (h->n1) {
    case 2: m1 ? VSOR(2, 2) : VSOR(2, 3); break;
    case 3: m1 ? VSOR(3, 3) : VSOR(3, 4); break;
    case 4: m1 ? VSOR(4, 4) : VSOR(4, 5); break;
    default: return fail(h, PDELAB_ERR_UNSUPPORTED, "degree");
  }
#undef VSOR
  HIPCHK(h, hipGetLastError());
  *out = tmp;
  return PDELAB_OK;
}
template int vel_sor_rhs<double>(H*, const double*, const double*, int, int, bool, const double**);
template int vel_sor_rhs<float>(H*, const float*, const float*, int, int, bool, const float**);

// N1 points per direction: conv_var_sq_kernel (per-lane items and tables in registers);
// N1 + 1: the general conv_var_kernel.
template <int N1, int M, typename T>
static void vel_launch_nm(H* h, const T* z, T* y, const int* stop, bool trans) {
  const dim3 grid((unsigned)h->ncells), block(64);
  if constexpr (M == N1) {
    if (trans)
      hipLaunchKernelGGL((conv_var_sq_kernel<N1, T, true>), grid, block, 0, h->stream, z, y, bv_for<T>(h), bf_for<T>(h),
                         va_for<T>(h), stop);
    else
      hipLaunchKernelGGL((conv_var_sq_kernel<N1, T, false>), grid, block, 0, h->stream, z, y, bv_for<T>(h), bf_for<T>(h),
                         va_for<T>(h), stop);
  } else {
    if (trans)
      hipLaunchKernelGGL((conv_var_kernel<N1, M, T, true>), grid, block, 0, h->stream, z, y, bv_for<T>(h), bf_for<T>(h),
                         va_for<T>(h), stop);
    else
      hipLaunchKernelGGL((conv_var_kernel<N1, M, T, false>), grid, block, 0, h->stream, z, y, bv_for<T>(h), bf_for<T>(h),
                         va_for<T>(h), stop);
  }
}

template <typename T>
int vel_launch(H* h, const T* z, T* y, const int* stop, bool trans) {
  if (!h->varb) return PDELAB_OK;
  const bool m1 = h->tab.m == h->n1;
  switch (h->n1) {
    case 2: m1 ? vel_launch_nm<2, 2, T>(h, z, y, stop, trans) : vel_launch_nm<2, 3, T>(h, z, y, stop, trans); break;
    case 3: m1 ? vel_launch_nm<3, 3, T>(h, z, y, stop, trans) : vel_launch_nm<3, 4, T>(h, z, y, stop, trans); break;
    case 4: m1 ? vel_launch_nm<4, 4, T>(h, z, y, stop, trans) : vel_launch_nm<4, 5, T>(h, z, y, stop, trans); break;
    default: return fail(h, PDELAB_ERR_UNSUPPORTED, "degree");
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}
template int vel_launch<double>(H*, const double*, double*, const int*, bool);
template int vel_launch<float>(H*, const float*, float*, const int*, bool);

template <int N1, int M>
static void vel_diag_nm(H* h, double* diag) {
  const long long n = h->N;
  hipLaunchKernelGGL((conv_var_diag_kernel<N1, M>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, diag,
                     (const double*)h->d_bv64, (const double*)h->d_bf64, h->va64);
}

int vel_diag(H* h, double* diag) {
  if (!h->varb) return PDELAB_OK;
  const bool m1 = h->tab.m == h->n1;
  switch (h->n1) {
    case 2: m1 ? vel_diag_nm<2, 2>(h, diag) : vel_diag_nm<2, 3>(h, diag); break;
    case 3: m1 ? vel_diag_nm<3, 3>(h, diag) : vel_diag_nm<3, 4>(h, diag); break;
    case 4: m1 ? vel_diag_nm<4, 4>(h, diag) : vel_diag_nm<4, 5>(h, diag); break;
    default: return fail(h, PDELAB_ERR_UNSUPPORTED, "degree");
  }
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}
