// apply_launch.hh -- kernel selection of the DG apply (included by apply_q1/2/3.hip).
#ifndef PDELAB_APPLY_LAUNCH_HH
#define PDELAB_APPLY_LAUNCH_HH

#include "handle.hh"

// the row kernel walking ZW z-chunks per workgroup (nz % ZW == 0, checked by apply_walk)
template <int N1, typename T, bool TRANS, bool REACT, bool ACCUM, bool DOT, int ZW>
static int launch_apply_walk(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, const LineMats<T>* lm,
                             long long nb, T* dotp, const int* stop) {
  hipLaunchKernelGGL((dg_apply_rowreg_kernel<N1, T, TRANS, false, REACT, false, ACCUM, DOT, ZW>),
                     dim3((unsigned)(nb / ZW)), dim3(ApplyShape<N1>::THREADS), 0, h->stream, z, y, fc, a, lm, nb / ZW,
                     dotp, stop);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <int N1, typename T, bool TRANS, bool CONV, bool REACT, bool VARK, bool ACCUM, bool DOT>
static int launch_apply_k(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, T* dotp, const int* stop) {
  using SH = ApplyShape<N1>;
  const long long nb = apply_blocks(h);
  const LineMats<T>* lm = line_mats<T>(h, TRANS);
  const int zw = apply_walk(h);
  if constexpr (!VARK && !CONV && N1 == 4) {
    if (zw == 2) return launch_apply_walk<N1, T, TRANS, REACT, ACCUM, DOT, 2>(h, z, y, fc, a, lm, nb, dotp, stop);
  }
  if constexpr (!VARK && !CONV && N1 == 3) {
    if (zw == 8) return launch_apply_walk<N1, T, TRANS, REACT, ACCUM, DOT, 8>(h, z, y, fc, a, lm, nb, dotp, stop);
    if (zw == 4) return launch_apply_walk<N1, T, TRANS, REACT, ACCUM, DOT, 4>(h, z, y, fc, a, lm, nb, dotp, stop);
  }
  if (row_kernel(h))
    hipLaunchKernelGGL((dg_apply_rowreg_kernel<N1, T, TRANS, CONV, REACT, VARK, ACCUM, DOT>), dim3((unsigned)nb),
                       dim3(SH::THREADS), 0, h->stream, z, y, fc, a, lm, nb, dotp, stop);
  else
    hipLaunchKernelGGL((dg_apply_kernel<N1, T, TRANS, CONV, REACT, VARK, ACCUM, DOT>), dim3((unsigned)nb),
                       dim3(SH::THREADS), 0, h->stream, z, y, fc, a, lm, nb, dotp, stop);
  HIPCHK(h, hipGetLastError());
  return PDELAB_OK;
}

template <int N1, typename T, bool TRANS, bool CONV, bool VARK, bool ACCUM, bool DOT>
static int launch_apply_t(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool react, T* dotp,
                          const int* stop) {
  if (react) return launch_apply_k<N1, T, TRANS, CONV, true, VARK, ACCUM, DOT>(h, z, y, fc, a, dotp, stop);
  return launch_apply_k<N1, T, TRANS, CONV, false, VARK, ACCUM, DOT>(h, z, y, fc, a, dotp, stop);
}

// mode 0: y = J z; 1: y += J z; 2: y = J z with the fused z.(Jz) partials
template <int N1, typename T, bool TRANS, bool CONV, bool VARK>
static int launch_apply_a(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool react, int mode, T* dotp,
                          const int* stop) {
  if (mode == 1) return launch_apply_t<N1, T, TRANS, CONV, VARK, true, false>(h, z, y, fc, a, react, nullptr, stop);
  if (mode == 2) return launch_apply_t<N1, T, TRANS, CONV, VARK, false, true>(h, z, y, fc, a, react, dotp, stop);
  return launch_apply_t<N1, T, TRANS, CONV, VARK, false, false>(h, z, y, fc, a, react, nullptr, stop);
}
template <int N1, typename T, bool TRANS, bool CONV>
static int launch_apply_v(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool react, int mode, T* dotp,
                          const int* stop) {
  return h->vark ? launch_apply_a<N1, T, TRANS, CONV, true>(h, z, y, fc, a, react, mode, dotp, stop)
                 : launch_apply_a<N1, T, TRANS, CONV, false>(h, z, y, fc, a, react, mode, dotp, stop);
}
template <int N1, typename T, bool TRANS>
static int launch_apply_c(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool react, int mode, T* dotp,
                          const int* stop) {
  // CONV selects the general line operator: convection, or constant coefficients
  // without the symmetric structure of line_op_sym
  const bool general = h->conv || (!h->vark && !h->sym[TRANS ? 1 : 0]);
  return general ? launch_apply_v<N1, T, TRANS, true>(h, z, y, fc, a, react, mode, dotp, stop)
                 : launch_apply_v<N1, T, TRANS, false>(h, z, y, fc, a, react, mode, dotp, stop);
}
template <int N1, typename T>
int launch_apply_n(H* h, const T* z, T* y, const T* fc, const DGArgs<T>& a, bool trans, int mode, T* dotp,
                          const int* stop) {
  const bool react = h->cfg.c + h->shift != 0.0;
  return trans ? launch_apply_c<N1, T, true>(h, z, y, fc, a, react, mode, dotp, stop)
               : launch_apply_c<N1, T, false>(h, z, y, fc, a, react, mode, dotp, stop);
}


#endif
