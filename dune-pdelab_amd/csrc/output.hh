// output.hh -- evaluation of the DG function at tensor sets of local points of every cell:
// the subsampled values a SubsamplingVTKWriter writes and the L2 error integral.
//
// u_h(xi) = sum_i x_i phi_i(xi) with the QkDG Lagrange basis (nodes j/k, x fastest,
// R:dune/pdelab/finiteelement/qkdglagrange.hh:55-220), the DiscreteGridFunction evaluation of
// R:dune/pdelab/gridfunctionspace/gridfunctionspaceutilities.hh:54.
#ifndef PDELAB_OUTPUT_HH
#define PDELAB_OUTPUT_HH

#include "block_reduce.hh"
#include "dg_load.hh"

namespace pdelab_amd {

constexpr int kMaxPts = 12;

// 1D points on [0,1], their weights (for integration) and the 1D basis values P[q][a].
struct PointSet {
  int m;
  double xi[kMaxPts], w[kMaxPts];
  double P[kMaxPts][4];
};

struct CellGeom {
  int n[3];
  double h[3];
  long long ncells;
};

// MODE 0: values[c m^3 + p] = u_h at point p = qa + m qb + m^2 qc of cell c.
// MODE 1: partial[c] = sum_p w_p (u_h - u)^2 |cell| with u the exact solution of the
// built-in problem (DifferenceSquaredAdapter integrated as integrateGridFunction does,
// R:dune/pdelab/common/functionutilities.hh:51-83).
template <int N1, int MODE>
__global__ __launch_bounds__(256) void cell_points_kernel(const double* __restrict__ x, const PointSet ps,
                                                          const CellGeom g, const LoadArgs la,
                                                          double* __restrict__ out) {
  constexpr int NL = N1 * N1 * N1;
  __shared__ double xc[NL];
  const long long c = blockIdx.x;
  const int tid = threadIdx.x;
  for (int i = tid; i < NL; i += blockDim.x) xc[i] = x[c * NL + i];
  __syncthreads();
  const int m = ps.m, np = m * m * m;
  const unsigned cu = (unsigned)c, n0 = (unsigned)g.n[0], n1 = (unsigned)g.n[1];
  const unsigned q = cu / n0, qz = q / n1;
  const int co[3] = {(int)(cu - q * n0), (int)(q - qz * n1), (int)qz};
  const double V = g.h[0] * g.h[1] * g.h[2];
  double acc = 0.0;
  for (int p = tid; p < np; p += blockDim.x) {
    const int qa = p % m, qb = (p / m) % m, qc = p / (m * m);
    double u = 0.0;
#pragma unroll
    for (int cc = 0; cc < N1; ++cc)
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < N1; ++a) s += xc[a + N1 * b + N1 * N1 * cc] * ps.P[qa][a];
        u += s * ps.P[qb][b] * ps.P[qc][cc];
      }
    if (MODE == 0) {
      out[c * np + p] = u;
    } else {
      const double X[3] = {(co[0] + ps.xi[qa]) * g.h[0], (co[1] + ps.xi[qb]) * g.h[1], (co[2] + ps.xi[qc]) * g.h[2]};
      const double e = u - exact_u(la, X);
      acc += ps.w[qa] * ps.w[qb] * ps.w[qc] * V * e * e;
    }
  }
  if (MODE == 1) {
    double v[1] = {acc};
    block_sum_n<256, double, 1>(v);
    if (tid == 0) out[c] = v[0];
  }
}

}  // namespace pdelab_amd

#endif
