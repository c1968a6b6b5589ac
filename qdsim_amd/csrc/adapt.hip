// Refinement indicator of the adaptive mesh driver.
//
// Reference: AdaptiveMesh::computeRefinementFlags (backend/src/adaptive_mesh.cpp:431-475)
// fits psi ~ a + b x + c y over the nodes of each element by the pseudo-inverse of
// [1 x y] and flags the element when |(b, c)| exceeds a threshold; FEMSolver::adapt_mesh
// (backend/src/fem.cpp:486-542) refines the flagged elements.  Here one thread per
// element forms the 3x3 normal equations of the same least-squares fit (coordinates
// taken relative to the element's first vertex, which leaves the gradient unchanged
// and keeps the system well scaled) and solves them by Cramer's rule; for npe >= 3
// nodes that are not collinear the normal-equation solution is the pseudo-inverse
// solution.
#include "qds_common.h"

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
grad_indicator_kernel(int npe, int64_t n_elem, const int32_t* __restrict__ conn, const double* __restrict__ x,
                      const double* __restrict__ y, const double* __restrict__ psi, double coord_scale,
                      double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* ce = conn + e * npe;
    const double x0 = x[ce[0]], y0 = y[ce[0]];
    // sums of the normal equations  [n sx sy; sx sxx sxy; sy sxy syy] (a b c)' = (s sxp syp)'
    double n = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, s = 0, sxp = 0, syp = 0;
    for (int k = 0; k < npe; ++k) {
      const int32_t g = ce[k];
      const double dx = (x[g] - x0) * coord_scale, dy = (y[g] - y0) * coord_scale, p = psi[g];
      n += 1.0;
      sx += dx;
      sy += dy;
      sxx += dx * dx;
      sxy += dx * dy;
      syy += dy * dy;
      s += p;
      sxp += dx * p;
      syp += dy * p;
    }
    const double det = n * (sxx * syy - sxy * sxy) - sx * (sx * syy - sxy * sy) + sy * (sx * sxy - sxx * sy);
    double gb = 0.0, gc = 0.0;
    if (det != 0.0) {
      // Cramer's rule for b and c
      gb = (n * (sxp * syy - sxy * syp) - s * (sx * syy - sxy * sy) + sy * (sx * syp - sxp * sy)) / det;
      gc = (n * (sxx * syp - sxp * sxy) - sx * (sx * syp - sxp * sy) + s * (sx * sxy - sxx * sy)) / det;
    }
    out[e] = sqrt(gb * gb + gc * gc);
  }
}

}  // namespace

// out[e] = |least-squares gradient of psi over element e's npe nodes|, with the
// node coordinates multiplied by coord_scale (e.g. nm -> m) before the fit.
QDS_API int qds_gradient_indicator(int32_t npe, int64_t n_elem, const int32_t* conn, const double* x,
                                   const double* y, const double* psi, double coord_scale, double* out,
                                   void* stream) {
  QDS_CHECK_ARG(npe == 3 || npe == 6 || npe == 10, "qds_gradient_indicator: npe must be 3, 6 or 10");
  QDS_CHECK_ARG(conn && x && y && psi && out, "qds_gradient_indicator: null pointer");
  if (n_elem == 0) return QDS_OK;
  hipLaunchKernelGGL(grad_indicator_kernel, dim3(qds_blocks(n_elem, kBlock, 65536)), dim3(kBlock), 0,
                     qds_stream(stream), npe, n_elem, conn, x, y, psi, coord_scale, out);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
