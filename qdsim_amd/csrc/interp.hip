// FE field evaluation at arbitrary points of a structured P1-P3 mesh.
//
// Reference: FEInterpolator (backend/src/fe_interpolator.cpp) -- interpolate
// (:54-160), interpolateWithGradient (:162-360), findElement (:363-410, a linear
// scan over all elements taking the first whose barycentrics are all >= -1e-10),
// computeBarycentricCoordinates (:433-455), shape values (:474-518) and gradients
// (:542-613); the CUDA interpolate_field_kernel (backend/src/gpu_kernels.cu:352-423)
// also scans every element per point.
//
// Here the mesh is the structured SimpleMesh, so the containing element follows
// from arithmetic: the candidates are the two triangles of the 3 x 3 cells around
// floor(x / dx), floor(y / dy), tested in increasing element index with the same
// barycentric test, which returns the element the linear scan would return.
// Points outside every element take the value of the nearest node and, for the
// gradient, the inverse-distance estimate over the nodes sharing an element with
// it, as the reference does.  One thread per point.
#include "qds_common.h"

namespace {

constexpr int kBlock = 256;

struct Mesh {
  int order, npe;
  int64_t cx, cy;        // cells
  double dx, dy;         // cell sizes (coordinate units of x, y)
  const int32_t* conn;   // [2 cx cy][npe]; element 2 (j cx + i) + t, t = 1 the upper triangle
  const double* x;
  const double* y;
  const int32_t* lnode;  // [(p cy + 1) (p cx + 1)] lowest node id at each lattice position
  const double* f;       // nodal fields, row-major [node][nf]
  int nf;
};

// barycentric coordinates with the reference's formula and inside test
__device__ __forceinline__ bool bary(double px, double py, double x0, double y0, double x1, double y1,
                                     double x2, double y2, double l[3]) {
  const double area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
  l[0] = 0.5 * ((x1 * y2 - x2 * y1) + (y1 - y2) * px + (x2 - x1) * py) / area;
  l[1] = 0.5 * ((x2 * y0 - x0 * y2) + (y2 - y0) * px + (x0 - x2) * py) / area;
  l[2] = 0.5 * ((x0 * y1 - x1 * y0) + (y0 - y1) * px + (x1 - x0) * py) / area;
  const double eps = 1e-10;
  return l[0] >= -eps && l[1] >= -eps && l[2] >= -eps;
}

__device__ int64_t locate(const Mesh& M, double px, double py, double l[3]) {
  const int64_t ci = (int64_t)floor(px / M.dx), cj = (int64_t)floor(py / M.dy);
  // fast path: a point well inside one of the two triangles of its own cell lies in
  // no other element, so the ordered scan below would return the same element
  if (ci >= 0 && ci < M.cx && cj >= 0 && cj < M.cy) {
    for (int t = 0; t < 2; ++t) {
      const int64_t e = 2 * (cj * M.cx + ci) + t;
      const int32_t* c = M.conn + e * M.npe;
      if (bary(px, py, M.x[c[0]], M.y[c[0]], M.x[c[1]], M.y[c[1]], M.x[c[2]], M.y[c[2]], l) &&
          l[0] > 1e-8 && l[1] > 1e-8 && l[2] > 1e-8)
        return e;
    }
  }
  for (int64_t j = cj - 1; j <= cj + 1; ++j) {
    if (j < 0 || j >= M.cy) continue;
    for (int64_t i = ci - 1; i <= ci + 1; ++i) {
      if (i < 0 || i >= M.cx) continue;
      for (int t = 0; t < 2; ++t) {
        const int64_t e = 2 * (j * M.cx + i) + t;
        const int32_t* c = M.conn + e * M.npe;
        if (bary(px, py, M.x[c[0]], M.y[c[0]], M.x[c[1]], M.y[c[1]], M.x[c[2]], M.y[c[2]], l)) return e;
      }
    }
  }
  return -1;
}

// shape values N[a] and derivatives dN[a][i] = dN_a / d lambda_i (fe_interpolator.cpp
// :474-518 node order: vertices, edges (01, 12, 20; P3 "a" node nearer the first
// vertex), bubble)
__device__ void shape(int order, const double l[3], double N[10], double dN[10][3]) {
  for (int a = 0; a < 10; ++a) dN[a][0] = dN[a][1] = dN[a][2] = 0.0;
  if (order == 1) {
    for (int a = 0; a < 3; ++a) { N[a] = l[a]; dN[a][a] = 1.0; }
    return;
  }
  const int ei[3] = {0, 1, 2}, ej[3] = {1, 2, 0};
  if (order == 2) {
    for (int a = 0; a < 3; ++a) { N[a] = l[a] * (2.0 * l[a] - 1.0); dN[a][a] = 4.0 * l[a] - 1.0; }
    for (int s = 0; s < 3; ++s) {
      const int i = ei[s], j = ej[s];
      N[3 + s] = 4.0 * l[i] * l[j];
      dN[3 + s][i] = 4.0 * l[j];
      dN[3 + s][j] = 4.0 * l[i];
    }
    return;
  }
  for (int a = 0; a < 3; ++a) {
    N[a] = 0.5 * l[a] * (3.0 * l[a] - 1.0) * (3.0 * l[a] - 2.0);
    dN[a][a] = 0.5 * (27.0 * l[a] * l[a] - 18.0 * l[a] + 2.0);
  }
  for (int s = 0; s < 3; ++s) {
    const int i = ei[s], j = ej[s];
    // node nearer vertex i, then nearer vertex j
    N[3 + 2 * s] = 4.5 * l[i] * l[j] * (3.0 * l[i] - 1.0);
    dN[3 + 2 * s][i] = 4.5 * l[j] * (6.0 * l[i] - 1.0);
    dN[3 + 2 * s][j] = 4.5 * l[i] * (3.0 * l[i] - 1.0);
    N[4 + 2 * s] = 4.5 * l[i] * l[j] * (3.0 * l[j] - 1.0);
    dN[4 + 2 * s][j] = 4.5 * l[i] * (6.0 * l[j] - 1.0);
    dN[4 + 2 * s][i] = 4.5 * l[j] * (3.0 * l[j] - 1.0);
  }
  N[9] = 27.0 * l[0] * l[1] * l[2];
  dN[9][0] = 27.0 * l[1] * l[2];
  dN[9][1] = 27.0 * l[0] * l[2];
  dN[9][2] = 27.0 * l[0] * l[1];
}

// nearest node of a point outside the mesh: clamp into the rectangle, then the
// nearest refined-lattice position; equal distances go to the lower node id
__device__ int32_t nearest_node(const Mesh& M, double px, double py) {
  const int64_t NX = M.order * M.cx + 1, NY = M.order * M.cy + 1;
  const double hx = M.dx / M.order, hy = M.dy / M.order;
  const double qx = fmin(fmax(px, 0.0), M.cx * M.dx), qy = fmin(fmax(py, 0.0), M.cy * M.dy);
  const int64_t fx = (int64_t)floor(qx / hx), fy = (int64_t)floor(qy / hy);
  int32_t best = -1;
  double bd = 0.0;
  for (int64_t ly = fy; ly <= fy + 1; ++ly) {
    if (ly < 0 || ly >= NY) continue;
    for (int64_t lx = fx; lx <= fx + 1; ++lx) {
      if (lx < 0 || lx >= NX) continue;
      const int32_t nd = M.lnode[ly * NX + lx];
      const double ddx = px - M.x[nd], ddy = py - M.y[nd];
      const double d = ddx * ddx + ddy * ddy;
      if (best < 0 || d < bd || (d == bd && nd < best)) { best = nd; bd = d; }
    }
  }
  return best;
}

// val, gx, gy: row-major [point][nf]
__global__ void __launch_bounds__(kBlock)
interp_points_kernel(Mesh M, int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                     double* __restrict__ val, double* __restrict__ gx, double* __restrict__ gy,
                     int32_t* __restrict__ elem) {
  const int nf = M.nf;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double X = px[t], Y = py[t];
    double l[3];
    const int64_t e = locate(M, X, Y, l);
    if (elem) elem[t] = (int32_t)e;
    if (e >= 0) {
      const int32_t* c = M.conn + e * M.npe;
      double N[10], dN[10][3];
      shape(M.order, l, N, dN);
      double ax[10], ay[10];
      if (gx) {
        const double x0 = M.x[c[0]], y0 = M.y[c[0]], x1 = M.x[c[1]], y1 = M.y[c[1]];
        const double x2 = M.x[c[2]], y2 = M.y[c[2]];
        const double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double lx[3] = {(y1 - y2) / det, (y2 - y0) / det, (y0 - y1) / det};
        const double ly[3] = {(x2 - x1) / det, (x0 - x2) / det, (x1 - x0) / det};
        for (int a = 0; a < M.npe; ++a) {
          ax[a] = dN[a][0] * lx[0] + dN[a][1] * lx[1] + dN[a][2] * lx[2];
          ay[a] = dN[a][0] * ly[0] + dN[a][1] * ly[1] + dN[a][2] * ly[2];
        }
      }
      for (int f = 0; f < nf; ++f) {
        double v = 0.0, sx = 0.0, sy = 0.0;
        for (int a = 0; a < M.npe; ++a) {
          const double fa = M.f[(int64_t)c[a] * nf + f];
          v += fa * N[a];
          if (gx) { sx += fa * ax[a]; sy += fa * ay[a]; }
        }
        val[t * nf + f] = v;
        if (gx) { gx[t * nf + f] = sx; gy[t * nf + f] = sy; }
      }
      continue;
    }
    const int32_t nd = nearest_node(M, X, Y);
    for (int f = 0; f < nf; ++f) val[t * nf + f] = M.f[(int64_t)nd * nf + f];
    if (!gx) continue;
    // nodes sharing an element with nd, first appearance in (element, local) order
    int32_t seen[64];
    int ns = 0;
    // lattice position of nd; the cells around it hold every element containing it
    const int p = M.order;
    const int64_t lxn = (int64_t)llround(M.x[nd] / (M.dx / p)), lyn = (int64_t)llround(M.y[nd] / (M.dy / p));
    for (int64_t j = lyn / p - 1; j <= lyn / p; ++j) {
      if (j < 0 || j >= M.cy) continue;
      for (int64_t i = lxn / p - 1; i <= lxn / p; ++i) {
        if (i < 0 || i >= M.cx) continue;
        for (int tt = 0; tt < 2; ++tt) {
          const int32_t* c = M.conn + (2 * (j * M.cx + i) + tt) * M.npe;
          bool has = false;
          for (int a = 0; a < M.npe; ++a) has = has || c[a] == nd;
          if (!has) continue;
          for (int a = 0; a < M.npe; ++a) {
            if (c[a] == nd) continue;
            bool dup = false;
            for (int q = 0; q < ns; ++q) dup = dup || seen[q] == c[a];
            if (!dup && ns < 64) seen[ns++] = c[a];
          }
        }
      }
    }
    for (int f = 0; f < nf; ++f) {
      double sx = 0.0, sy = 0.0, tw = 0.0;
      for (int q = 0; q < ns; ++q) {
        const double ddx = M.x[seen[q]] - M.x[nd], ddy = M.y[seen[q]] - M.y[nd];
        const double dist = sqrt(ddx * ddx + ddy * ddy);
        if (dist > 1e-10) {
          const double w = 1.0 / dist;
          const double df = M.f[(int64_t)seen[q] * nf + f] - M.f[(int64_t)nd * nf + f];
          sx += w * (df * ddx / dist);
          sy += w * (df * ddy / dist);
          tw += w;
        }
      }
      gx[t * nf + f] = tw > 0.0 ? sx / tw : 0.0;
      gy[t * nf + f] = tw > 0.0 ? sy / tw : 0.0;
    }
  }
}

// Values only (no gradients, no element ids), many fields: a block locates
// kVPts points (one thread each) and keeps their element and shape values in
// LDS; then all threads produce the (point, field) outputs with consecutive
// threads on consecutive fields of a point, so the stores and the reads of the
// field rows are contiguous.  Same arithmetic per output as interp_points_kernel
// (the nodal sum in local-node order; the nearest node outside the mesh).
constexpr int kVPts = kBlock;   // every thread locates one point
__global__ void __launch_bounds__(kBlock)
interp_values_kernel(Mesh M, int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                     double* __restrict__ val) {
  __shared__ int64_t s_e[kVPts];
  __shared__ int32_t s_nd[kVPts];
  __shared__ double s_N[kVPts][10];
  const int nf = M.nf;
  for (int64_t p0 = (int64_t)blockIdx.x * kVPts; p0 < n; p0 += (int64_t)gridDim.x * kVPts) {
    const int np = (int)min((int64_t)kVPts, n - p0);
    if ((int)threadIdx.x < np) {
      const int64_t t = p0 + threadIdx.x;
      double l[3];
      const int64_t e = locate(M, px[t], py[t], l);
      s_e[threadIdx.x] = e;
      if (e >= 0) {
        double N[10], dN[10][3];
        shape(M.order, l, N, dN);
        for (int a = 0; a < M.npe; ++a) s_N[threadIdx.x][a] = N[a];
      } else {
        s_nd[threadIdx.x] = nearest_node(M, px[t], py[t]);
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < np * nf; q += kBlock) {
      const int pt = q / nf, f = q - pt * nf;
      const int64_t e = s_e[pt];
      double v;
      if (e >= 0) {
        const int32_t* c = M.conn + e * M.npe;
        v = 0.0;
        for (int a = 0; a < M.npe; ++a) v += M.f[(int64_t)c[a] * nf + f] * s_N[pt][a];
      } else {
        v = M.f[(int64_t)s_nd[pt] * nf + f];
      }
      val[(p0 + pt) * nf + f] = v;
    }
    __syncthreads();
  }
}

}  // namespace

QDS_API int qds_interpolate_points(int32_t order, int64_t cx, int64_t cy, double dx, double dy,
                                   const int32_t* conn, const double* x, const double* y,
                                   const int32_t* lattice_node, const double* field, int32_t nfield,
                                   int64_t n_points, const double* px, const double* py, double* val,
                                   double* grad_x, double* grad_y, int32_t* elem, void* stream) {
  QDS_CHECK_ARG(order >= 1 && order <= 3 && cx >= 1 && cy >= 1 && dx > 0 && dy > 0 && nfield >= 1,
                "qds_interpolate_points: bad mesh description");
  QDS_CHECK_ARG(conn && x && y && lattice_node && field && (n_points == 0 || (px && py && val)),
                "qds_interpolate_points: null pointer");
  QDS_CHECK_ARG((grad_x == nullptr) == (grad_y == nullptr), "qds_interpolate_points: give both gradients or none");
  if (n_points == 0) return QDS_OK;
  Mesh M{order, order == 1 ? 3 : (order == 2 ? 6 : 10), cx, cy, dx, dy, conn, x, y, lattice_node, field, nfield};
  if (grad_x == nullptr && elem == nullptr && nfield >= 4) {
    hipLaunchKernelGGL(interp_values_kernel, dim3(qds_blocks((n_points + kVPts - 1) / kVPts, 1, 65536)),
                       dim3(kBlock), 0, qds_stream(stream), M, n_points, px, py, val);
    QDS_LAUNCH_CHECK();
    return QDS_OK;
  }
  hipLaunchKernelGGL(interp_points_kernel, dim3(qds_blocks(n_points, kBlock, 65536)), dim3(kBlock), 0,
                     qds_stream(stream), M, n_points, px, py, val, grad_x, grad_y, elem);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
