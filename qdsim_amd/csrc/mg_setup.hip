// Construction of the lattice multigrid hierarchy (the eigensolver's shifted
// preconditioner) on the device.
//
// No reference counterpart: the reference has no preconditioner.  The hierarchy is
// rebuilt for every solve because its coefficients follow the assembled potential;
// built from dozens of small tensor operations per level it cost ~25 ms of host
// launch time per benchmark solve.  Three kernels build it instead:
//
//  * lattice_coef_kernel: per lattice point (X, Y) of the full (NX x NY) lattice,
//    the edge weights WE(X -> X+1) = (c_a + c_b)/2 h_y/h_x and WN(Y -> Y+1) =
//    (c_a + c_b)/2 h_x/h_y from the kinetic coefficient of the two FE elements
//    touching the edge (sampled a quarter lattice step on either side), and the
//    mass term (mean of V over the four quarter points around the point - sigma)
//    h_x h_y, with the lumped mass h_x h_y;
//  * coarsen_kernel: vertex-centred coarsening (coarse point J at fine 2J): coarse
//    edge weight = harmonic mean of the two fine edges it spans, coarse mass terms =
//    R m with the bilinear weights (1/2, 1, 1/2) in each direction;
//  * interior_kernel: the level's unknowns (interior points): diagonal = the four
//    edge weights + mass term, east / north weights (zero across the last column /
//    row), in the hierarchy's storage type, plus FP64 copies of the diagonal and
//    lumped mass for later shift changes.
//
// Element lookup: element index 2 (j cx + i) + (u + v > 1) of the cell holding
// (x, y), with f = x * (1 / dxc) as the tensor form computed it, so the hierarchy is
// bitwise the one of that form.
#include "qds_common.h"

// every product and sum rounded on its own (no fused multiply-adds), as the tensor
// operations round them: the quarter-point positions decide element choices
#pragma clang fp contract(off)

namespace {

constexpr int kBlock = 256;

struct Geom {
  int64_t NX, NY, cx, cy;
  double hx, hy, inv_dxc, inv_dyc;
};

__device__ __forceinline__ int64_t element_at(const Geom& g, double x, double y) {
  const double fx = x * g.inv_dxc, fy = y * g.inv_dyc;
  const double i = fmin(fmax(floor(fx), 0.0), (double)(g.cx - 1));
  const double j = fmin(fmax(floor(fy), 0.0), (double)(g.cy - 1));
  const double u = fx - i, v = fy - j;
  return 2 * ((int64_t)j * g.cx + (int64_t)i) + (u + v > 1.0 ? 1 : 0);
}

__global__ void __launch_bounds__(kBlock)
lattice_coef_kernel(Geom g, const double* __restrict__ c, const double* __restrict__ v, double sigma,
                    double* __restrict__ WE, double* __restrict__ WN, double* __restrict__ Mt,
                    double* __restrict__ Mass) {
  const int64_t n = g.NX * g.NY;
  const double qx = 0.25 * g.hx, qy = 0.25 * g.hy;
  const double area = g.hx * g.hy;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t Y = t / g.NX, X = t - Y * g.NX;
    const double Xd = (double)X, Yd = (double)Y;
    if (X < g.NX - 1) {
      const double xm = (Xd + 0.5) * g.hx, yy = Yd * g.hy;
      const double ca = c[element_at(g, xm, yy + qy)], cb = c[element_at(g, xm, yy - qy)];
      WE[Y * (g.NX - 1) + X] = 0.5 * (ca + cb) * (g.hy / g.hx);
    }
    if (Y < g.NY - 1) {
      const double ym = (Yd + 0.5) * g.hy, xx = Xd * g.hx;
      const double ca = c[element_at(g, xx + qx, ym)], cb = c[element_at(g, xx - qx, ym)];
      WN[Y * g.NX + X] = 0.5 * (ca + cb) * (g.hx / g.hy);
    }
    const double xp = Xd * g.hx, yp = Yd * g.hy;
    double vs = v[element_at(g, xp + -qx, yp + -qy)];
    vs += v[element_at(g, xp + -qx, yp + qy)];
    vs += v[element_at(g, xp + qx, yp + -qy)];
    vs += v[element_at(g, xp + qx, yp + qy)];
    vs = vs * 0.25;
    Mt[t] = (vs - sigma) * area;
    Mass[t] = area;
  }
}

__device__ __forceinline__ double hmean(double a, double b) { return 2.0 * a * b / (a + b); }

// fine level NX x NY (odd) -> coarse NXc = (NX + 1) / 2, NYc = (NY + 1) / 2
__global__ void __launch_bounds__(kBlock)
coarsen_kernel(int64_t NX, int64_t NY, const double* __restrict__ WE, const double* __restrict__ WN,
               const double* __restrict__ Mt, const double* __restrict__ Mass, double* __restrict__ WEc,
               double* __restrict__ WNc, double* __restrict__ Mtc, double* __restrict__ Massc) {
  const int64_t NXc = (NX + 1) / 2, NYc = (NY + 1) / 2, n = NXc * NYc;
  const double w[3] = {0.5, 1.0, 0.5};
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t J = t / NXc, I = t - J * NXc;
    if (I < NXc - 1) WEc[J * (NXc - 1) + I] = hmean(WE[2 * J * (NX - 1) + 2 * I], WE[2 * J * (NX - 1) + 2 * I + 1]);
    if (J < NYc - 1) WNc[J * NXc + I] = hmean(WN[2 * J * NX + 2 * I], WN[(2 * J + 1) * NX + 2 * I]);
    double m = 0.0, s = 0.0;
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3; ++b) {
        const int64_t y = 2 * J + a - 1, x = 2 * I + b - 1;
        const bool in = y >= 0 && y < NY && x >= 0 && x < NX;
        const double wab = w[a] * w[b];
        m += wab * (in ? Mt[y * NX + x] : 0.0);
        s += wab * (in ? Mass[y * NX + x] : 0.0);
      }
    }
    Mtc[t] = m;
    Massc[t] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
interior_kernel(int64_t NX, int64_t NY, const double* __restrict__ WE, const double* __restrict__ WN,
                const double* __restrict__ Mt, const double* __restrict__ Mass, T* __restrict__ d,
                T* __restrict__ we, T* __restrict__ wn, double* __restrict__ d64, double* __restrict__ mass) {
  const int64_t nx = NX - 2, ny = NY - 2, n = nx * ny;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t y = t / nx, x = t - y * nx;
    const int64_t Yf = y + 1, Xf = x + 1;
    const double dd = WE[Yf * (NX - 1) + Xf - 1] + WE[Yf * (NX - 1) + Xf] + WN[(Yf - 1) * NX + Xf] +
                      WN[Yf * NX + Xf] + Mt[Yf * NX + Xf];
    d64[t] = dd;
    d[t] = (T)dd;
    we[t] = x == nx - 1 ? (T)0 : (T)WE[Yf * (NX - 1) + Xf];
    wn[t] = y == ny - 1 ? (T)0 : (T)WN[Yf * NX + Xf];
    mass[t] = Mass[Yf * NX + Xf];
  }
}

inline dim3 grid_for(int64_t n) { return dim3(qds_blocks(n, kBlock, 1 << 16)); }

}  // namespace

QDS_API int qds_mg_lattice_coefficients(int64_t NX, int64_t NY, int64_t cx, int64_t cy, double hx, double hy,
                                        double inv_dxc, double inv_dyc, const double* c_elem, const double* v_elem,
                                        double sigma, double* WE, double* WN, double* Mterm, double* Mass,
                                        void* stream) {
  QDS_CHECK_ARG(NX >= 2 && NY >= 2 && cx >= 1 && cy >= 1 && c_elem && v_elem && WE && WN && Mterm && Mass,
                "qds_mg_lattice_coefficients: bad argument");
  Geom g{NX, NY, cx, cy, hx, hy, inv_dxc, inv_dyc};
  hipLaunchKernelGGL(lattice_coef_kernel, grid_for(NX * NY), dim3(kBlock), 0, qds_stream(stream), g, c_elem, v_elem,
                     sigma, WE, WN, Mterm, Mass);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_mg_lattice_coarsen(int64_t NX, int64_t NY, const double* WE, const double* WN, const double* Mterm,
                                   const double* Mass, double* WEc, double* WNc, double* Mterm_c, double* Mass_c,
                                   void* stream) {
  QDS_CHECK_ARG(NX >= 3 && NY >= 3 && NX % 2 == 1 && NY % 2 == 1, "qds_mg_lattice_coarsen: NX, NY must be odd >= 3");
  QDS_CHECK_ARG(WE && WN && Mterm && Mass && WEc && WNc && Mterm_c && Mass_c, "qds_mg_lattice_coarsen: null pointer");
  const int64_t n = ((NX + 1) / 2) * ((NY + 1) / 2);
  hipLaunchKernelGGL(coarsen_kernel, grid_for(n), dim3(kBlock), 0, qds_stream(stream), NX, NY, WE, WN, Mterm, Mass,
                     WEc, WNc, Mterm_c, Mass_c);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_mg_lattice_interior(int32_t dtype, int64_t NX, int64_t NY, const double* WE, const double* WN,
                                    const double* Mterm, const double* Mass, void* d, void* we, void* wn, double* d64,
                                    double* mass, void* stream) {
  QDS_CHECK_ARG(NX >= 3 && NY >= 3 && WE && WN && Mterm && Mass && d && we && wn && d64 && mass,
                "qds_mg_lattice_interior: bad argument");
  const int64_t n = (NX - 2) * (NY - 2);
  hipStream_t st = qds_stream(stream);
  if (dtype == qds::F32)
    hipLaunchKernelGGL(interior_kernel<float>, grid_for(n), dim3(kBlock), 0, st, NX, NY, WE, WN, Mterm, Mass,
                       (float*)d, (float*)we, (float*)wn, d64, mass);
  else if (dtype == qds::F64)
    hipLaunchKernelGGL(interior_kernel<double>, grid_for(n), dim3(kBlock), 0, st, NX, NY, WE, WN, Mterm, Mass,
                       (double*)d, (double*)we, (double*)wn, d64, mass);
  else
    QDS_CHECK_ARG(false, "qds_mg_lattice_interior: dtype %d", dtype);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
