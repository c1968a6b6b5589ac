// FEM assembly on gfx950: element matrices, then an atomic-free segmented gather
// into CSR values.
//
// Reference: per-element quadrature loop of backend/src/fem.cpp:924-989
// (H_e += w (hbar^2/2m grad phi_i . grad phi_j + V phi_i phi_j), M_e += w phi_i phi_j)
// and the P1 centroid formulas of qdsim_cython/solvers/fem_solver.pyx:277-320,
// followed by scipy's COO->CSR duplicate summation (fem_solver.pyx:142-145).
//
// Kernel 1 (element matrices): one 64-lane wavefront owns a batch of elements.
// Each lane owns one entry (i <= j) of the packed symmetric element matrix of one
// element: P1 has 6 entries (10 elements per wave), P2 21 (3 per wave), P3 55 (1
// per wave).  The reference-triangle tables (weights, phi, dphi/dxi, dphi/deta at
// the quadrature points) live in LDS.  Up to three element matrices come out of
// one pass: out0 = stiffness(c) + mass(v), out1 = mass(r), out2 = mass(w), each
// coefficient sampled per quadrature point (stride_q = 1) or per element
// (stride_q = 0, the centroid rule of the Cython solvers).
//
// Kernel 2 (gather): one thread per CSR nonzero sums its segment of packed element
// entries in the fixed (element, i, j) order given by the host-built inverse map.
#include "qds_common.h"

namespace {

constexpr int kMaxQ = 16;
constexpr int kMaxNpe = 10;
constexpr int kWave = 64;
constexpr int kBlock = 256;
typedef double v4d __attribute__((ext_vector_type(4)));

struct Coef {
  const double* p;
  int64_t se;  // element stride
  int32_t sq;  // quadrature-point stride (0: one value per element)
};

__device__ __forceinline__ double coef_at(const Coef& c, int64_t e, int q) {
  return c.p[e * c.se + (int64_t)q * c.sq];
}

// Coefficient samples of one wave's elements, staged in LDS: up to four fields
// (c, v, r, w) x elements per wave x quadrature points.
constexpr int kMaxCoefSlots = 4 * 10 * kMaxQ;

// wave-level barrier with LDS visibility (the waves of a block run different
// numbers of elements, so a block barrier cannot sit inside the element loop)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(kBlock)
element_kernel(int npe, int nq, int64_t n_elem, const int32_t* __restrict__ conn,
               const double* __restrict__ x, const double* __restrict__ y,
               const double* __restrict__ tables, Coef c, Coef v, Coef r, Coef w,
               double* __restrict__ out0, double* __restrict__ out1, double* __restrict__ out2) {
  __shared__ double s_w[kMaxQ];
  __shared__ double s_phi[kMaxQ * kMaxNpe];
  __shared__ double s_dx[kMaxQ * kMaxNpe];
  __shared__ double s_dy[kMaxQ * kMaxNpe];
  __shared__ double s_cf[kBlock / kWave][kMaxCoefSlots];
  const int tq = nq * npe;
  for (int t = threadIdx.x; t < nq; t += blockDim.x) s_w[t] = tables[t];
  for (int t = threadIdx.x; t < tq; t += blockDim.x) {
    s_phi[t] = tables[nq + t];
    s_dx[t] = tables[nq + tq + t];
    s_dy[t] = tables[nq + 2 * tq + t];
  }
  __syncthreads();

  const int nsym = npe * (npe + 1) / 2;
  const int per_wave = kWave / nsym;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int sub = lane / nsym;
  const int s = lane - sub * nsym;
  // (i, j) of packed entry s: row-major upper triangle
  int i = 0, rem = s;
  while (rem >= npe - i) { rem -= npe - i; ++i; }
  const int j = i + rem;
  // staged fields: c, v, r, w in that order, each per_wave * nq samples
  const bool use[4] = {c.p != nullptr, v.p != nullptr, out1 != nullptr && r.p != nullptr,
                       out2 != nullptr && w.p != nullptr};
  const int per_field = per_wave * nq;
  int foff[4], nf = 0;
#pragma unroll
  for (int f = 0; f < 4; ++f) foff[f] = use[f] ? (nf++) * per_field : -1;
  double* cf = s_cf[wave];

  const int64_t waves_total = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t wb = (int64_t)blockIdx.x * (kBlock / kWave) + wave; wb * per_wave < n_elem;
       wb += waves_total) {
    const int64_t e_first = wb * per_wave;
    // one batched load of every coefficient sample of this wave's elements: the
    // quadrature loop below then reads LDS instead of issuing a dependent global
    // load per point
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (!use[f]) continue;
      const Coef& cw = f == 0 ? c : f == 1 ? v : f == 2 ? r : w;
      for (int t = lane; t < per_field; t += kWave) {
        const int se = t / nq, q = t - se * nq;
        const int64_t e = e_first + se;
        cf[foff[f] + t] = e < n_elem ? coef_at(cw, e, q) : 0.0;
      }
    }
    wave_sync();
    const int64_t e = e_first + sub;
    if (sub < per_wave && e < n_elem) {
      const int32_t* ce = conn + e * npe;
      const double x0 = x[ce[0]], y0 = y[ce[0]];
      const double a = x[ce[1]] - x0, b = x[ce[2]] - x0;
      const double cc = y[ce[1]] - y0, d = y[ce[2]] - y0;
      const double det = a * d - b * cc;
      const double area = 0.5 * fabs(det);
      const double inv2 = 1.0 / (det * det);
      const double g11 = (b * b + d * d) * inv2;
      const double g12 = -(a * b + cc * d) * inv2;
      const double g22 = (a * a + cc * cc) * inv2;
      const double* sc = cf + sub * nq;
      double k0 = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double wq = s_w[q] * area;
        const double pp = s_phi[q * npe + i] * s_phi[q * npe + j];
        if (use[0]) {
          const double xi = s_dx[q * npe + i], eti = s_dy[q * npe + i];
          const double xj = s_dx[q * npe + j], etj = s_dy[q * npe + j];
          const double gg = xi * xj * g11 + (xi * etj + eti * xj) * g12 + eti * etj * g22;
          k0 += wq * sc[foff[0] + q] * gg;
        }
        if (use[1]) m0 += wq * sc[foff[1] + q] * pp;
        if (out1) m1 += wq * (use[2] ? sc[foff[2] + q] : 1.0) * pp;
        if (use[3]) m2 += wq * sc[foff[3] + q] * pp;
      }
      const int64_t o = e * nsym + s;
      if (out0) out0[o] = k0 + m0;
      if (out1) out1[o] = m1;
      if (out2) out2[o] = m2;
    }
    wave_sync();   // every lane has read this element group's samples
  }
}

// ---------------------------------------------------------------------------
// Element kernel on FP64 MFMA.  Per element e and packed entry s = (i, j):
//   out0[e][s] = sum_q w_q |J|/2 [ c_q (g11 A_q + g12 B_q + g22 C_q) + v_q P_q ](s)
// with A = dxi_i dxi_j, B = dxi_i deta_j + deta_i dxi_j, C = deta_i deta_j,
// P = phi_i phi_j (reference-triangle tables, element independent) and
// g = J^-1 J^-T / det^2 the element metric.  That is a product
//   out0 (E x nsym) = Acoef (E x 4nq) . T (4nq x nsym),
//   Acoef[e][part nq + q] = w_q |J|/2 (g11 c, g12 c, g22 c, v)[part],
// and likewise out1 / out2 with the P rows of T and r / w.  A wave computes a
// tile of 16 elements with v_mfma_f64_16x16x4f64 (rows: elements, columns: 16
// packed entries, k: 4 of the 4nq coefficients); T lives in LDS, built once per
// workgroup; each lane forms its A operands from its element's geometry and
// coefficient samples.  Compared with one lane per entry looping over the
// quadrature points (9 LDS reads per point and lane), LDS traffic drops ~30x.
constexpr int kTile = 16;

template <int NPE, int NQ>
__global__ void __launch_bounds__(kBlock)
element_mfma_kernel(int64_t n_elem, const int32_t* __restrict__ conn, const double* __restrict__ x,
                    const double* __restrict__ y, const double* __restrict__ tables, Coef c, Coef v, Coef r,
                    Coef w, double* __restrict__ out0, double* __restrict__ out1, double* __restrict__ out2) {
  constexpr int NSYM = NPE * (NPE + 1) / 2;
  constexpr int NCT = (NSYM + 15) / 16;          // 16-column tiles of packed entries
  constexpr int NSP = 16 * NCT;
  constexpr int K0 = 4 * NQ;                     // rows of T used by out0
  constexpr int KS0 = K0 / 4;                    // k-steps of out0 (4 NQ is a multiple of 4)
  constexpr int KS1 = (NQ + 3) / 4;              // k-steps of out1 / out2
  constexpr int TROWS = K0 + 4;                  // 4 zero rows pad the last out1 k-step
  __shared__ double s_w[NQ];
  __shared__ double s_t[TROWS * NSP];
  for (int t = threadIdx.x; t < NQ; t += blockDim.x) s_w[t] = tables[t];
  for (int t = threadIdx.x; t < TROWS * NSP; t += blockDim.x) {
    const int k = t / NSP, sc = t - k * NSP;
    double val = 0.0;
    if (k < K0 && sc < NSYM) {
      int i = 0, rem = sc;
      while (rem >= NPE - i) { rem -= NPE - i; ++i; }
      const int j = i + rem;
      const int part = k / NQ, q = k - part * NQ;
      const double* phi = tables + NQ;
      const double* dx = phi + NQ * NPE;
      const double* dy = dx + NQ * NPE;
      const double xi = dx[q * NPE + i], xj = dx[q * NPE + j];
      const double ei = dy[q * NPE + i], ej = dy[q * NPE + j];
      val = part == 0 ? xi * xj : part == 1 ? xi * ej + ei * xj : part == 2 ? ei * ej
                                                                             : phi[q * NPE + i] * phi[q * NPE + j];
    }
    s_t[t] = val;
  }
  __syncthreads();

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int li = lane & 15, kg = lane >> 4;
  const bool has_c = c.p != nullptr, has_v = v.p != nullptr;
  const int64_t n_tiles = (n_elem + kTile - 1) / kTile;
  for (int64_t tile = (int64_t)blockIdx.x * (kBlock / kWave) + wave; tile < n_tiles;
       tile += (int64_t)gridDim.x * (kBlock / kWave)) {
    const int64_t e = tile * kTile + li;
    const bool live = e < n_elem;
    double g11 = 0.0, g12 = 0.0, g22 = 0.0, area = 0.0;
    if (live) {
      const int32_t* ce = conn + e * NPE;
      const double x0 = x[ce[0]], y0 = y[ce[0]];
      const double a = x[ce[1]] - x0, b = x[ce[2]] - x0;
      const double cc = y[ce[1]] - y0, d = y[ce[2]] - y0;
      const double det = a * d - b * cc;
      area = 0.5 * fabs(det);
      const double inv2 = 1.0 / (det * det);
      g11 = (b * b + d * d) * inv2;
      g12 = -(a * b + cc * d) * inv2;
      g22 = (a * a + cc * cc) * inv2;
    }
    const v4d zero = {0.0, 0.0, 0.0, 0.0};
    if (out0) {
      v4d acc[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[ct] = zero;
#pragma unroll
      for (int kk = 0; kk < KS0; ++kk) {
        const int k = 4 * kk + kg;
        const int part = k / NQ, q = k - part * NQ;
        double a = 0.0;
        if (live) {
          const double wa = s_w[q] * area;
          if (part < 3) {
            if (has_c) a = (part == 0 ? g11 : part == 1 ? g12 : g22) * wa * coef_at(c, e, q);
          } else if (has_v) {
            a = wa * coef_at(v, e, q);
          }
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s_t[k * NSP + 16 * ct + li], acc[ct], 0, 0, 0);
      }
      // result: lane l, register rr holds D[l/16 + 4 rr][l%16]
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int sc = 16 * ct + li;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int64_t er = tile * kTile + kg + 4 * rr;
          if (sc < NSYM && er < n_elem) out0[er * NSYM + sc] = acc[ct][rr];
        }
      }
    }
    // mass-type outputs: the P rows of T with coefficient r (1 when absent) or w
#pragma unroll
    for (int o = 1; o <= 2; ++o) {
      double* out = o == 1 ? out1 : out2;
      if (!out) continue;
      const Coef& cw = o == 1 ? r : w;
      v4d acc[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[ct] = zero;
#pragma unroll
      for (int kk = 0; kk < KS1; ++kk) {
        const int q = 4 * kk + kg;
        double a = 0.0;
        if (live && q < NQ) a = s_w[q] * area * (cw.p ? coef_at(cw, e, q) : 1.0);
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s_t[(3 * NQ + q) * NSP + 16 * ct + li], acc[ct], 0, 0, 0);
      }
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int sc = 16 * ct + li;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int64_t er = tile * kTile + kg + 4 * rr;
          if (sc < NSYM && er < n_elem) out[er * NSYM + sc] = acc[ct][rr];
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock)
gather_kernel(int64_t nnz, const int32_t* __restrict__ seg_ptr, const int32_t* __restrict__ seg_src,
              const double* __restrict__ e0, const double* __restrict__ e1,
              const double* __restrict__ e2, double* __restrict__ v0, double* __restrict__ v1,
              double* __restrict__ v2) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = seg_ptr[t], en = seg_ptr[t + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int32_t k = b; k < en; ++k) {
      const int32_t s = seg_src[k];
      if (e0) a0 += e0[s];
      if (e1) a1 += e1[s];
      if (e2) a2 += e2[s];
    }
    if (v0) v0[t] = a0;
    if (v1) v1[t] = a1;
    if (v2) v2[t] = a2;
  }
}

__global__ void quad_points_kernel(int npe, int nq, int64_t n_elem, const int32_t* __restrict__ conn,
                                   const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ bary, double* __restrict__ xq,
                                   double* __restrict__ yq) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_elem * nq;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / nq;
    const int q = (int)(t - e * nq);
    const int32_t* ce = conn + e * npe;
    const double l0 = bary[3 * q], l1 = bary[3 * q + 1], l2 = bary[3 * q + 2];
    xq[t] = l0 * x[ce[0]] + l1 * x[ce[1]] + l2 * x[ce[2]];
    yq[t] = l0 * y[ce[0]] + l1 * y[ce[1]] + l2 * y[ce[2]];
  }
}

// reduced[k] = full[src[k]] * scale (+ optional imaginary part), converted to the
// storage type of the eigensolver; src[k] < 0 marks a structural zero (padding of
// the lattice stencil form)
template <typename T>
__global__ void take_real_kernel(int64_t n, const int32_t* __restrict__ src,
                                 const double* __restrict__ vals, double scale, T* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = src[t];
    out[t] = s < 0 ? T(0) : static_cast<T>(vals[s] * scale);
  }
}

template <typename T>
__global__ void take_complex_kernel(int64_t n, const int32_t* __restrict__ src,
                                    const double* __restrict__ re, const double* __restrict__ im,
                                    double im_scale, T* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = src[t];
    out[2 * t] = s < 0 ? T(0) : static_cast<T>(re[s]);
    out[2 * t + 1] = (s < 0 || !im) ? T(0) : static_cast<T>(im[s] * im_scale);
  }
}

}  // namespace

// tables = [w(nq) | phi(nq*npe) | dphi_dxi(nq*npe) | dphi_deta(nq*npe)], row-major by q.
// Coefficient pointers may be NULL; stride_q 0 means one value per element.
QDS_API int qds_assemble_elements(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                                  const double* x, const double* y, const double* tables,
                                  const double* c_q, int64_t c_se, int32_t c_sq,
                                  const double* v_q, int64_t v_se, int32_t v_sq,
                                  const double* r_q, int64_t r_se, int32_t r_sq,
                                  const double* w_q, int64_t w_se, int32_t w_sq,
                                  double* out0, double* out1, double* out2, void* stream) {
  QDS_CHECK_ARG(npe == 3 || npe == 6 || npe == 10, "qds_assemble_elements: npe must be 3, 6 or 10");
  QDS_CHECK_ARG(nq >= 1 && nq <= kMaxQ, "qds_assemble_elements: nq must be in [1, %d]", kMaxQ);
  QDS_CHECK_ARG(conn && x && y && tables, "qds_assemble_elements: null geometry");
  QDS_CHECK_ARG(!out2 || w_q, "qds_assemble_elements: out2 needs w_q");
  QDS_CHECK_ARG(out0 || out1 || out2, "qds_assemble_elements: no output");
  QDS_CHECK_ARG(!out0 || c_q || v_q, "qds_assemble_elements: out0 needs c_q or v_q");
  if (n_elem == 0) return QDS_OK;
  const int nsym = npe * (npe + 1) / 2;
  Coef C{c_q, c_se, c_sq}, V{v_q, v_se, v_sq}, R{r_q, r_se, r_sq}, W{w_q, w_se, w_sq};
  hipStream_t st = qds_stream(stream);
  // the default rules of P1-P3 (degree 2p + 2: 6, 12, 16 points) run on MFMA
  const int64_t tiles = (n_elem + kTile - 1) / kTile;
  const unsigned mblocks = qds_blocks(tiles, kBlock / kWave, 4096);
#define QDS_ELEM_MFMA(P, Q)                                                                          \
  if (npe == P && nq == Q) {                                                                         \
    hipLaunchKernelGGL((element_mfma_kernel<P, Q>), dim3(mblocks), dim3(kBlock), 0, st, n_elem, conn, x, y, \
                       tables, C, V, R, W, out0, out1, out2);                                        \
    QDS_LAUNCH_CHECK();                                                                              \
    return QDS_OK;                                                                                   \
  }
  QDS_ELEM_MFMA(3, 6)
  QDS_ELEM_MFMA(6, 12)
  QDS_ELEM_MFMA(10, 16)
#undef QDS_ELEM_MFMA
  const int per_wave = kWave / nsym;
  const int64_t waves = (n_elem + per_wave - 1) / per_wave;
  const unsigned blocks = qds_blocks(waves, kBlock / kWave, 65536);
  hipLaunchKernelGGL(element_kernel, dim3(blocks), dim3(kBlock), 0, st, npe, nq,
                     n_elem, conn, x, y, tables, C, V, R, W, out0, out1, out2);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_assemble_gather(int64_t nnz, const int32_t* seg_ptr, const int32_t* seg_src,
                                const double* e0, const double* e1, const double* e2, double* v0,
                                double* v1, double* v2, void* stream) {
  QDS_CHECK_ARG(seg_ptr && seg_src, "qds_assemble_gather: null map");
  QDS_CHECK_ARG((!v0 || e0) && (!v1 || e1) && (!v2 || e2), "qds_assemble_gather: output without input");
  if (nnz == 0) return QDS_OK;
  hipLaunchKernelGGL(gather_kernel, dim3(qds_blocks(nnz, kBlock, 131072)), dim3(kBlock), 0,
                     qds_stream(stream), nnz, seg_ptr, seg_src, e0, e1, e2, v0, v1, v2);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_quad_points(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                            const double* x, const double* y, const double* bary, double* xq,
                            double* yq, void* stream) {
  QDS_CHECK_ARG(conn && x && y && bary && xq && yq, "qds_quad_points: null");
  if (n_elem == 0) return QDS_OK;
  hipLaunchKernelGGL(quad_points_kernel, dim3(qds_blocks(n_elem * nq, kBlock, 131072)),
                     dim3(kBlock), 0, qds_stream(stream), npe, nq, n_elem, conn, x, y, bary, xq, yq);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// dtype F32/F64: out[k] = vals[src[k]] * scale.
// dtype C64/C128: out[k] = re[src[k]] + i * im_scale * im[src[k]] (im may be NULL).
QDS_API int qds_csr_take(int32_t dtype, int64_t n, const int32_t* src, const double* re,
                         const double* im, double scale, void* out, void* stream) {
  QDS_CHECK_ARG(src && re && out, "qds_csr_take: null");
  if (n == 0) return QDS_OK;
  const dim3 g(qds_blocks(n, kBlock, 131072)), blk(kBlock);
  hipStream_t st = qds_stream(stream);
  switch (dtype) {
    case qds::F32:
      hipLaunchKernelGGL(take_real_kernel<float>, g, blk, 0, st, n, src, re, scale, (float*)out);
      break;
    case qds::F64:
      hipLaunchKernelGGL(take_real_kernel<double>, g, blk, 0, st, n, src, re, scale, (double*)out);
      break;
    case qds::C64:
      hipLaunchKernelGGL(take_complex_kernel<float>, g, blk, 0, st, n, src, re, im, scale, (float*)out);
      break;
    case qds::C128:
      hipLaunchKernelGGL(take_complex_kernel<double>, g, blk, 0, st, n, src, re, im, scale, (double*)out);
      break;
    default:
      QDS_CHECK_ARG(false, "qds_csr_take: bad dtype %d", dtype);
  }
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// ---------------------------------------------------------------------------
// Quadrature-point fields used by the self-consistent loop and the adjoint.
namespace {

// u_q[e, q] = sum_a phi_a(q) u[conn[e, a]]  (FE interpolation of a nodal field)
__global__ void interp_q_kernel(int npe, int nq, int64_t n_elem, const int32_t* __restrict__ conn,
                                const double* __restrict__ phi, const double* __restrict__ u,
                                double* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_elem * nq;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / nq;
    const int q = (int)(t - e * nq);
    double s = 0.0;
    for (int a = 0; a < npe; ++a) s += phi[q * npe + a] * u[conn[e * npe + a]];
    out[t] = s;
  }
}

// rho[e, q, s] = w_q |det J_e| / 2 * psi_s(x_q)^2, psi_s = sum_a phi_a(q) Psi[conn[e,a], s]
// (the Hellmann-Feynman weight: dE_s / dV(x_q) for M-normalised real eigenvectors)
__global__ void state_density_q_kernel(int npe, int nq, int64_t n_elem, const int32_t* __restrict__ conn,
                                       const double* __restrict__ x, const double* __restrict__ y,
                                       const double* __restrict__ tables, const double* __restrict__ psi,
                                       int64_t ldpsi, int k, double* __restrict__ out) {
  const double* w = tables;
  const double* phi = tables + nq;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_elem * nq * k;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t eq = t / k;
    const int s = (int)(t - eq * k);
    const int64_t e = eq / nq;
    const int q = (int)(eq - e * nq);
    const int32_t* ce = conn + e * npe;
    const double x0 = x[ce[0]], y0 = y[ce[0]];
    const double det = (x[ce[1]] - x0) * (y[ce[2]] - y0) - (x[ce[2]] - x0) * (y[ce[1]] - y0);
    double v = 0.0;
    for (int a = 0; a < npe; ++a) v += phi[q * npe + a] * psi[(int64_t)ce[a] * ldpsi + s];
    out[t] = w[q] * 0.5 * fabs(det) * v * v;
  }
}

}  // namespace

QDS_API int qds_interpolate_q(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                              const double* phi, const double* u, double* out, void* stream) {
  QDS_CHECK_ARG(conn && phi && u && out, "qds_interpolate_q: null");
  if (n_elem == 0) return QDS_OK;
  hipLaunchKernelGGL(interp_q_kernel, dim3(qds_blocks(n_elem * nq, kBlock, 131072)), dim3(kBlock), 0,
                     qds_stream(stream), npe, nq, n_elem, conn, phi, u, out);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_state_density_q(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                                const double* x, const double* y, const double* tables,
                                const double* psi, int64_t ldpsi, int32_t k, double* out, void* stream) {
  QDS_CHECK_ARG(conn && x && y && tables && psi && out && k >= 1 && ldpsi >= k,
                "qds_state_density_q: bad arguments");
  if (n_elem == 0) return QDS_OK;
  hipLaunchKernelGGL(state_density_q_kernel, dim3(qds_blocks(n_elem * nq * k, kBlock, 131072)),
                     dim3(kBlock), 0, qds_stream(stream), npe, nq, n_elem, conn, x, y, tables, psi,
                     ldpsi, k, out);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
