// CSR sparse-matrix x dense-block products on gfx950.
//
// Reference: the only reference SpMV is a thread-per-row complex128 kernel with no
// caller (backend/src/fused_gpu_kernels.cu:97-125); the reference solvers use scipy
// csr @ x inside ARPACK (qdsim_cython/solvers/eigen_solver.pyx:225-235).  Here:
//
// * Block products Y = A X (X is n x b, row-major with leading dimension ldx): a
//   group of lanes owns one row; each lane owns VEC consecutive columns of the
//   block, so the loads of X[col, :] are contiguous 16-byte vectors and every
//   nonzero is read once per row group.  With FUSED, two value arrays that share
//   the pattern (H and M) are applied in the same pass: the column indices are
//   read once and both Y_H and Y_M are written.
// * Single vectors (b == 1): LPR lanes split the nonzeros of a row and reduce with
//   cross-lane shuffles inside the 64-wide wavefront.
// Complex values are interleaved (re, im) pairs; products are complex-symmetric
// (no conjugation), as needed for the CAP Hamiltonian H - iW.
#include "qds_common.h"

#include <type_traits>

namespace {

constexpr int kBlock = 256;

template <typename T, int VEC>
struct VecT;
template <> struct VecT<float, 4> { using type = float4; };
template <> struct VecT<float, 2> { using type = float2; };
template <> struct VecT<float, 1> { using type = float; };
template <> struct VecT<double, 2> { using type = double2; };
template <> struct VecT<double, 1> { using type = double; };

template <typename T, int VEC>
__device__ __forceinline__ void load_vec(const T* p, T (&v)[VEC]) {
  using V = typename VecT<T, VEC>::type;
  V t = *reinterpret_cast<const V*>(p);
  const T* tp = reinterpret_cast<const T*>(&t);
#pragma unroll
  for (int i = 0; i < VEC; ++i) v[i] = tp[i];
}

template <typename T, int VEC>
__device__ __forceinline__ void store_vec(T* p, const T (&v)[VEC]) {
  using V = typename VecT<T, VEC>::type;
  V t;
  T* tp = reinterpret_cast<T*>(&t);
#pragma unroll
  for (int i = 0; i < VEC; ++i) tp[i] = v[i];
  *reinterpret_cast<V*>(p) = t;
}

// Rows are spread over the 8 XCDs in contiguous ranges: the hardware hands
// workgroups to the XCDs round-robin, so block id g runs on XCD g % 8; logical
// block (g % 8) * (G / 8) + g / 8 makes each XCD sweep its own eighth of the rows
// in order, and the X rows that a window of concurrently running blocks touches
// (the band of the matrix) stay in that XCD's L2 instead of being fetched by all
// eight.  G is a multiple of 8.
__device__ __forceinline__ int64_t xcd_block(int64_t g, int64_t G) {
  return (g % 8) * (G / 8) + g / 8;
}

// Real block product.  lpr lanes own one row (lpr = ceil(b / VEC), any value up to
// 64); a wave holds rpw = 64 / lpr rows, the remaining lanes idle.
template <typename T, int VEC, bool FUSED>
__global__ void __launch_bounds__(kBlock)
spmm_real(int64_t n, int b, int lpr, int rpw, const int32_t* __restrict__ indptr,
          const int32_t* __restrict__ indices, const T* __restrict__ va, const T* __restrict__ vb,
          const T* __restrict__ X, int64_t ldx, T* __restrict__ Ya, int64_t ldya,
          T* __restrict__ Yb, int64_t ldyb) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rg = lane / lpr;
  const int li = lane - rg * lpr;
  const int c0 = li * VEC;
  if (rg >= rpw || c0 >= b) return;
  const int64_t rows_per_block = (int64_t)(kBlock / 64) * rpw;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t r = lb * rows_per_block + (int64_t)wave * rpw + rg;
  if (r >= n) return;
  T acc_a[VEC], acc_b[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) { acc_a[v] = T(0); acc_b[v] = T(0); }
  const int32_t kb = indptr[r], ke = indptr[r + 1];
  int32_t k = kb;
  // groups of 8 nonzeros: all 8 index/value loads, then all 8 X loads, then the
  // FMAs -- two dependent memory round trips per 8 nonzeros
  for (; k + 8 <= ke; k += 8) {
    int32_t cc[8];
    T a[8], bb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      cc[u] = indices[k + u];
      a[u] = va[k + u];
      bb[u] = FUSED ? vb[k + u] : T(0);
    }
    T x[8][VEC];
#pragma unroll
    for (int u = 0; u < 8; ++u) load_vec<T, VEC>(X + (int64_t)cc[u] * ldx + c0, x[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        acc_a[v] = fma(a[u], x[u][v], acc_a[v]);
        if (FUSED) acc_b[v] = fma(bb[u], x[u][v], acc_b[v]);
      }
    }
  }
  for (; k + 2 <= ke; k += 2) {
    const int32_t c_0 = indices[k], c_1 = indices[k + 1];
    const T a0 = va[k], a1 = va[k + 1];
    T b0 = T(0), b1 = T(0);
    if (FUSED) { b0 = vb[k]; b1 = vb[k + 1]; }
    T x0[VEC], x1[VEC];
    load_vec<T, VEC>(X + (int64_t)c_0 * ldx + c0, x0);
    load_vec<T, VEC>(X + (int64_t)c_1 * ldx + c0, x1);
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      acc_a[v] = fma(a0, x0[v], acc_a[v]);
      acc_a[v] = fma(a1, x1[v], acc_a[v]);
      if (FUSED) {
        acc_b[v] = fma(b0, x0[v], acc_b[v]);
        acc_b[v] = fma(b1, x1[v], acc_b[v]);
      }
    }
  }
  for (; k < ke; ++k) {
    const int32_t c_0 = indices[k];
    const T a0 = va[k];
    T x0[VEC];
    load_vec<T, VEC>(X + (int64_t)c_0 * ldx + c0, x0);
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      acc_a[v] = fma(a0, x0[v], acc_a[v]);
      if (FUSED) acc_b[v] = fma(vb[k], x0[v], acc_b[v]);
    }
  }
  store_vec<T, VEC>(Ya + r * ldya + c0, acc_a);
  if (FUSED) store_vec<T, VEC>(Yb + r * ldyb + c0, acc_b);
}

// Complex block product; T is the real component type, VEC counts complex numbers
// per lane (loaded as 2*VEC reals).
template <typename T, int VEC, bool FUSED>
__global__ void __launch_bounds__(kBlock)
spmm_cplx(int64_t n, int b, int lpr, const int32_t* __restrict__ indptr,
          const int32_t* __restrict__ indices, const T* __restrict__ va, const T* __restrict__ vb,
          const T* __restrict__ X, int64_t ldx, T* __restrict__ Ya, int64_t ldya,
          T* __restrict__ Yb, int64_t ldyb) {
  constexpr int R = 2 * VEC;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int li = (int)(gtid % lpr);
  const int64_t groups = (int64_t)gridDim.x * blockDim.x / lpr;
  const int c0 = li * VEC;  // complex column
  const bool active = c0 < b;
  for (int64_t r = gtid / lpr; r < n; r += groups) {
    T aa[R], ab[R];
#pragma unroll
    for (int v = 0; v < R; ++v) { aa[v] = T(0); ab[v] = T(0); }
    const int32_t kb = indptr[r], ke = indptr[r + 1];
    if (active) {
      for (int32_t k = kb; k < ke; ++k) {
        const int32_t c = indices[k];
        const T ar = va[2 * k], ai = va[2 * k + 1];
        T br = T(0), bi = T(0);
        if (FUSED) { br = vb[2 * k]; bi = vb[2 * k + 1]; }
        T x[R];
        load_vec<T, R>(X + 2 * ((int64_t)c * ldx + c0), x);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const T xr = x[2 * v], xi = x[2 * v + 1];
          aa[2 * v] = fma(ar, xr, fma(-ai, xi, aa[2 * v]));
          aa[2 * v + 1] = fma(ar, xi, fma(ai, xr, aa[2 * v + 1]));
          if (FUSED) {
            ab[2 * v] = fma(br, xr, fma(-bi, xi, ab[2 * v]));
            ab[2 * v + 1] = fma(br, xi, fma(bi, xr, ab[2 * v + 1]));
          }
        }
      }
      store_vec<T, R>(Ya + 2 * (r * ldya + c0), aa);
      if (FUSED) store_vec<T, R>(Yb + 2 * (r * ldyb + c0), ab);
    }
  }
}

template <typename T>
__device__ __forceinline__ T group_sum(T v, int lpr) {
  for (int o = lpr / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Single-vector product: LPR lanes per row (power of two <= 64).
template <typename T, bool FUSED>
__global__ void __launch_bounds__(kBlock)
spmv_real(int64_t n, int lpr, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
          const T* __restrict__ va, const T* __restrict__ vb, const T* __restrict__ x,
          T* __restrict__ ya, T* __restrict__ yb) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int li = (int)(gtid % lpr);
  const int64_t groups = (int64_t)gridDim.x * blockDim.x / lpr;
  // all lanes of a wave iterate the same number of times so the shuffles stay converged
  const int64_t n_iter = (n + groups - 1) / groups;
  int64_t r = gtid / lpr;
  for (int64_t it = 0; it < n_iter; ++it, r += groups) {
    T sa = T(0), sb = T(0);
    if (r < n) {
      const int32_t ke = indptr[r + 1];
      for (int32_t k = indptr[r] + li; k < ke; k += lpr) {
        const T xv = x[indices[k]];
        sa = fma(va[k], xv, sa);
        if (FUSED) sb = fma(vb[k], xv, sb);
      }
    }
    sa = group_sum(sa, lpr);
    if (FUSED) sb = group_sum(sb, lpr);
    if (r < n && li == 0) {
      ya[r] = sa;
      if (FUSED) yb[r] = sb;
    }
  }
}

template <typename T, int VEC, bool FUSED>
int launch_real(int64_t n, int b, const int32_t* ip, const int32_t* ix, const void* va,
                const void* vb, const void* X, int64_t ldx, void* Ya, int64_t ldya, void* Yb,
                int64_t ldyb, hipStream_t st) {
  const int lpr = (b + VEC - 1) / VEC;
  QDS_CHECK_ARG(lpr <= 64, "qds_spmm: block width %d too large", b);
  const int rpw = 64 / lpr;
  const int64_t rows_per_block = (int64_t)(kBlock / 64) * rpw;
  int64_t blocks = (n + rows_per_block - 1) / rows_per_block;
  blocks = (blocks + 7) / 8 * 8;
  QDS_CHECK_ARG(blocks < (int64_t(1) << 31), "qds_spmm: too many rows");
  hipLaunchKernelGGL((spmm_real<T, VEC, FUSED>), dim3((unsigned)blocks), dim3(kBlock), 0, st, n, b, lpr, rpw,
                     ip, ix, (const T*)va, (const T*)vb, (const T*)X, ldx, (T*)Ya, ldya, (T*)Yb, ldyb);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

template <typename T, int VEC, bool FUSED>
int launch_cplx(int64_t n, int b, const int32_t* ip, const int32_t* ix, const void* va,
                const void* vb, const void* X, int64_t ldx, void* Ya, int64_t ldya, void* Yb,
                int64_t ldyb, hipStream_t st) {
  int lpr = 1;
  while (lpr * VEC < b) lpr <<= 1;
  QDS_CHECK_ARG(lpr <= 64, "qds_spmm: block width %d too large", b);
  const int64_t rows_per_block = kBlock / lpr;
  hipLaunchKernelGGL((spmm_cplx<T, VEC, FUSED>), dim3(qds_blocks(n, (int)rows_per_block, 262144)),
                     dim3(kBlock), 0, st, n, b, lpr, ip, ix, (const T*)va, (const T*)vb,
                     (const T*)X, ldx, (T*)Ya, ldya, (T*)Yb, ldyb);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

template <typename T>
int launch_spmv(int64_t n, int64_t nnz, const int32_t* ip, const int32_t* ix, const void* va,
                const void* vb, const void* x, void* ya, void* yb, hipStream_t st) {
  const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
  int lpr = 2;
  while (lpr < 16 && lpr < avg) lpr <<= 1;
  const int rows_per_block = kBlock / lpr;
  const dim3 g(qds_blocks(n, rows_per_block, 262144));
  if (vb)
    hipLaunchKernelGGL((spmv_real<T, true>), g, dim3(kBlock), 0, st, n, lpr, ip, ix, (const T*)va,
                       (const T*)vb, (const T*)x, (T*)ya, (T*)yb);
  else
    hipLaunchKernelGGL((spmv_real<T, false>), g, dim3(kBlock), 0, st, n, lpr, ip, ix, (const T*)va,
                       (const T*)nullptr, (const T*)x, (T*)ya, (T*)nullptr);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

bool aligned(const void* p, int bytes) { return (reinterpret_cast<uintptr_t>(p) % bytes) == 0; }

}  // namespace

// Y_a = A_a X and, when vals_b != NULL, Y_b = A_b X with the same pattern.
// X: n_cols x b block, row-major, leading dimension ldx (in elements of dtype).
QDS_API int qds_spmm(int32_t dtype, int64_t n_rows, int64_t nnz, int32_t b, const int32_t* indptr,
                     const int32_t* indices, const void* vals_a, const void* vals_b, const void* X,
                     int64_t ldx, void* Y_a, int64_t ldy_a, void* Y_b, int64_t ldy_b, void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals_a && X && Y_a, "qds_spmm: null pointer");
  QDS_CHECK_ARG(!vals_b || Y_b, "qds_spmm: vals_b without Y_b");
  QDS_CHECK_ARG(b >= 1 && ldx >= b && ldy_a >= b && (!vals_b || ldy_b >= b), "qds_spmm: bad block shape");
  if (n_rows == 0) return QDS_OK;
  hipStream_t st = qds_stream(stream);
  const bool fused = vals_b != nullptr;
  const int es = qds::dtype_size(dtype);
  QDS_CHECK_ARG(es > 0, "qds_spmm: bad dtype %d", dtype);
  if (b == 1 && (dtype == qds::F32 || dtype == qds::F64)) {
    if (dtype == qds::F32)
      return launch_spmv<float>(n_rows, nnz, indptr, indices, vals_a, vals_b, X, Y_a, Y_b, st);
    return launch_spmv<double>(n_rows, nnz, indptr, indices, vals_a, vals_b, X, Y_a, Y_b, st);
  }
  // vector width: widest that divides b and keeps 16-byte alignment of every row
  auto ok = [&](int vec_bytes, int vec) {
    if (b % vec) return false;
    bool a = aligned(X, vec_bytes) && aligned(Y_a, vec_bytes) && (ldx * es) % vec_bytes == 0 &&
             (ldy_a * es) % vec_bytes == 0;
    if (fused) a = a && aligned(Y_b, vec_bytes) && (ldy_b * es) % vec_bytes == 0;
    return a;
  };
  switch (dtype) {
    case qds::F32:
      if (ok(16, 4))
        return fused ? launch_real<float, 4, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                     : launch_real<float, 4, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
      if (ok(8, 2))
        return fused ? launch_real<float, 2, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                     : launch_real<float, 2, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
      return fused ? launch_real<float, 1, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                   : launch_real<float, 1, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
    case qds::F64:
      if (ok(16, 2))
        return fused ? launch_real<double, 2, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                     : launch_real<double, 2, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
      return fused ? launch_real<double, 1, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                   : launch_real<double, 1, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
    case qds::C64:
      if (ok(16, 2))
        return fused ? launch_cplx<float, 2, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                     : launch_cplx<float, 2, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
      return fused ? launch_cplx<float, 1, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                   : launch_cplx<float, 1, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
    case qds::C128:
      return fused ? launch_cplx<double, 1, true>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st)
                   : launch_cplx<double, 1, false>(n_rows, b, indptr, indices, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, st);
    default:
      QDS_CHECK_ARG(false, "qds_spmm: bad dtype %d", dtype);
  }
  return QDS_OK;
}

// ---------------------------------------------------------------------------
// CAP Hamiltonian product on complex blocks with real value arrays sharing one
// pattern: Y_H = (H - i W) X, Y_M = M X  (complex-symmetric, no conjugation).
// Reference: the CAP enters as V + i eta in backend/src/fem.cpp:942,974 (and is
// faked by a constant imaginary shift in the Cython solvers,
// schrodinger_solver.pyx:548-557); here -i W is applied exactly.
namespace {
template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
spmm_cap_kernel(int64_t n, int b, int lpr, const int32_t* __restrict__ indptr,
                const int32_t* __restrict__ indices, const T* __restrict__ vh, const T* __restrict__ vw,
                const T* __restrict__ vm, const T* __restrict__ X, int64_t ldx, T* __restrict__ YH,
                int64_t ldh, T* __restrict__ YM, int64_t ldm) {
  constexpr int R = 2 * VEC;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int li = (int)(gtid % lpr);
  const int64_t groups = (int64_t)gridDim.x * blockDim.x / lpr;
  const int c0 = li * VEC;
  if (c0 >= b) return;
  for (int64_t r = gtid / lpr; r < n; r += groups) {
    T ah[R], am[R];
#pragma unroll
    for (int v = 0; v < R; ++v) { ah[v] = T(0); am[v] = T(0); }
    const int32_t ke = indptr[r + 1];
    for (int32_t k = indptr[r]; k < ke; ++k) {
      const int32_t c = indices[k];
      const T h = vh[k], w = vw[k], m = vm[k];
      T x[R];
      load_vec<T, R>(X + 2 * ((int64_t)c * ldx + c0), x);
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const T xr = x[2 * v], xi = x[2 * v + 1];
        ah[2 * v] = fma(h, xr, fma(w, xi, ah[2 * v]));          // (h - iw)(xr + i xi)
        ah[2 * v + 1] = fma(h, xi, fma(-w, xr, ah[2 * v + 1]));
        am[2 * v] = fma(m, xr, am[2 * v]);
        am[2 * v + 1] = fma(m, xi, am[2 * v + 1]);
      }
    }
    store_vec<T, R>(YH + 2 * (r * ldh + c0), ah);
    store_vec<T, R>(YM + 2 * (r * ldm + c0), am);
  }
}
}  // namespace

// dtype C64 (real arrays FP32) or C128 (FP64); X, Y complex interleaved blocks.
QDS_API int qds_spmm_cap(int32_t dtype, int64_t n_rows, int32_t b, const int32_t* indptr,
                         const int32_t* indices, const void* vals_h, const void* vals_w,
                         const void* vals_m, const void* X, int64_t ldx, void* Y_h, int64_t ldy_h,
                         void* Y_m, int64_t ldy_m, void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals_h && vals_w && vals_m && X && Y_h && Y_m && b >= 1,
                "qds_spmm_cap: bad arguments");
  if (n_rows == 0) return QDS_OK;
  hipStream_t st = qds_stream(stream);
  const int es = qds::dtype_size(dtype);
  const bool v2 = (b % 2 == 0) && aligned(X, 16) && aligned(Y_h, 16) && aligned(Y_m, 16) &&
                  (ldx * es) % 16 == 0 && (ldy_h * es) % 16 == 0 && (ldy_m * es) % 16 == 0;
  auto launch = [&](auto tag, auto vec) -> int {
    using T = decltype(tag);
    constexpr int VEC = decltype(vec)::value;
    int lpr = 1;
    while (lpr * VEC < b) lpr <<= 1;
    QDS_CHECK_ARG(lpr <= 64, "qds_spmm_cap: block width %d too large", b);
    hipLaunchKernelGGL((spmm_cap_kernel<T, VEC>), dim3(qds_blocks(n_rows, kBlock / lpr, 262144)),
                       dim3(kBlock), 0, st, n_rows, b, lpr, indptr, indices, (const T*)vals_h,
                       (const T*)vals_w, (const T*)vals_m, (const T*)X, ldx, (T*)Y_h, ldy_h, (T*)Y_m, ldy_m);
    QDS_LAUNCH_CHECK();
    return QDS_OK;
  };
  if (dtype == qds::C64)
    return v2 ? launch(float(0), std::integral_constant<int, 2>()) : launch(float(0), std::integral_constant<int, 1>());
  if (dtype == qds::C128)
    return launch(double(0), std::integral_constant<int, 1>());
  QDS_CHECK_ARG(false, "qds_spmm_cap: dtype must be C64 or C128");
  return QDS_EINVAL;
}
