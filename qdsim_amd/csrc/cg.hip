// Jacobi-preconditioned conjugate gradients in FP64 for the Poisson step of the
// self-consistent loop.
//
// Reference: CythonPoissonSolver.solve (qdsim_cython/solvers/poisson_solver.pyx:213-252)
// factorises the Dirichlet-modified stiffness with scipy spsolve.  Here the
// Dirichlet values are lifted into the right-hand side by the caller, so the
// system on the remaining DOFs is symmetric positive definite and CG applies.
// The whole iteration runs on the device: alpha and beta live in device memory
// and the host only reads the residual norm every `check_every` iterations.
#include "qds_common.h"
#include "../../include/qdsim_hip.h"   // qds_spmm, qds_mg_level, qds_mg_vcycle

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

// scalars layout (double): 0 rz, 1 pq, 2 rr, 3 alpha, 4 beta, 5 rz_new, 6 bnorm2
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

// partial[blk*2 + {0,1}] = sum a0*b0, sum a1*b1 over this block's slab
__global__ void __launch_bounds__(kBlock)
dot2_partial(int64_t n, const double* __restrict__ a0, const double* __restrict__ b0,
             const double* __restrict__ a1, const double* __restrict__ b1, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s0 += a0[i] * b0[i];
    if (a1) s1 += a1[i] * b1[i];
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

// final reduction of the partials into scalars[i0], scalars[i1], then the CG
// scalar update selected by `stage`:
//   stage 0: init   rz = s0, rr = s1
//   stage 1: alpha  pq = s0, alpha = rz / pq
//   stage 2: beta   rz_new = s0, rr = s1, beta = rz_new / rz, rz = rz_new
__global__ void __launch_bounds__(kBlock)
dot2_final(int nblk, const double* __restrict__ partial, double* __restrict__ sc, int stage) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    s0 += partial[2 * i];
    s1 += partial[2 * i + 1];
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    if (stage == 0) {
      sc[0] = s0;
      sc[2] = s1;
    } else if (stage == 1) {
      sc[1] = s0;
      sc[3] = s0 != 0.0 ? sc[0] / s0 : 0.0;
    } else {
      sc[5] = s0;
      sc[2] = s1;
      sc[4] = sc[0] != 0.0 ? s0 / sc[0] : 0.0;
      sc[0] = s0;
    }
  }
}

// x += alpha p ; r -= alpha q ; z = r * dinv
__global__ void __launch_bounds__(kBlock)
cg_update_xr(int64_t n, const double* __restrict__ sc, double* __restrict__ x, double* __restrict__ r,
             double* __restrict__ z, const double* __restrict__ p, const double* __restrict__ q,
             const double* __restrict__ dinv) {
  const double alpha = sc[3];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    z[i] = ri * dinv[i];
  }
}

// p = z + beta p
__global__ void __launch_bounds__(kBlock)
cg_update_p(int64_t n, const double* __restrict__ sc, const double* __restrict__ z, double* __restrict__ p) {
  const double beta = sc[4];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// x += alpha p ; r -= alpha q (the preconditioner runs separately)
__global__ void __launch_bounds__(kBlock)
cg_update_x_r(int64_t n, const double* __restrict__ sc, double* __restrict__ x, double* __restrict__ r,
              const double* __restrict__ p, const double* __restrict__ q) {
  const double alpha = sc[3];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

// r = b - r (r holds A x on entry)
__global__ void __launch_bounds__(kBlock)
cg_residual(int64_t n, const double* __restrict__ b, double* __restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - r[i];
}

// r = b - r (r holds A x on entry); z = r * dinv; p = z
__global__ void __launch_bounds__(kBlock)
cg_init(int64_t n, const double* __restrict__ b, double* __restrict__ r, double* __restrict__ z,
        double* __restrict__ p, const double* __restrict__ dinv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = b[i] - r[i];
    r[i] = ri;
    const double zi = ri * dinv[i];
    z[i] = zi;
    p[i] = zi;
  }
}

// dinv = 1 / diag(A)
__global__ void diag_inv(int64_t n, const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                         const double* __restrict__ v, double* __restrict__ dinv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int32_t k = ip[i]; k < ip[i + 1]; ++k)
      if (ix[k] == i) d = v[k];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
}

}  // namespace

QDS_API int64_t qds_cg_workspace(int64_t n) {
  // r, z, p, q, dinv vectors + partials + 8 scalars
  return (5 * n + 2 * kRedBlocks + 8) * (int64_t)sizeof(double);
}

// Solve A x = b (A SPD, CSR FP64).  x holds the initial guess on entry.
// Stops when ||r|| <= rtol * ||b|| (checked every check_every iterations) or after
// max_iter iterations.  iters_out / relres_out (host pointers) report the result.
QDS_API int qds_cg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                   const double* vals, const double* b, double* x, double rtol, int32_t max_iter,
                   int32_t check_every, void* workspace, int64_t ws_bytes, int32_t* iters_out,
                   double* relres_out, void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals && b && x && workspace, "qds_cg: null pointer");
  QDS_CHECK_ARG(ws_bytes >= qds_cg_workspace(n), "qds_cg: workspace too small");
  if (check_every < 1) check_every = 1;
  hipStream_t st = qds_stream(stream);
  double* r = static_cast<double*>(workspace);
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  double* dinv = q + n;
  double* partial = dinv + n;
  double* sc = partial + 2 * kRedBlocks;
  const dim3 g(qds_blocks(n, kBlock, kRedBlocks)), blk(kBlock);
  const int nblk = (int)g.x;

  hipLaunchKernelGGL(diag_inv, g, blk, 0, st, n, indptr, indices, vals, dinv);
  QDS_LAUNCH_CHECK();
  // ||b||^2
  hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, b, b, (const double*)nullptr,
                     (const double*)nullptr, partial);
  hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 1);  // sc[1] = b.b
  QDS_LAUNCH_CHECK();
  double bb = 0.0;
  QDS_HIP(hipMemcpyAsync(&bb, sc + 1, sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  if (bb == 0.0) {
    QDS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), st));
    if (iters_out) *iters_out = 0;
    if (relres_out) *relres_out = 0.0;
    return QDS_OK;
  }
  int rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, x, 1, r, 1, nullptr, 1, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(cg_init, g, blk, 0, st, n, b, r, z, p, dinv);
  hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, r, z, r, r, partial);
  hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 0);
  QDS_LAUNCH_CHECK();
  int it = 0;
  double rr = 0.0;
  const double tol2 = rtol * rtol * bb;
  QDS_HIP(hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  while (rr > tol2 && it < max_iter) {
    for (int k = 0; k < check_every && it < max_iter; ++k, ++it) {
      rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, p, 1, q, 1, nullptr, 1, stream);
      if (rc) return rc;
      hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, p, q, (const double*)nullptr,
                         (const double*)nullptr, partial);
      hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 1);
      hipLaunchKernelGGL(cg_update_xr, g, blk, 0, st, n, sc, x, r, z, p, q, dinv);
      hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, r, z, r, r, partial);
      hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 2);
      hipLaunchKernelGGL(cg_update_p, g, blk, 0, st, n, sc, z, p);
      QDS_LAUNCH_CHECK();
    }
    QDS_HIP(hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    QDS_HIP(hipStreamSynchronize(st));
  }
  if (iters_out) *iters_out = it;
  if (relres_out) *relres_out = sqrt(rr / bb);
  return QDS_OK;
}

// Conjugate gradients preconditioned by one multigrid V(nu, nu)-cycle per
// iteration (qds_mg_vcycle, FP64 hierarchy, one column).  levels[0] describes the
// unknowns of the system in lattice order (levels[0].nx * levels[0].ny == n); its
// f and u are replaced by the CG residual and preconditioned residual.  Same
// stopping rule, workspace and outputs as qds_cg.
QDS_API int qds_cg_mg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
                      const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
                      int32_t n_levels, const qds_mg_level* levels, int32_t nu, int32_t coarse_sweeps,
                      double omega, void* workspace, int64_t ws_bytes, int32_t* iters_out, double* relres_out,
                      void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals && b && x && workspace && levels, "qds_cg_mg: null pointer");
  QDS_CHECK_ARG(n_levels >= 1 && n_levels <= 32 && levels[0].nx * levels[0].ny == n,
                "qds_cg_mg: level 0 (%lld x %lld) does not hold the %lld unknowns", (long long)levels[0].nx,
                (long long)levels[0].ny, (long long)n);
  QDS_CHECK_ARG(ws_bytes >= qds_cg_workspace(n), "qds_cg_mg: workspace too small");
  if (check_every < 1) check_every = 1;
  hipStream_t st = qds_stream(stream);
  double* r = static_cast<double*>(workspace);
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  double* partial = q + 2 * n;   // after the (unused) Jacobi diagonal slot
  double* sc = partial + 2 * kRedBlocks;
  const dim3 g(qds_blocks(n, kBlock, kRedBlocks)), blk(kBlock);
  const int nblk = (int)g.x;
  qds_mg_level lv[32];
  for (int l = 0; l < n_levels; ++l) lv[l] = levels[l];
  lv[0].f = r;
  lv[0].u = z;
  lv[0].ldf = lv[0].ldu = 1;
  auto precondition = [&]() {
    return qds_mg_vcycle(qds::F64, n_levels, lv, 1, nu, nu, coarse_sweeps, omega, stream);
  };

  hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, b, b, (const double*)nullptr, (const double*)nullptr,
                     partial);
  hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 1);  // sc[1] = b.b
  QDS_LAUNCH_CHECK();
  double bb = 0.0;
  QDS_HIP(hipMemcpyAsync(&bb, sc + 1, sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  if (bb == 0.0) {
    QDS_HIP(hipMemsetAsync(x, 0, n * sizeof(double), st));
    if (iters_out) *iters_out = 0;
    if (relres_out) *relres_out = 0.0;
    return QDS_OK;
  }
  int rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, x, 1, r, 1, nullptr, 1, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(cg_residual, g, blk, 0, st, n, b, r);
  QDS_LAUNCH_CHECK();
  if ((rc = precondition())) return rc;
  QDS_HIP(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, r, z, r, r, partial);
  hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 0);   // rz, rr
  QDS_LAUNCH_CHECK();
  int it = 0;
  double rr = 0.0;
  const double tol2 = rtol * rtol * bb;
  QDS_HIP(hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  while (rr > tol2 && it < max_iter) {
    for (int k = 0; k < check_every && it < max_iter; ++k, ++it) {
      rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, p, 1, q, 1, nullptr, 1, stream);
      if (rc) return rc;
      hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, p, q, (const double*)nullptr,
                         (const double*)nullptr, partial);
      hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 1);   // alpha
      hipLaunchKernelGGL(cg_update_x_r, g, blk, 0, st, n, sc, x, r, p, q);
      QDS_LAUNCH_CHECK();
      if ((rc = precondition())) return rc;
      hipLaunchKernelGGL(dot2_partial, g, blk, 0, st, n, r, z, r, r, partial);
      hipLaunchKernelGGL(dot2_final, dim3(1), blk, 0, st, nblk, partial, sc, 2);   // beta, rr
      hipLaunchKernelGGL(cg_update_p, g, blk, 0, st, n, sc, z, p);
      QDS_LAUNCH_CHECK();
    }
    QDS_HIP(hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    QDS_HIP(hipStreamSynchronize(st));
  }
  if (iters_out) *iters_out = it;
  if (relres_out) *relres_out = sqrt(rr / bb);
  return QDS_OK;
}
