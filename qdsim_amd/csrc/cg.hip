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

// ---- BiCGStab (nonsymmetric systems) -----------------------------------------------
//
// Jacobi left-preconditioned BiCGStab in FP64 for the drift-diffusion systems.
// Reference: full_poisson_dd_solver.cpp:1018-1040 (and :473-480 for the Poisson
// step) factorises those matrices with SimplicialLDLT and solves directly; here
// the row-replaced system is solved iteratively on the device.  The iteration runs
// on D^-1 A x = D^-1 b (D = diag A, 1 where it is zero), so the residual it
// measures and stops on is the row-scaled one: a replaced Dirichlet row (diagonal
// 1) and an interior row (diagonal of the operator's scale, e.g. eps ~ 1e-12 F/cm)
// weigh alike.  Scalars live in device memory; a breakdown (rho, omega or <r0, v>
// at zero, or a non-finite step) raises a flag that stops every later update in
// the burst, and the host restarts from the last iterate with a fresh shadow
// residual.
namespace {

// sc layout (double): 0 rho, 1 alpha, 2 omega, 3 beta, 4 rr, 5 flag, 6 r0v, 7 ts, 8 tt, 9 bb
__device__ __forceinline__ double bs_block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;
}

// d_i of the left scaling (1 without one)
__device__ __forceinline__ double dscale(const double* __restrict__ dinv, int64_t i) {
  return dinv ? dinv[i] : 1.0;
}

// r = dinv (b - r) (r holds A x); r0 = r; p = v = 0; partial (r.r)
__global__ void __launch_bounds__(kBlock)
bs_init(int64_t n, const double* __restrict__ b, const double* __restrict__ dinv, double* __restrict__ r,
        double* __restrict__ r0, double* __restrict__ p, double* __restrict__ v, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = dscale(dinv, i) * (b[i] - r[i]);
    r[i] = ri;
    r0[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
    s += ri * ri;
  }
  s = bs_block_sum(s, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s;
    partial[2 * blockIdx.x + 1] = s;
  }
}

// stage 0 (restart): rho = r0.r = rr, alpha = omega = 1, beta = 0, flag = (rho == 0)
// stage 1: r0v = s0, alpha = rho / r0v
// stage 2: ts = s0, tt = s1, omega = ts / tt
// stage 3: rho_new = s0, rr = s1, beta = (rho_new / rho) (alpha / omega), rho = rho_new
// stage 4: bb = s0
__global__ void __launch_bounds__(kBlock)
bs_final(int nblk, const double* __restrict__ partial, double* __restrict__ sc, int stage) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    s0 += partial[2 * i];
    s1 += partial[2 * i + 1];
  }
  s0 = bs_block_sum(s0, sh);
  s1 = bs_block_sum(s1, sh);
  if (threadIdx.x != 0) return;
  if (stage == 4) {
    sc[9] = s1;
    return;
  }
  if (stage == 0) {
    sc[0] = s0;
    sc[1] = 1.0;
    sc[2] = 1.0;
    sc[3] = 0.0;
    sc[4] = s1;
    sc[5] = s0 != 0.0 ? 0.0 : 1.0;
    return;
  }
  if (sc[5] != 0.0) return;
  if (stage == 1) {
    sc[6] = s0;
    const double a = sc[0] / s0;
    if (s0 == 0.0 || !isfinite(a)) sc[5] = 1.0;
    else sc[1] = a;
  } else if (stage == 2) {
    sc[7] = s0;
    sc[8] = s1;
    const double w = s0 / s1;
    if (s1 == 0.0 || w == 0.0 || !isfinite(w)) sc[5] = 1.0;
    else sc[2] = w;
  } else {
    sc[4] = s1;
    const double bt = (s0 / sc[0]) * (sc[1] / sc[2]);
    if (s0 == 0.0 || !isfinite(bt)) sc[5] = 1.0;
    else {
      sc[3] = bt;
      sc[0] = s0;
    }
  }
}

// p = r + beta (p - omega v)
__global__ void __launch_bounds__(kBlock)
bs_update_p(int64_t n, const double* __restrict__ sc, const double* __restrict__ r, double* __restrict__ p,
            const double* __restrict__ v) {
  if (sc[5] != 0.0) return;
  const double beta = sc[3], omega = sc[2];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// partial (a . (d b)) and, with c, ((d c) . (d c)) ; d = dinv scales the raw products
__global__ void __launch_bounds__(kBlock)
bs_dot_scaled(int64_t n, const double* __restrict__ sc, const double* __restrict__ a, const double* __restrict__ b,
              const double* __restrict__ c, const double* __restrict__ dinv, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  if (sc == nullptr || sc[5] == 0.0) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double d = dscale(dinv, i);
      s0 += a[i] * (d * b[i]);
      if (c) {
        const double dc = d * c[i];
        s1 += dc * dc;
      }
    }
  }
  s0 = bs_block_sum(s0, sh);
  s1 = bs_block_sum(s1, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

// v = dinv w ; s = r - alpha v (in r)
__global__ void __launch_bounds__(kBlock)
bs_update_s(int64_t n, const double* __restrict__ sc, double* __restrict__ r, const double* __restrict__ w,
            double* __restrict__ v, const double* __restrict__ dinv) {
  if (sc[5] != 0.0) return;
  const double alpha = sc[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double vi = dscale(dinv, i) * w[i];
    v[i] = vi;
    r[i] = r[i] - alpha * vi;
  }
}

// x += alpha P + omega S ; r = s - omega dinv t ; partial (r0.r), (r.r)   (s in r;
// P, S = p, s with the left scaling, the preconditioned p^, s^ with a right one)
__global__ void __launch_bounds__(kBlock)
bs_update_x(int64_t n, const double* __restrict__ sc, double* __restrict__ x, double* r,
            const double* __restrict__ r0, const double* __restrict__ P, const double* S,
            const double* __restrict__ t, const double* __restrict__ dinv, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0;
  if (sc[5] == 0.0) {
    const double alpha = sc[1], omega = sc[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double si = r[i];
      x[i] += alpha * P[i] + omega * S[i];
      const double ri = si - omega * dscale(dinv, i) * t[i];
      r[i] = ri;
      s0 += r0[i] * ri;
      s1 += ri * ri;
    }
  }
  s0 = bs_block_sum(s0, sh);
  s1 = bs_block_sum(s1, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

}  // namespace

QDS_API int64_t qds_bicgstab_workspace(int64_t n) {
  // r, r0, p, v, w, t, dinv or p^, s^ + partials + scalars
  return (8 * n + 2 * kRedBlocks + 16) * (int64_t)sizeof(double);
}

namespace {

// BiCGStab on A x = b.  dinv != NULL: left scaling D^-1 (Jacobi), residual measured
// scaled; precondition != NULL: right preconditioner u = M^-1 f (f, u device
// vectors of length n), residual measured unscaled.
template <class Precond>
int bicgstab_impl(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
                  const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every, bool jacobi,
                  Precond precondition, void* workspace, int32_t* iters_out, double* relres_out,
                  int32_t* restarts_out, hipStream_t st) {
  if (check_every < 1) check_every = 1;
  double* r = static_cast<double*>(workspace);
  double* r0 = r + n;
  double* p = r0 + n;
  double* v = p + n;
  double* w = v + n;
  double* t = w + n;
  double* dinv = jacobi ? t + n : nullptr;
  double* ph = jacobi ? p : t + n;         // the operand of the first product
  double* shat = jacobi ? r : t + 2 * n;   // and of the second
  double* partial = t + 3 * n;
  double* sc = partial + 2 * kRedBlocks;
  const dim3 g(qds_blocks(n, kBlock, kRedBlocks)), blk(kBlock);
  const int nblk = (int)g.x;

  if (jacobi) hipLaunchKernelGGL(diag_inv, g, blk, 0, st, n, indptr, indices, vals, dinv);
  hipLaunchKernelGGL(bs_dot_scaled, g, blk, 0, st, n, (const double*)nullptr, b, b, b, (const double*)dinv, partial);
  hipLaunchKernelGGL(bs_final, dim3(1), blk, 0, st, nblk, partial, sc, 4);
  QDS_LAUNCH_CHECK();
  double hs[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  QDS_HIP(hipMemcpyAsync(hs, sc, 10 * sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  const double bb = hs[9];   // ||D^-1 b||^2 (||b||^2 without the scaling)
  int restarts = 0;
  if (!(bb > 0.0)) {
    // b = 0: x = 0.  A non-finite b gives a non-finite x (all-ones bytes are a NaN),
    // as a direct factorisation would; relres_out reports NaN
    QDS_HIP(hipMemsetAsync(x, bb == 0.0 ? 0 : 0xff, n * sizeof(double), st));
    if (iters_out) *iters_out = 0;
    if (relres_out) *relres_out = bb == 0.0 ? 0.0 : bb;
    if (restarts_out) *restarts_out = 0;
    return QDS_OK;
  }
  const double tol2 = rtol * rtol * bb;
  auto restart = [&]() -> int {
    int rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, x, 1, r, 1, nullptr, 1, st);
    if (rc) return rc;
    hipLaunchKernelGGL(bs_init, g, blk, 0, st, n, b, (const double*)dinv, r, r0, p, v, partial);
    hipLaunchKernelGGL(bs_final, dim3(1), blk, 0, st, nblk, partial, sc, 0);
    QDS_LAUNCH_CHECK();
    return QDS_OK;
  };
  int rc = restart();
  if (rc) return rc;
  QDS_HIP(hipMemcpyAsync(hs, sc, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  QDS_HIP(hipStreamSynchronize(st));
  int it = 0;
  while (hs[4] > tol2 && it < max_iter) {
    for (int k = 0; k < check_every && it < max_iter; ++k, ++it) {
      hipLaunchKernelGGL(bs_update_p, g, blk, 0, st, n, sc, r, p, v);
      if (!jacobi && (rc = precondition(p, ph))) return rc;
      rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, ph, 1, w, 1, nullptr, 1, st);
      if (rc) return rc;
      hipLaunchKernelGGL(bs_dot_scaled, g, blk, 0, st, n, sc, r0, w, (const double*)nullptr, (const double*)dinv,
                         partial);
      hipLaunchKernelGGL(bs_final, dim3(1), blk, 0, st, nblk, partial, sc, 1);
      hipLaunchKernelGGL(bs_update_s, g, blk, 0, st, n, sc, r, w, v, (const double*)dinv);
      if (!jacobi && (rc = precondition(r, shat))) return rc;
      rc = qds_spmm(qds::F64, n, nnz, 1, indptr, indices, vals, nullptr, shat, 1, t, 1, nullptr, 1, st);
      if (rc) return rc;
      hipLaunchKernelGGL(bs_dot_scaled, g, blk, 0, st, n, sc, r, t, t, (const double*)dinv, partial);
      hipLaunchKernelGGL(bs_final, dim3(1), blk, 0, st, nblk, partial, sc, 2);
      hipLaunchKernelGGL(bs_update_x, g, blk, 0, st, n, sc, x, r, r0, ph, shat, t, (const double*)dinv, partial);
      hipLaunchKernelGGL(bs_final, dim3(1), blk, 0, st, nblk, partial, sc, 3);
      QDS_LAUNCH_CHECK();
    }
    QDS_HIP(hipMemcpyAsync(hs, sc, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    QDS_HIP(hipStreamSynchronize(st));
    if (hs[5] != 0.0 && it < max_iter) {
      // breakdown inside the burst: the updates after it were skipped and x is the
      // last good iterate; restart from it
      if (++restarts > 50) break;
      if ((rc = restart())) return rc;
      QDS_HIP(hipMemcpyAsync(hs, sc, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
      QDS_HIP(hipStreamSynchronize(st));
    }
  }
  if (iters_out) *iters_out = it;
  if (relres_out) *relres_out = sqrt(hs[4] / bb);
  if (restarts_out) *restarts_out = restarts;
  return QDS_OK;
}

}  // namespace

// Solve A x = b (A general, CSR FP64); x holds the initial guess.  Stops when
// ||D^-1 (b - A x)|| <= rtol ||D^-1 b|| (recursive residual, read every
// check_every iterations) or after max_iter iterations; iters_out / relres_out
// (host) report the result, restarts_out (host, optional) the breakdown restarts.
QDS_API int qds_bicgstab(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
                         const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
                         void* workspace, int64_t ws_bytes, int32_t* iters_out, double* relres_out,
                         int32_t* restarts_out, void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals && b && x && workspace && n > 0, "qds_bicgstab: null pointer");
  QDS_CHECK_ARG(ws_bytes >= qds_bicgstab_workspace(n), "qds_bicgstab: workspace too small");
  auto none = [](const double*, double*) { return QDS_OK; };
  return bicgstab_impl(n, nnz, indptr, indices, vals, b, x, rtol, max_iter, check_every, true, none, workspace,
                       iters_out, relres_out, restarts_out, qds_stream(stream));
}

// The same iteration right-preconditioned by one multigrid V(nu, nu)-cycle per
// product (qds_mg_vcycle, FP64 hierarchy, one column; levels[0] holds the n
// unknowns in lattice order, as for qds_cg_mg); the residual is measured unscaled:
// ||b - A x|| <= rtol ||b||.
QDS_API int qds_bicgstab_mg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                            const double* vals, const double* b, double* x, double rtol, int32_t max_iter,
                            int32_t check_every, int32_t n_levels, const qds_mg_level* levels, int32_t nu,
                            int32_t coarse_sweeps, double omega, void* workspace, int64_t ws_bytes,
                            int32_t* iters_out, double* relres_out, int32_t* restarts_out, void* stream) {
  QDS_CHECK_ARG(indptr && indices && vals && b && x && workspace && levels && n > 0, "qds_bicgstab_mg: null pointer");
  QDS_CHECK_ARG(n_levels >= 1 && n_levels <= 32 && levels[0].nx * levels[0].ny == n,
                "qds_bicgstab_mg: level 0 (%lld x %lld) does not hold the %lld unknowns", (long long)levels[0].nx,
                (long long)levels[0].ny, (long long)n);
  QDS_CHECK_ARG(ws_bytes >= qds_bicgstab_workspace(n), "qds_bicgstab_mg: workspace too small");
  qds_mg_level lv[32];
  for (int l = 0; l < n_levels; ++l) lv[l] = levels[l];
  lv[0].ldf = lv[0].ldu = 1;
  auto vcycle = [&](const double* f, double* u) {
    lv[0].f = const_cast<double*>(f);
    lv[0].u = u;
    return qds_mg_vcycle(qds::F64, n_levels, lv, 1, nu, nu, coarse_sweeps, omega, stream);
  };
  return bicgstab_impl(n, nnz, indptr, indices, vals, b, x, rtol, max_iter, check_every, false, vcycle, workspace,
                       iters_out, relres_out, restarts_out, qds_stream(stream));
}
