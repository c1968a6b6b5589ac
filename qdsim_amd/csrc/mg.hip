// Geometric multigrid on the DOF lattice, used as the block preconditioner of the
// eigensolver (no reference counterpart: the reference relies on ARPACK
// shift-invert with a sparse direct factorisation,
// qdsim_cython/advanced_eigenvalue_solvers.py:106-122).
//
// Every DOF of a P1/P2/P3 structured mesh sits on the lattice of the p-times
// refined grid, so the interior DOFs (homogeneous Dirichlet) form an nx x ny array.
// The preconditioner is the low-order-refined operator on that lattice: the P1
// stiffness of the refined triangulation, which for right triangles is exactly a
// five-point stencil with edge weights w = c * h_perp / h_par, plus a lumped mass
// term.  Per level: d[p] (diagonal, including the couplings to Dirichlet
// neighbours), we[p] (edge p -> east), wn[p] (edge p -> north), zero on edges that
// leave the interior.  Coarse levels come from vertex-centred coarsening,
// bilinear prolongation P and restriction R = P^T.  Per direction a level either
// has Dirichlet neighbours outside (coarse point j at fine 2j + 1, fine count
// 2 nc + 1: the eigensolver's interior lattice) or Neumann boundary rows that are
// unknowns themselves (coarse point j at fine 2j, fine count 2 nc - 1: the
// insulating top and bottom of the Poisson problem).
//
// Vectors hold b columns per lattice point (row-major point x column), matching
// the eigensolver's block layout.  One V(nu1, nu2)-cycle with damped Jacobi
// smoothing is launched by qds_mg_vcycle; all kernels run on the caller's stream.
#include "qds_common.h"

#include <type_traits>

namespace {

constexpr int kBlock = 256;

template <typename T>
struct Level {
  int64_t nx, ny;
  const T* d;
  const T* we;
  const T* wn;
  T* u;
  T* f;
  T* t;
  int64_t ldu, ldf;  // per-point strides of u and f (t is dense, stride b)
  int32_t nbx, nby;  // 1: coarsening of this level keeps its boundary rows (Neumann), 0: Dirichlet
  // batch of problems on one lattice (blockIdx.y): element strides between the
  // problems' d, we / wn, u, f and t (0: shared by every problem)
  int64_t sd, sw, su, sf, st;
};

// this workgroup's problem of the batch
template <typename T>
__device__ __forceinline__ void to_set(Level<T>& L) {
  const int64_t y = blockIdx.y;
  L.d += y * L.sd;
  L.we += y * L.sw;
  L.wn += y * L.sw;
  L.u += y * L.su;
  L.f += y * L.sf;
  L.t += y * L.st;
}

// Column vectors: each thread owns VEC consecutive columns of one lattice point,
// so every stencil access is a 16-byte load (32-bit index math throughout; the
// host guarantees np * b < 2^31).
template <typename T, int VEC> struct V;
template <> struct V<float, 4> { typedef float4 t; };
template <> struct V<float, 1> { typedef float t; };
template <> struct V<double, 2> { typedef double2 t; };
template <> struct V<double, 1> { typedef double t; };

template <typename T, int VEC>
struct Vec {
  T v[VEC];
  __device__ __forceinline__ void load(const T* p) {
    typename V<T, VEC>::t x = *reinterpret_cast<const typename V<T, VEC>::t*>(p);
    const T* q = reinterpret_cast<const T*>(&x);
#pragma unroll
    for (int i = 0; i < VEC; ++i) v[i] = q[i];
  }
  __device__ __forceinline__ void store(T* p) const {
    typename V<T, VEC>::t x;
    T* q = reinterpret_cast<T*>(&x);
#pragma unroll
    for (int i = 0; i < VEC; ++i) q[i] = v[i];
    *reinterpret_cast<typename V<T, VEC>::t*>(p) = x;
  }
};

// 1-D bilinear weights of fine index i: coarse point j sits at fine 2j + o
// (o = 1 Dirichlet, 0 Neumann); a coincident point has weight 1, a point between
// two coarse points 1/2 each (indices outside the coarse range are skipped by the
// caller)
__device__ __forceinline__ int coarse_nbrs(int i, int o, int* js) {
  const int r = i - o;
  if ((r & 1) == 0) {
    js[0] = r / 2;
    return 1;
  }
  js[0] = (r - 1) / 2;   // r >= -1, so (r - 1) / 2 >= -1
  js[1] = (r + 1) / 2;
  return 2;
}

// VEC values of a block stored in type S (the level-0 right-hand side and result
// may be FP64 blocks of an FP64 eigensolver phase while the hierarchy is FP32)
template <typename T, int VEC, typename S>
__device__ __forceinline__ void load_as(Vec<T, VEC>& v, const S* p) {
  if constexpr (std::is_same<T, S>::value) {
    v.load(p);
  } else {
#pragma unroll
    for (int i = 0; i < VEC; ++i) v.v[i] = static_cast<T>(p[i]);
  }
}

template <typename T, int VEC, typename S>
__device__ __forceinline__ void store_as(const Vec<T, VEC>& v, S* p) {
  if constexpr (std::is_same<T, S>::value) {
    v.store(p);
  } else {
#pragma unroll
    for (int i = 0; i < VEC; ++i) p[i] = static_cast<S>(v.v[i]);
  }
}

// (A u)[p, c0:c0+VEC]
template <typename T, int VEC>
__device__ __forceinline__ Vec<T, VEC> apply_at(const Level<T>& L, uint32_t ld, uint32_t p, uint32_t ix,
                                                uint32_t iy, uint32_t c0, const T* __restrict__ u) {
  Vec<T, VEC> s, x;
  s.load(u + p * ld + c0);
  const T d = L.d[p];
#pragma unroll
  for (int i = 0; i < VEC; ++i) s.v[i] *= d;
  if (ix + 1 < (uint32_t)L.nx) {
    const T w = L.we[p];
    x.load(u + (p + 1) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] -= w * x.v[i];
  }
  if (ix > 0) {
    const T w = L.we[p - 1];
    x.load(u + (p - 1) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] -= w * x.v[i];
  }
  if (iy + 1 < (uint32_t)L.ny) {
    const T w = L.wn[p];
    x.load(u + (p + (uint32_t)L.nx) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] -= w * x.v[i];
  }
  if (iy > 0) {
    const T w = L.wn[p - (uint32_t)L.nx];
    x.load(u + (p - (uint32_t)L.nx) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] -= w * x.v[i];
  }
  return s;
}

// Work-item loops of the three level operations.  Item t covers VEC columns of one
// lattice point; a launch walks items t0, t0 + step, ... (grid-stride in the
// multi-block kernels, block-stride in the single-block coarse cycle).

// out = u + omega D^-1 (f - A u); zero_init: u = 0 so out = omega D^-1 f
template <typename T, int VEC>
__device__ __forceinline__ void jacobi_items(const Level<T>& L, int b, const T* __restrict__ u, int64_t ldu_in,
                                             T* __restrict__ out, int64_t ldo, T omega, int zero_init,
                                             uint32_t t0, uint32_t step) {
  const uint32_t nv = (uint32_t)(b / VEC);
  const uint32_t total = (uint32_t)(L.nx * L.ny) * nv;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t p = t / nv;
    const uint32_t c0 = (t - p * nv) * VEC;
    const T sc = omega / L.d[p];
    Vec<T, VEC> fv;
    fv.load(L.f + p * (uint32_t)L.ldf + c0);
    if (zero_init) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) fv.v[i] *= sc;
      fv.store(out + p * (uint32_t)ldo + c0);
    } else {
      const uint32_t iy = p / (uint32_t)L.nx, ix = p - iy * (uint32_t)L.nx;
      Vec<T, VEC> au = apply_at<T, VEC>(L, (uint32_t)ldu_in, p, ix, iy, c0, u);
      Vec<T, VEC> uv;
      uv.load(u + p * (uint32_t)ldu_in + c0);
#pragma unroll
      for (int i = 0; i < VEC; ++i) uv.v[i] += sc * (fv.v[i] - au.v[i]);
      uv.store(out + p * (uint32_t)ldo + c0);
    }
  }
}

// coarse f = R (f - A u), R = P^T of bilinear interpolation
template <typename T, int VEC>
__device__ __forceinline__ void restrict_items(const Level<T>& F, const Level<T>& C, int b, uint32_t t0,
                                               uint32_t step) {
  const uint32_t nv = (uint32_t)(b / VEC);
  const uint32_t total = (uint32_t)(C.nx * C.ny) * nv;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t pc = t / nv;
    const uint32_t c0 = (t - pc * nv) * VEC;
    const uint32_t jy = pc / (uint32_t)C.nx, jx = pc - jy * (uint32_t)C.nx;
    const uint32_t fx = 2 * jx + 1 - F.nbx, fy = 2 * jy + 1 - F.nby;
    Vec<T, VEC> acc;
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc.v[i] = T(0);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const uint32_t x = fx + dx, y = fy + dy;
        if (x >= (uint32_t)F.nx || y >= (uint32_t)F.ny) continue;   // outside a Neumann boundary
        const T w = (dx == 0 ? T(1) : T(0.5)) * (dy == 0 ? T(1) : T(0.5));
        const uint32_t p = y * (uint32_t)F.nx + x;
        Vec<T, VEC> fv;
        fv.load(F.f + p * (uint32_t)F.ldf + c0);
        Vec<T, VEC> au = apply_at<T, VEC>(F, (uint32_t)F.ldu, p, x, y, c0, F.u);
#pragma unroll
        for (int i = 0; i < VEC; ++i) acc.v[i] += w * (fv.v[i] - au.v[i]);
      }
    }
    acc.store(C.f + pc * (uint32_t)C.ldf + c0);
  }
}

// fine u += P e (bilinear); coarse point (jx, jy) sits at fine (2jx+1, 2jy+1)
template <typename T, int VEC>
__device__ __forceinline__ void prolong_items(const Level<T>& F, const Level<T>& C, int b, uint32_t t0,
                                              uint32_t step) {
  const uint32_t nv = (uint32_t)(b / VEC);
  const uint32_t total = (uint32_t)(F.nx * F.ny) * nv;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t p = t / nv;
    const uint32_t c0 = (t - p * nv) * VEC;
    const uint32_t iy = p / (uint32_t)F.nx, ix = p - iy * (uint32_t)F.nx;
    int xs[2], ys[2];
    const int nxp = coarse_nbrs((int)ix, 1 - F.nbx, xs);
    const int nyp = coarse_nbrs((int)iy, 1 - F.nby, ys);
    const T wx = nxp == 1 ? T(1) : T(0.5), wy = nyp == 1 ? T(1) : T(0.5);
    Vec<T, VEC> acc, e;
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc.v[i] = T(0);
    for (int a = 0; a < nyp; ++a) {
      if (ys[a] < 0 || ys[a] >= (int)C.ny) continue;
      for (int q = 0; q < nxp; ++q) {
        if (xs[q] < 0 || xs[q] >= (int)C.nx) continue;
        e.load(C.u + ((uint32_t)ys[a] * (uint32_t)C.nx + (uint32_t)xs[q]) * (uint32_t)C.ldu + c0);
        const T w = wx * wy;
#pragma unroll
        for (int i = 0; i < VEC; ++i) acc.v[i] += w * e.v[i];
      }
    }
    Vec<T, VEC> uv;
    uv.load(F.u + p * (uint32_t)F.ldu + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) uv.v[i] += acc.v[i];
    uv.store(F.u + p * (uint32_t)F.ldu + c0);
  }
}

// dense copy of u into t (first sweep of an odd count on a live u)
template <typename T>
__device__ __forceinline__ void copy_items(int64_t np, int b, const T* __restrict__ src, int64_t lds,
                                           T* __restrict__ dst, uint32_t t0, uint32_t step) {
  const uint32_t total = (uint32_t)np * (uint32_t)b;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t p = t / (uint32_t)b;
    const uint32_t c = t - p * (uint32_t)b;
    dst[t] = src[p * (uint32_t)lds + c];
  }
}

#define QDS_GRID_T0 (blockIdx.x * blockDim.x + threadIdx.x)
#define QDS_GRID_STEP (gridDim.x * blockDim.x)

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
jacobi_kernel(Level<T> L, int b, const T* __restrict__ u, int64_t ldu_in, T* __restrict__ out,
              int64_t ldo, T omega, int zero_init, int64_t s_in, int64_t s_out) {
  to_set(L);
  u += blockIdx.y * s_in;
  out += blockIdx.y * s_out;
  jacobi_items<T, VEC>(L, b, u, ldu_in, out, ldo, omega, zero_init, QDS_GRID_T0, QDS_GRID_STEP);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
restrict_residual_kernel(Level<T> F, Level<T> C, int b) {
  to_set(F);
  to_set(C);
  restrict_items<T, VEC>(F, C, b, QDS_GRID_T0, QDS_GRID_STEP);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kBlock)
prolong_add_kernel(Level<T> F, Level<T> C, int b) {
  to_set(F);
  to_set(C);
  prolong_items<T, VEC>(F, C, b, QDS_GRID_T0, QDS_GRID_STEP);
}

// ---- fused V(1,1) passes ---------------------------------------------------
// With a zero initial guess and one damped-Jacobi sweep before and after the
// coarse correction, a level needs only two passes over its lattice:
//   down:  u1 = omega D^-1 f,  r = f - A u1 = (1 - omega) f + omega sum_q w_pq f_q / d_q,
//          coarse f = R r                        (reads f; u1 is never stored)
//   up:    u2 = u1 + P e,  u = u2 + omega D^-1 (f - A u2)
//          (reads f and the coarse correction e; writes u)
// This is the same arithmetic as the sweep/restrict/prolong/sweep sequence.

// sum over the lattice neighbours q of p of w_pq * src_q / d_q
template <typename T, int VEC, typename S>
__device__ __forceinline__ Vec<T, VEC> nbr_scaled(const Level<T>& L, uint32_t ld, uint32_t p, uint32_t ix,
                                                  uint32_t iy, uint32_t c0, const S* __restrict__ src) {
  Vec<T, VEC> s, x;
#pragma unroll
  for (int i = 0; i < VEC; ++i) s.v[i] = T(0);
  const uint32_t nx = (uint32_t)L.nx;
  if (ix + 1 < nx) {
    const T w = L.we[p] / L.d[p + 1];
    load_as<T, VEC, S>(x, src + (p + 1) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] += w * x.v[i];
  }
  if (ix > 0) {
    const T w = L.we[p - 1] / L.d[p - 1];
    load_as<T, VEC, S>(x, src + (p - 1) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] += w * x.v[i];
  }
  if (iy + 1 < (uint32_t)L.ny) {
    const T w = L.wn[p] / L.d[p + nx];
    load_as<T, VEC, S>(x, src + (p + nx) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] += w * x.v[i];
  }
  if (iy > 0) {
    const T w = L.wn[p - nx] / L.d[p - nx];
    load_as<T, VEC, S>(x, src + (p - nx) * ld + c0);
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] += w * x.v[i];
  }
  return s;
}

// (P e)[p] for fine point (ix, iy) of level F: bilinear weights of the coarse
// points around it
template <typename T, int VEC>
__device__ __forceinline__ Vec<T, VEC> prolong_at(const Level<T>& C, int nbx, int nby, uint32_t ix, uint32_t iy,
                                                  uint32_t c0) {
  int xs[2], ys[2];
  const int nxp = coarse_nbrs((int)ix, 1 - nbx, xs);
  const int nyp = coarse_nbrs((int)iy, 1 - nby, ys);
  const T wx = nxp == 1 ? T(1) : T(0.5), wy = nyp == 1 ? T(1) : T(0.5);
  Vec<T, VEC> acc, e;
#pragma unroll
  for (int i = 0; i < VEC; ++i) acc.v[i] = T(0);
  for (int a = 0; a < nyp; ++a) {
    if (ys[a] < 0 || ys[a] >= (int)C.ny) continue;
    for (int q = 0; q < nxp; ++q) {
      if (xs[q] < 0 || xs[q] >= (int)C.nx) continue;
      e.load(C.u + ((uint32_t)ys[a] * (uint32_t)C.nx + (uint32_t)xs[q]) * (uint32_t)C.ldu + c0);
      const T w = wx * wy;
#pragma unroll
      for (int i = 0; i < VEC; ++i) acc.v[i] += w * e.v[i];
    }
  }
  return acc;
}

template <typename T, int VEC, typename S = T>
__device__ __forceinline__ void down_items(const Level<T>& F, const Level<T>& C, int b, T omega, uint32_t t0,
                                           uint32_t step, const S* __restrict__ ff = nullptr) {
  if constexpr (std::is_same<T, S>::value) {
    if (!ff) ff = F.f;
  }
  const uint32_t nv = (uint32_t)(b / VEC);
  const uint32_t total = (uint32_t)(C.nx * C.ny) * nv;
  const uint32_t ldf = (uint32_t)F.ldf;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t pc = t / nv;
    const uint32_t c0 = (t - pc * nv) * VEC;
    const uint32_t jy = pc / (uint32_t)C.nx, jx = pc - jy * (uint32_t)C.nx;
    Vec<T, VEC> acc;
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc.v[i] = T(0);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        // fine neighbour of the coarse point; beside a Neumann boundary row the
        // outer neighbour does not exist (unsigned wrap makes it >= nx / ny)
        const uint32_t x = 2 * jx + 1 - F.nbx + dx, y = 2 * jy + 1 - F.nby + dy;
        if (x >= (uint32_t)F.nx || y >= (uint32_t)F.ny) continue;
        const T w = (dx == 0 ? T(1) : T(0.5)) * (dy == 0 ? T(1) : T(0.5));
        const uint32_t p = y * (uint32_t)F.nx + x;
        Vec<T, VEC> fv;
        load_as<T, VEC, S>(fv, ff + p * ldf + c0);
        const Vec<T, VEC> ns = nbr_scaled<T, VEC, S>(F, ldf, p, x, y, c0, ff);
#pragma unroll
        for (int i = 0; i < VEC; ++i) acc.v[i] += w * ((T(1) - omega) * fv.v[i] + omega * ns.v[i]);
      }
    }
    acc.store(C.f + pc * (uint32_t)C.ldf + c0);
  }
}

template <typename T, int VEC, typename S = T>
__device__ __forceinline__ void up_items(const Level<T>& F, const Level<T>& C, int b, T omega, uint32_t t0,
                                         uint32_t step, const S* __restrict__ ff = nullptr, S* __restrict__ uu = nullptr) {
  if constexpr (std::is_same<T, S>::value) {
    if (!ff) ff = F.f;
    if (!uu) uu = F.u;
  }
  const uint32_t nv = (uint32_t)(b / VEC);
  const uint32_t total = (uint32_t)(F.nx * F.ny) * nv;
  const uint32_t ldf = (uint32_t)F.ldf;
  const uint32_t nx = (uint32_t)F.nx;
  for (uint32_t t = t0; t < total; t += step) {
    const uint32_t p = t / nv;
    const uint32_t c0 = (t - p * nv) * VEC;
    const uint32_t iy = p / nx, ix = p - iy * nx;
    // u2 at p and its neighbours, then one Jacobi step at p
    auto u2 = [&](uint32_t q, uint32_t qx, uint32_t qy) {
      Vec<T, VEC> fv, e = prolong_at<T, VEC>(C, F.nbx, F.nby, qx, qy, c0);
      load_as<T, VEC, S>(fv, ff + q * ldf + c0);
      const T sc = omega / F.d[q];
#pragma unroll
      for (int i = 0; i < VEC; ++i) e.v[i] += sc * fv.v[i];
      return e;
    };
    Vec<T, VEC> up = u2(p, ix, iy), fp, s;
    load_as<T, VEC, S>(fp, ff + p * ldf + c0);
    const T dp = F.d[p];
#pragma unroll
    for (int i = 0; i < VEC; ++i) s.v[i] = fp.v[i] - dp * up.v[i];
    if (ix + 1 < nx) {
      const T w = F.we[p];
      const Vec<T, VEC> uq = u2(p + 1, ix + 1, iy);
#pragma unroll
      for (int i = 0; i < VEC; ++i) s.v[i] += w * uq.v[i];
    }
    if (ix > 0) {
      const T w = F.we[p - 1];
      const Vec<T, VEC> uq = u2(p - 1, ix - 1, iy);
#pragma unroll
      for (int i = 0; i < VEC; ++i) s.v[i] += w * uq.v[i];
    }
    if (iy + 1 < (uint32_t)F.ny) {
      const T w = F.wn[p];
      const Vec<T, VEC> uq = u2(p + nx, ix, iy + 1);
#pragma unroll
      for (int i = 0; i < VEC; ++i) s.v[i] += w * uq.v[i];
    }
    if (iy > 0) {
      const T w = F.wn[p - nx];
      const Vec<T, VEC> uq = u2(p - nx, ix, iy - 1);
#pragma unroll
      for (int i = 0; i < VEC; ++i) s.v[i] += w * uq.v[i];
    }
    const T sc = omega / dp;
#pragma unroll
    for (int i = 0; i < VEC; ++i) up.v[i] += sc * s.v[i];
    store_as<T, VEC, S>(up, uu + p * (uint32_t)F.ldu + c0);
  }
}

template <typename T, int VEC, typename S>
__global__ void __launch_bounds__(kBlock)
mg_down_kernel(Level<T> F, Level<T> C, int b, T omega, const S* __restrict__ ff, int64_t sff) {
  to_set(F);
  to_set(C);
  ff += blockIdx.y * sff;
  down_items<T, VEC, S>(F, C, b, omega, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, ff);
}

template <typename T, int VEC, typename S>
__global__ void __launch_bounds__(kBlock)
mg_up_kernel(Level<T> F, Level<T> C, int b, T omega, const S* __restrict__ ff, S* __restrict__ uu, int64_t sff,
             int64_t suu) {
  to_set(F);
  to_set(C);
  ff += blockIdx.y * sff;
  uu += blockIdx.y * suu;
  up_items<T, VEC, S>(F, C, b, omega, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, ff, uu);
}

// ---- LDS-tiled fused passes for large levels ---------------------------------
// The grid-stride passes above re-read f, d, we, wn of every neighbour from global
// memory (the down pass ~45 16-byte loads and ~36 scalar loads per coarse item,
// the up pass ~27 per fine item): on the level-0 lattice of the 2.36M-DOF problem
// they ran at 1.1-1.7 TB/s, bound by the address unit.  The tiled passes stage a
// small tile with its halo in LDS -- all b columns of a point, loaded as
// consecutive 16-byte chunks so every wave reads whole contiguous row segments --
// and then compute every item from LDS.  Same arithmetic as down_items / up_items.
constexpr int kDX = 4, kDY = 8;            // coarse tile of the down pass
constexpr int kDFX = 2 * kDX + 3, kDFY = 2 * kDY + 3;  // its fine f tile (with halo)
constexpr int kUX = 8, kUY = 8;            // fine tile of the up pass
constexpr int kUFX = kUX + 2, kUFY = kUY + 2;          // its halo tile
constexpr int kUCX = kUX / 2 + 2, kUCY = kUY / 2 + 2;  // its coarse e tile

// rows of 16-byte chunks: tile points (i, j) of a (w x h) window starting at
// (x0, y0) of a (nx x ny) lattice, b columns each; zero outside the lattice
template <typename T, typename S>
__device__ __forceinline__ void stage_window(T* dst, const S* __restrict__ src, int64_t ld, int nx, int ny,
                                             int x0, int y0, int w, int h, int b) {
  constexpr int CE = 16 / sizeof(T);
  const int cpp = b / CE;
  const int total = w * h * cpp;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int pt = e / cpp;
    const int part = e - pt * cpp;
    const int i = pt % w, j = pt / w;
    const int x = x0 + i, y = y0 + j;
    Vec<T, CE> v;
#pragma unroll
    for (int q = 0; q < CE; ++q) v.v[q] = T(0);
    if (x >= 0 && x < nx && y >= 0 && y < ny) {
      const S* p = src + ((int64_t)y * nx + x) * ld + part * CE;
      if constexpr (std::is_same<T, S>::value) {
        v.load(p);
      } else {
        // FP64 block into an FP32 tile: two 16-byte loads per 16-byte chunk
        Vec<S, 2> a, c;
        a.load(p);
        c.load(p + 2);
        v.v[0] = static_cast<T>(a.v[0]);
        v.v[1] = static_cast<T>(a.v[1]);
        v.v[2] = static_cast<T>(c.v[0]);
        v.v[3] = static_cast<T>(c.v[1]);
      }
    }
    v.store(dst + (int64_t)pt * b + part * CE);
  }
}

// down: coarse f on a kDX x kDY coarse tile
template <typename T, int VEC, typename S>
__global__ void __launch_bounds__(256)
mg_down_tiled(Level<T> F, Level<T> C, int b, T omega, const S* __restrict__ ff, int64_t sff) {
  to_set(F);
  to_set(C);
  ff += blockIdx.y * sff;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* fs = reinterpret_cast<T*>(lds_raw);                  // kDFX * kDFY points x b
  T* cinv = fs + kDFX * kDFY * b;                          // 1 / d
  T* cwe = cinv + kDFX * kDFY;
  T* cwn = cwe + kDFX * kDFY;
  const int cnx = (int)C.nx, cny = (int)C.ny, fnx = (int)F.nx, fny = (int)F.ny;
  const int ntx = (cnx + kDX - 1) / kDX;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int cx0 = tx * kDX, cy0 = ty * kDY;
  const int fx0 = 2 * cx0 - 1, fy0 = 2 * cy0 - 1;   // fine point of tile entry (0, 0)
  for (int e = threadIdx.x; e < kDFX * kDFY; e += blockDim.x) {
    const int x = fx0 + e % kDFX, y = fy0 + e / kDFX;
    T inv = T(0), we = T(0), wn = T(0);
    if (x >= 0 && x < fnx && y >= 0 && y < fny) {
      const uint32_t p = (uint32_t)y * (uint32_t)fnx + (uint32_t)x;
      inv = T(1) / F.d[p];
      we = F.we[p];
      wn = F.wn[p];
    }
    cinv[e] = inv;
    cwe[e] = we;
    cwn[e] = wn;
  }
  stage_window<T, S>(fs, ff, F.ldf, fnx, fny, fx0, fy0, kDFX, kDFY, b);
  __syncthreads();
  // r = (1 - omega) f + omega sum_q w_pq f_q / d_q on the inner (kDFX-2) x (kDFY-2)
  // points, kept in registers until every thread has read its neighbours
  const int nv = b / VEC;
  constexpr int kIn = (kDFX - 2) * (kDFY - 2);
  constexpr int kPer = (kIn * (32 / VEC) + 255) / 256;   // items per thread for b <= 32
  Vec<T, VEC> rr[kPer];
  int rq[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int it = threadIdx.x + 256 * k;
    rq[k] = -1;
    if (it < kIn * nv) {
      const int pt = it / nv;
      const int c0 = (it - pt * nv) * VEC;
      const int i = 1 + pt % (kDFX - 2), j = 1 + pt / (kDFX - 2);
      const int q = j * kDFX + i;
      Vec<T, VEC> fc, fe, fw, fn, fsou;
      fc.load(fs + q * b + c0);
      fe.load(fs + (q + 1) * b + c0);
      fw.load(fs + (q - 1) * b + c0);
      fn.load(fs + (q + kDFX) * b + c0);
      fsou.load(fs + (q - kDFX) * b + c0);
      const T ae = cwe[q] * cinv[q + 1], aw = cwe[q - 1] * cinv[q - 1];
      const T an = cwn[q] * cinv[q + kDFX], as = cwn[q - kDFX] * cinv[q - kDFX];
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        rr[k].v[v] = (T(1) - omega) * fc.v[v] + omega * (ae * fe.v[v] + aw * fw.v[v] + an * fn.v[v] + as * fsou.v[v]);
      rq[k] = q * b + c0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (rq[k] >= 0) rr[k].store(fs + rq[k]);
  __syncthreads();
  for (int it = threadIdx.x; it < kDX * kDY * nv; it += blockDim.x) {
    const int pc = it / nv;
    const int c0 = (it - pc * nv) * VEC;
    const int ci = pc % kDX, cj = pc / kDX;
    const int jx = cx0 + ci, jy = cy0 + cj;
    if (jx >= cnx || jy >= cny) continue;
    // fine centre (2 jx + 1, 2 jy + 1) is tile entry (2 ci + 2, 2 cj + 2)
    Vec<T, VEC> acc;
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc.v[v] = T(0);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const T w = (dx == 0 ? T(1) : T(0.5)) * (dy == 0 ? T(1) : T(0.5));
        Vec<T, VEC> r;
        r.load(fs + ((2 * cj + 2 + dy) * kDFX + 2 * ci + 2 + dx) * b + c0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc.v[v] += w * r.v[v];
      }
    }
    acc.store(C.f + ((uint32_t)jy * (uint32_t)cnx + (uint32_t)jx) * (uint32_t)C.ldf + c0);
  }
}

// up: fine u on a kUX x kUY fine tile
template <typename T, int VEC, typename S>
__global__ void __launch_bounds__(256)
mg_up_tiled(Level<T> F, Level<T> C, int b, T omega, const S* __restrict__ ff, S* __restrict__ uu, int64_t sff,
            int64_t suu) {
  to_set(F);
  to_set(C);
  ff += blockIdx.y * sff;
  uu += blockIdx.y * suu;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* fs = reinterpret_cast<T*>(lds_raw);                  // kUFX * kUFY points x b
  T* us = fs + kUFX * kUFY * b;                            // u2 on the halo tile
  T* es = us + kUFX * kUFY * b;                            // coarse e, kUCX * kUCY points x b
  T* cd = es + kUCX * kUCY * b;
  T* cinv = cd + kUFX * kUFY;
  T* cwe = cinv + kUFX * kUFY;
  T* cwn = cwe + kUFX * kUFY;
  const int fnx = (int)F.nx, fny = (int)F.ny, cnx = (int)C.nx, cny = (int)C.ny;
  const int ntx = (fnx + kUX - 1) / kUX;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int fx0 = tx * kUX, fy0 = ty * kUY;              // tile origin (even)
  const int hx0 = fx0 - 1, hy0 = fy0 - 1;                // halo origin
  const int ex0 = fx0 / 2 - 1, ey0 = fy0 / 2 - 1;        // coarse window origin
  for (int e = threadIdx.x; e < kUFX * kUFY; e += blockDim.x) {
    const int x = hx0 + e % kUFX, y = hy0 + e / kUFX;
    T d = T(0), inv = T(0), we = T(0), wn = T(0);
    if (x >= 0 && x < fnx && y >= 0 && y < fny) {
      const uint32_t p = (uint32_t)y * (uint32_t)fnx + (uint32_t)x;
      d = F.d[p];
      inv = T(1) / d;
      we = F.we[p];
      wn = F.wn[p];
    }
    cd[e] = d;
    cinv[e] = inv;
    cwe[e] = we;
    cwn[e] = wn;
  }
  stage_window<T, S>(fs, ff, F.ldf, fnx, fny, hx0, hy0, kUFX, kUFY, b);
  stage_window<T, T>(es, C.u, C.ldu, cnx, cny, ex0, ey0, kUCX, kUCY, b);
  __syncthreads();
  const int nv = b / VEC;
  // u2 = P e + omega f / d on the halo tile (0 outside the lattice)
  for (int it = threadIdx.x; it < kUFX * kUFY * nv; it += blockDim.x) {
    const int e = it / nv;
    const int c0 = (it - e * nv) * VEC;
    const int x = hx0 + e % kUFX, y = hy0 + e / kUFX;
    Vec<T, VEC> acc;
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc.v[v] = T(0);
    if (x >= 0 && x < fnx && y >= 0 && y < fny) {
      int xs0, ys0, nxp, nyp;
      T wx, wy;
      if (x & 1) { xs0 = (x - 1) / 2 - ex0; nxp = 1; wx = T(1); }
      else { xs0 = x / 2 - 1 - ex0; nxp = 2; wx = T(0.5); }
      if (y & 1) { ys0 = (y - 1) / 2 - ey0; nyp = 1; wy = T(1); }
      else { ys0 = y / 2 - 1 - ey0; nyp = 2; wy = T(0.5); }
      for (int a = 0; a < nyp; ++a)
        for (int q = 0; q < nxp; ++q) {
          Vec<T, VEC> ev;
          ev.load(es + ((ys0 + a) * kUCX + xs0 + q) * b + c0);
#pragma unroll
          for (int v = 0; v < VEC; ++v) acc.v[v] += (wx * wy) * ev.v[v];
        }
      Vec<T, VEC> fv;
      fv.load(fs + e * b + c0);
      const T sc = omega * cinv[e];
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc.v[v] += sc * fv.v[v];
    }
    acc.store(us + e * b + c0);
  }
  __syncthreads();
  for (int it = threadIdx.x; it < kUX * kUY * nv; it += blockDim.x) {
    const int pt = it / nv;
    const int c0 = (it - pt * nv) * VEC;
    const int li = pt % kUX, lj = pt / kUX;
    const int px = fx0 + li, py = fy0 + lj;
    if (px >= fnx || py >= fny) continue;
    const int q = (lj + 1) * kUFX + li + 1;
    Vec<T, VEC> up, fp, ue, uw, un, uso;
    up.load(us + q * b + c0);
    fp.load(fs + q * b + c0);
    ue.load(us + (q + 1) * b + c0);
    uw.load(us + (q - 1) * b + c0);
    un.load(us + (q + kUFX) * b + c0);
    uso.load(us + (q - kUFX) * b + c0);
    const T dp = cd[q], sc = omega * cinv[q];
    const T we = cwe[q], ww = cwe[q - 1], wn = cwn[q], ws = cwn[q - kUFX];
    Vec<T, VEC> out;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      const T s = fp.v[v] - dp * up.v[v] + we * ue.v[v] + ww * uw.v[v] + wn * un.v[v] + ws * uso.v[v];
      out.v[v] = up.v[v] + sc * s;
    }
    store_as<T, VEC, S>(out, uu + ((uint32_t)py * (uint32_t)fnx + (uint32_t)px) * (uint32_t)F.ldu + c0);
  }
}

template <typename T>
size_t down_tiled_lds(int b) { return sizeof(T) * ((size_t)kDFX * kDFY * b + 3 * kDFX * kDFY); }
template <typename T>
size_t up_tiled_lds(int b) {
  return sizeof(T) * ((size_t)2 * kUFX * kUFY * b + (size_t)kUCX * kUCY * b + 4 * kUFX * kUFY);
}
// the tiled passes need whole 16-byte chunks per point, b <= 32 and a level of at
// least a few tiles; smaller or odd-width levels keep the grid-stride passes
template <typename T>
bool use_tiled(const Level<T>& L, int b, int vec) {
  // the tiled passes assume Dirichlet coarsening in both directions
  return vec * (int)sizeof(T) == 16 && b % vec == 0 && b <= 32 && L.nx >= 32 && L.ny >= 32 && !L.nbx && !L.nby;
}

// Levels handled by the single-workgroup coarse cycle (level l0 and below).
constexpr int kMaxCoarseLevels = 12;
constexpr int kCoarseBlock = 1024;
constexpr int64_t kCoarseItems = 24576;  // work items of the finest level it takes
// and lattice points: the item cap alone handed a 150 x 150 level of the Poisson
// preconditioner (one column) to one workgroup (290 us per cycle, most of a
// Poisson solve) and 4-6k points of the eigensolver's (b = 16-24) to one
// workgroup (180-250 us); 2048 points measured best (512: same within 1%)
constexpr int64_t kCoarsePoints = 2048;

template <typename T>
struct LevelPack {
  Level<T> L[kMaxCoarseLevels];
};

// Jacobi sweeps on one level within one workgroup, same ping-pong schedule as the
// multi-launch path (the result ends in u).
template <typename T, int VEC>
__device__ void block_smooth(const Level<T>& L, int b, int sweeps, T omega, bool zero_first) {
  T* src = L.u;
  T* dst = (sweeps % 2 == 0) ? L.t : L.u;
  for (int s = 0; s < sweeps; ++s) {
    const bool zi = zero_first && s == 0;
    if (!zi && src == dst) {
      copy_items<T>(L.nx * L.ny, b, L.u, L.ldu, L.t, threadIdx.x, blockDim.x);
      __syncthreads();
      src = L.t;
    }
    jacobi_items<T, VEC>(L, b, src, src == L.u ? L.ldu : (int64_t)b, dst, dst == L.u ? L.ldu : (int64_t)b,
                         omega, zi ? 1 : 0, threadIdx.x, blockDim.x);
    __syncthreads();
    src = dst;
    dst = (dst == L.u) ? L.t : L.u;
  }
}

// The V-cycle below (and including) pack level 0, whose f is already set: one
// workgroup, phases separated by workgroup barriers (global-memory results of a
// phase are visible to the whole workgroup after the barrier).
template <typename T, int VEC>
__global__ void __launch_bounds__(kCoarseBlock)
coarse_cycle_kernel(LevelPack<T> P, int nl, int b, int nu1, int nu2, int ncoarse, T omega) {
  // this workgroup's problem of the batch: local copies of its levels
  auto lev = [&](int l) {
    Level<T> L = P.L[l];
    to_set(L);
    return L;
  };
  if (nu1 == 1 && nu2 == 1) {
    for (int l = 0; l < nl - 1; ++l) {
      const Level<T> F = lev(l), C = lev(l + 1);
      down_items<T, VEC>(F, C, b, omega, threadIdx.x, blockDim.x);
      __syncthreads();
    }
    block_smooth<T, VEC>(lev(nl - 1), b, ncoarse, omega, true);
    for (int l = nl - 2; l >= 0; --l) {
      const Level<T> F = lev(l), C = lev(l + 1);
      up_items<T, VEC>(F, C, b, omega, threadIdx.x, blockDim.x);
      __syncthreads();
    }
    return;
  }
  for (int l = 0; l < nl - 1; ++l) {
    const Level<T> F = lev(l), C = lev(l + 1);
    block_smooth<T, VEC>(F, b, nu1, omega, true);
    restrict_items<T, VEC>(F, C, b, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  block_smooth<T, VEC>(lev(nl - 1), b, ncoarse, omega, true);
  for (int l = nl - 2; l >= 0; --l) {
    const Level<T> F = lev(l), C = lev(l + 1);
    prolong_items<T, VEC>(F, C, b, threadIdx.x, blockDim.x);
    __syncthreads();
    block_smooth<T, VEC>(F, b, nu2, omega, false);
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
apply_kernel(Level<T> L, int b, const T* __restrict__ u, T* __restrict__ out) {
  to_set(L);
  const uint32_t total = (uint32_t)(L.nx * L.ny) * (uint32_t)b;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const uint32_t p = t / (uint32_t)b;
    const uint32_t c = t - p * (uint32_t)b;
    const uint32_t iy = p / (uint32_t)L.nx, ix = p - iy * (uint32_t)L.nx;
    out[t] = apply_at<T, 1>(L, (uint32_t)b, p, ix, iy, c, u).v[0];
  }
}

template <typename T>
__global__ void copy_kernel(int64_t np, int b, const T* __restrict__ src, int64_t lds, T* __restrict__ dst,
                            int64_t s_src, int64_t s_dst) {
  src += blockIdx.y * s_src;
  dst += blockIdx.y * s_dst;
  const uint32_t total = (uint32_t)np * (uint32_t)b;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const uint32_t p = t / (uint32_t)b;
    const uint32_t c = t - p * (uint32_t)b;
    dst[t] = src[p * (uint32_t)lds + c];
  }
}

template <typename T, int VEC>
int smooth(const Level<T>& L, int b, int sweeps, T omega, bool zero_first, hipStream_t st, int batch) {
  // Jacobi sweeps ping-pong between u and t; the result always ends in u
  const int64_t work = L.nx * L.ny * (b / VEC);
  const dim3 g(qds_blocks(work, kBlock, 65536), batch), blk(kBlock);
  auto s_of = [&](const T* q) { return q == L.u ? L.su : L.st; };
  // with an odd count the first sweep writes straight to u, so no final copy
  T* src = L.u;
  T* dst = (sweeps % 2 == 0) ? L.t : L.u;
  auto ld_of = [&](const T* q) { return q == L.u ? L.ldu : (int64_t)b; };
  for (int s = 0; s < sweeps; ++s) {
    const bool zi = zero_first && s == 0;
    if (!zi && src == dst) {
      // first sweep of an odd count on a live u: copy u into t (dense) first
      hipLaunchKernelGGL(copy_kernel<T>, dim3(qds_blocks(L.nx * L.ny * b, kBlock, 65536), batch), blk, 0, st,
                         L.nx * L.ny, b, (const T*)L.u, L.ldu, L.t, L.su, L.st);
      QDS_LAUNCH_CHECK();
      src = L.t;
    }
    hipLaunchKernelGGL((jacobi_kernel<T, VEC>), g, blk, 0, st, L, b, (const T*)src, ld_of(src), dst,
                       ld_of(dst), omega, zi ? 1 : 0, s_of(src), s_of(dst));
    QDS_LAUNCH_CHECK();
    src = dst;
    dst = (dst == L.u) ? L.t : L.u;
  }
  return QDS_OK;
}

// S: element type of level 0's right-hand side and result (T, or FP64 with an
// FP32 hierarchy; the mixed form needs the fused V(1,1) with level 0 outside the
// single-workgroup coarse cycle)
template <typename T, int VEC, typename S = T>
int vcycle(int nl, Level<T>* L, int b, int nu1, int nu2, int ncoarse, T omega, hipStream_t st,
           const S* f0 = nullptr, S* u0 = nullptr, int batch = 1, int64_t sf0 = 0, int64_t su0 = 0) {
  // first level small enough for the single-workgroup coarse cycle
  int lc = nl;
  for (int l = 0; l < nl; ++l)
    if (L[l].nx * L[l].ny * (b / VEC) <= kCoarseItems && L[l].nx * L[l].ny <= kCoarsePoints &&
        nl - l <= kMaxCoarseLevels) {
      lc = l;
      break;
    }
  const bool fused = nu1 == 1 && nu2 == 1;
  constexpr bool mixed = !std::is_same<T, S>::value;
  if (mixed)
    QDS_CHECK_ARG(fused && lc > 0 && nl > 1 && f0 && u0,
                  "qds_mg_vcycle_io: a level-0 block type other than the hierarchy's needs the fused "
                  "V(1,1) cycle and a level 0 outside the single-workgroup coarse cycle");
  // the tiled passes read level 0's FP64 blocks with 16-byte loads
  const bool io16 = !mixed || (reinterpret_cast<uintptr_t>(f0) % 16 == 0 && L[0].ldf % 2 == 0);
  for (int l = 0; l < lc && l < nl - 1; ++l) {
    const int64_t work = L[l + 1].nx * L[l + 1].ny * (b / VEC);
    if (fused) {
      if (use_tiled(L[l], b, VEC) && (l > 0 || io16)) {
        const unsigned tiles = (unsigned)(((L[l + 1].nx + kDX - 1) / kDX) * ((L[l + 1].ny + kDY - 1) / kDY));
        if (l == 0 && mixed)
          hipLaunchKernelGGL((mg_down_tiled<T, VEC, S>), dim3(tiles, batch), dim3(256), down_tiled_lds<T>(b), st,
                             L[l], L[l + 1], b, omega, f0, sf0);
        else
          hipLaunchKernelGGL((mg_down_tiled<T, VEC, T>), dim3(tiles, batch), dim3(256), down_tiled_lds<T>(b), st,
                             L[l], L[l + 1], b, omega, (const T*)L[l].f, L[l].sf);
      } else if (l == 0 && mixed) {
        hipLaunchKernelGGL((mg_down_kernel<T, VEC, S>), dim3(qds_blocks(work, kBlock, 65536), batch), dim3(kBlock),
                           0, st, L[l], L[l + 1], b, omega, f0, sf0);
      } else {
        hipLaunchKernelGGL((mg_down_kernel<T, VEC, T>), dim3(qds_blocks(work, kBlock, 65536), batch), dim3(kBlock),
                           0, st, L[l], L[l + 1], b, omega, (const T*)L[l].f, L[l].sf);
      }
      QDS_LAUNCH_CHECK();
      continue;
    }
    int rc = smooth<T, VEC>(L[l], b, nu1, omega, true, st, batch);
    if (rc) return rc;
    hipLaunchKernelGGL((restrict_residual_kernel<T, VEC>), dim3(qds_blocks(work, kBlock, 65536), batch),
                       dim3(kBlock), 0, st, L[l], L[l + 1], b);
    QDS_LAUNCH_CHECK();
  }
  if (lc < nl) {
    LevelPack<T> P;
    for (int l = lc; l < nl; ++l) P.L[l - lc] = L[l];
    hipLaunchKernelGGL((coarse_cycle_kernel<T, VEC>), dim3(1, batch), dim3(kCoarseBlock), 0, st, P, nl - lc, b,
                       nu1, nu2, ncoarse, omega);
    QDS_LAUNCH_CHECK();
  } else {
    int rc = smooth<T, VEC>(L[nl - 1], b, ncoarse, omega, true, st, batch);
    if (rc) return rc;
  }
  const int top = lc < nl ? lc - 1 : nl - 2;
  for (int l = top; l >= 0; --l) {
    const int64_t work = L[l].nx * L[l].ny * (b / VEC);
    if (fused) {
      if (use_tiled(L[l], b, VEC) && (l > 0 || io16)) {
        const unsigned tiles = (unsigned)(((L[l].nx + kUX - 1) / kUX) * ((L[l].ny + kUY - 1) / kUY));
        if (l == 0 && mixed)
          hipLaunchKernelGGL((mg_up_tiled<T, VEC, S>), dim3(tiles, batch), dim3(256), up_tiled_lds<T>(b), st, L[l],
                             L[l + 1], b, omega, f0, u0, sf0, su0);
        else
          hipLaunchKernelGGL((mg_up_tiled<T, VEC, T>), dim3(tiles, batch), dim3(256), up_tiled_lds<T>(b), st, L[l],
                             L[l + 1], b, omega, (const T*)L[l].f, L[l].u, L[l].sf, L[l].su);
      } else if (l == 0 && mixed) {
        hipLaunchKernelGGL((mg_up_kernel<T, VEC, S>), dim3(qds_blocks(work, kBlock, 65536), batch), dim3(kBlock), 0,
                           st, L[l], L[l + 1], b, omega, f0, u0, sf0, su0);
      } else {
        hipLaunchKernelGGL((mg_up_kernel<T, VEC, T>), dim3(qds_blocks(work, kBlock, 65536), batch), dim3(kBlock), 0,
                           st, L[l], L[l + 1], b, omega, (const T*)L[l].f, L[l].u, L[l].sf, L[l].su);
      }
      QDS_LAUNCH_CHECK();
      continue;
    }
    hipLaunchKernelGGL((prolong_add_kernel<T, VEC>), dim3(qds_blocks(work, kBlock, 65536), batch), dim3(kBlock),
                       0, st, L[l], L[l + 1], b);
    QDS_LAUNCH_CHECK();
    int rc = smooth<T, VEC>(L[l], b, nu2, omega, false, st, batch);
    if (rc) return rc;
  }
  return QDS_OK;
}

template <typename T, typename S = T>
int vcycle_any(int nl, Level<T>* L, int b, int nu1, int nu2, int ncoarse, T omega, hipStream_t st,
               const S* f0 = nullptr, S* u0 = nullptr, int batch = 1, int64_t sf0 = 0, int64_t su0 = 0) {
  constexpr int VW = 16 / sizeof(T);
  bool vec = b % VW == 0;
  for (int l = 0; l < nl && vec; ++l) {
    vec = vec && L[l].ldu % VW == 0 && L[l].ldf % VW == 0 &&
          reinterpret_cast<uintptr_t>(L[l].u) % 16 == 0 && reinterpret_cast<uintptr_t>(L[l].f) % 16 == 0 &&
          reinterpret_cast<uintptr_t>(L[l].t) % 16 == 0;
  }
  if (vec) return vcycle<T, VW, S>(nl, L, b, nu1, nu2, ncoarse, omega, st, f0, u0, batch, sf0, su0);
  return vcycle<T, 1, S>(nl, L, b, nu1, nu2, ncoarse, omega, st, f0, u0, batch, sf0, su0);
}

}  // namespace

extern "C" {
// Level descriptor shared with the host side (all pointers are device memory of the
// level's dtype).  u, f, t hold nx*ny*b values.
typedef struct {
  int64_t nx, ny;
  const void* d;
  const void* we;
  const void* wn;
  void* u;
  void* f;
  void* t;
  int64_t ldu, ldf;  // per-point strides of u and f (0 means b)
  int32_t neumann_x, neumann_y;  // 1: boundary rows are unknowns and coarse points (see the top of this file)
} qds_mg_level;
}

namespace {
// level l + 1 is the vertex-centred coarsening of level l in both directions
bool coarsening_ok(const qds_mg_level* lv, int n_levels) {
  for (int l = 0; l + 1 < n_levels; ++l) {
    const int64_t ex = 2 * lv[l + 1].nx + 1 - 2 * (lv[l].neumann_x ? 1 : 0);
    const int64_t ey = 2 * lv[l + 1].ny + 1 - 2 * (lv[l].neumann_y ? 1 : 0);
    if (lv[l].nx != ex || lv[l].ny != ey) return false;
  }
  return true;
}
}  // namespace

// One V-cycle for A z = r on level 0.  levels[0].f must hold r and levels[0].u
// receives z (callers may point them at their own buffers).  Requires
// levels[l].nx == 2 * levels[l+1].nx + 1 (same for ny), or 2 * levels[l+1].nx - 1
// in a direction where levels[l] sets neumann_x (neumann_y).
QDS_API int qds_mg_vcycle(int32_t dtype, int32_t n_levels, const qds_mg_level* levels, int32_t b,
                          int32_t nu_pre, int32_t nu_post, int32_t coarse_sweeps, double omega,
                          void* stream) {
  QDS_CHECK_ARG(levels && n_levels >= 1 && n_levels <= 32 && b >= 1, "qds_mg_vcycle: bad arguments");
  QDS_CHECK_ARG(coarsening_ok(levels, n_levels), "qds_mg_vcycle: levels are not vertex-centred coarsenings");
  for (int l = 0; l < n_levels; ++l)
    QDS_CHECK_ARG(levels[l].d && levels[l].we && levels[l].wn && levels[l].u && levels[l].f && levels[l].t,
                  "qds_mg_vcycle: null pointer on level %d", l);
  for (int l = 0; l < n_levels; ++l)
    QDS_CHECK_ARG(levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldu ? levels[l].ldu : b) < (int64_t(1) << 31) &&
                  levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldf ? levels[l].ldf : b) < (int64_t(1) << 31),
                  "qds_mg_vcycle: level %d exceeds 32-bit indexing", l);
  hipStream_t st = qds_stream(stream);
  if (dtype == qds::F32) {
    Level<float> L[32];
    for (int l = 0; l < n_levels; ++l)
      L[l] = Level<float>{levels[l].nx, levels[l].ny, (const float*)levels[l].d, (const float*)levels[l].we,
                          (const float*)levels[l].wn, (float*)levels[l].u, (float*)levels[l].f,
                          (float*)levels[l].t, levels[l].ldu ? levels[l].ldu : b,
                          levels[l].ldf ? levels[l].ldf : b, levels[l].neumann_x ? 1 : 0,
                          levels[l].neumann_y ? 1 : 0};
    return vcycle_any<float>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, (float)omega, st);
  }
  if (dtype == qds::F64) {
    Level<double> L[32];
    for (int l = 0; l < n_levels; ++l)
      L[l] = Level<double>{levels[l].nx, levels[l].ny, (const double*)levels[l].d,
                           (const double*)levels[l].we, (const double*)levels[l].wn,
                           (double*)levels[l].u, (double*)levels[l].f, (double*)levels[l].t,
                           levels[l].ldu ? levels[l].ldu : b, levels[l].ldf ? levels[l].ldf : b,
                           levels[l].neumann_x ? 1 : 0, levels[l].neumann_y ? 1 : 0};
    return vcycle_any<double>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, omega, st);
  }
  QDS_CHECK_ARG(false, "qds_mg_vcycle: dtype must be F32 or F64");
  return QDS_EINVAL;
}

// V-cycle with an FP32 hierarchy (dtype F32) on an FP64 right-hand side and result
// (io_dtype F64): levels[0].f / .u point at FP64 blocks (strides ldf / ldu in
// elements), every other level is FP32.  Equal dtypes behave as qds_mg_vcycle.
QDS_API int qds_mg_vcycle_io(int32_t dtype, int32_t io_dtype, int32_t n_levels, const qds_mg_level* levels,
                             int32_t b, int32_t nu_pre, int32_t nu_post, int32_t coarse_sweeps, double omega,
                             void* stream) {
  if (dtype == io_dtype)
    return qds_mg_vcycle(dtype, n_levels, levels, b, nu_pre, nu_post, coarse_sweeps, omega, stream);
  QDS_CHECK_ARG(dtype == qds::F32 && io_dtype == qds::F64, "qds_mg_vcycle_io: only an FP32 hierarchy with FP64 level-0 blocks");
  QDS_CHECK_ARG(levels && n_levels >= 2 && n_levels <= 32 && b >= 1, "qds_mg_vcycle_io: bad arguments");
  QDS_CHECK_ARG(coarsening_ok(levels, n_levels), "qds_mg_vcycle_io: levels are not vertex-centred coarsenings");
  for (int l = 0; l < n_levels; ++l)
    QDS_CHECK_ARG(levels[l].d && levels[l].we && levels[l].wn && levels[l].u && levels[l].f && levels[l].t,
                  "qds_mg_vcycle_io: null pointer on level %d", l);
  for (int l = 0; l < n_levels; ++l)
    QDS_CHECK_ARG(levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldu ? levels[l].ldu : b) < (int64_t(1) << 31) &&
                  levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldf ? levels[l].ldf : b) < (int64_t(1) << 31),
                  "qds_mg_vcycle_io: level %d exceeds 32-bit indexing", l);
  Level<float> L[32];
  for (int l = 0; l < n_levels; ++l)
    L[l] = Level<float>{levels[l].nx, levels[l].ny, (const float*)levels[l].d, (const float*)levels[l].we,
                        (const float*)levels[l].wn, (float*)levels[l].u, (float*)levels[l].f, (float*)levels[l].t,
                        levels[l].ldu ? levels[l].ldu : b, levels[l].ldf ? levels[l].ldf : b,
                        levels[l].neumann_x ? 1 : 0, levels[l].neumann_y ? 1 : 0};
  // level 0's FP32 pointers are never dereferenced in the mixed cycle; its f / u are
  // the FP64 blocks; its t is the hierarchy's own scratch
  const double* f0 = (const double*)levels[0].f;
  double* u0 = (double*)levels[0].u;
  L[0].f = L[0].t;
  L[0].u = L[0].t;
  return vcycle_any<float, double>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, (float)omega,
                                   qds_stream(stream), f0, u0);
}

extern "C" {
// Element strides between the problems of a batch, per level (0: shared).
typedef struct {
  int64_t sd, sw, su, sf, st;   // d; we and wn; u; f; t
} qds_mg_level_strides;
}

// One V-cycle for each of `batch` problems on the same lattice hierarchy in one
// launch per pass: problem i's level arrays are the level's pointers plus i times
// the level's strides.  dtype / io_dtype as qds_mg_vcycle_io (for an FP64 level 0
// on an FP32 hierarchy, strides[0].sf / .su are the FP64 blocks' strides).
QDS_API int qds_mg_vcycle_batched(int32_t dtype, int32_t io_dtype, int32_t n_levels, const qds_mg_level* levels,
                                  const qds_mg_level_strides* strides, int32_t batch, int32_t b, int32_t nu_pre,
                                  int32_t nu_post, int32_t coarse_sweeps, double omega, void* stream) {
  QDS_CHECK_ARG(levels && strides && n_levels >= 1 && n_levels <= 32 && b >= 1 && batch >= 1 && batch <= 65535,
                "qds_mg_vcycle_batched: bad arguments");
  QDS_CHECK_ARG(coarsening_ok(levels, n_levels), "qds_mg_vcycle_batched: levels are not vertex-centred coarsenings");
  for (int l = 0; l < n_levels; ++l) {
    QDS_CHECK_ARG(levels[l].d && levels[l].we && levels[l].wn && levels[l].u && levels[l].f && levels[l].t,
                  "qds_mg_vcycle_batched: null pointer on level %d", l);
    QDS_CHECK_ARG(levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldu ? levels[l].ldu : b) < (int64_t(1) << 31) &&
                  levels[l].nx * levels[l].ny * (int64_t)(levels[l].ldf ? levels[l].ldf : b) < (int64_t(1) << 31),
                  "qds_mg_vcycle_batched: level %d exceeds 32-bit indexing", l);
    const qds_mg_level_strides& z = strides[l];
    QDS_CHECK_ARG(z.sd >= 0 && z.sw >= 0 && z.su >= 0 && z.sf >= 0 && z.st >= 0,
                  "qds_mg_vcycle_batched: negative stride on level %d", l);
  }
  hipStream_t stm = qds_stream(stream);
  auto fill = [&](auto* L, auto tag) {
    using T = decltype(tag);
    for (int l = 0; l < n_levels; ++l) {
      L[l] = Level<T>{levels[l].nx, levels[l].ny, (const T*)levels[l].d, (const T*)levels[l].we,
                      (const T*)levels[l].wn, (T*)levels[l].u, (T*)levels[l].f, (T*)levels[l].t,
                      levels[l].ldu ? levels[l].ldu : b, levels[l].ldf ? levels[l].ldf : b,
                      levels[l].neumann_x ? 1 : 0, levels[l].neumann_y ? 1 : 0,
                      strides[l].sd, strides[l].sw, strides[l].su, strides[l].sf, strides[l].st};
    }
  };
  if (dtype == io_dtype && dtype == qds::F32) {
    Level<float> L[32];
    fill(L, 0.0f);
    return vcycle_any<float, float>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, (float)omega, stm, nullptr,
                                    nullptr, batch);
  }
  if (dtype == io_dtype && dtype == qds::F64) {
    Level<double> L[32];
    fill(L, 0.0);
    return vcycle_any<double, double>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, omega, stm, nullptr, nullptr,
                                      batch);
  }
  QDS_CHECK_ARG(dtype == qds::F32 && io_dtype == qds::F64 && n_levels >= 2,
                "qds_mg_vcycle_batched: only an FP32 hierarchy with FP64 level-0 blocks");
  Level<float> L[32];
  fill(L, 0.0f);
  const double* f0 = (const double*)levels[0].f;
  double* u0 = (double*)levels[0].u;
  L[0].f = L[0].t;
  L[0].u = L[0].t;
  L[0].sf = L[0].st;
  L[0].su = L[0].st;
  return vcycle_any<float, double>(n_levels, L, b, nu_pre, nu_post, coarse_sweeps, (float)omega, stm, f0, u0, batch,
                                   strides[0].sf, strides[0].su);
}

// out = A u on one level (tests of the lattice operator).
QDS_API int qds_mg_apply(int32_t dtype, const qds_mg_level* level, int32_t b, const void* u, void* out,
                         void* stream) {
  QDS_CHECK_ARG(level && u && out && level->d && level->we && level->wn, "qds_mg_apply: null");
  hipStream_t st = qds_stream(stream);
  const int64_t work = level->nx * level->ny * b;
  const dim3 g(qds_blocks(work, kBlock, 131072)), blk(kBlock);
  if (dtype == qds::F64) {
    Level<double> L{level->nx, level->ny, (const double*)level->d, (const double*)level->we,
                    (const double*)level->wn, nullptr, nullptr, nullptr, b, b};
    hipLaunchKernelGGL(apply_kernel<double>, g, blk, 0, st, L, b, (const double*)u, (double*)out);
  } else if (dtype == qds::F32) {
    Level<float> L{level->nx, level->ny, (const float*)level->d, (const float*)level->we,
                   (const float*)level->wn, nullptr, nullptr, nullptr, b, b};
    hipLaunchKernelGGL(apply_kernel<float>, g, blk, 0, st, L, b, (const float*)u, (float*)out);
  } else {
    QDS_CHECK_ARG(false, "qds_mg_apply: bad dtype");
  }
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
