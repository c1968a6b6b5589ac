// Rayleigh-Ritz of the block eigensolver on the GPU (FP64), one workgroup per
// problem of a batch.
//
// Reference: the dense steps inside scipy's lobpcg / ARPACK
// (qdsim_cython/solvers/eigen_solver.pyx:225-266): the projected pencil
// (S^T H S, S^T M S) of the current basis S = [X | W | P] is solved for its lowest
// b eigenpairs.  The projected matrices are m x m with m <= 96, so a whole problem
// fits one workgroup (registers and LDS), and the solve never leaves the device: the
// block iteration runs without a host round trip and can be captured in a HIP graph.
//
// Per problem, with G = [GH | GM] (m x 2m, row-major) from the Gram kernel:
//  1. residual test of the current block from its column norms (when given):
//     res_j = ||r_j|| / (max(|lam_j|, s) ||M x_j||), s = max_{j<k} |lam_j|; the
//     locking mask (res_j >= tol) scales the next residual; converged when
//     max_{j<k} res_j < tol, and with `freeze` a converged problem keeps its block
//     (C = [I; 0]) from then on;
//  2. scale by the M-Gram diagonal; columns whose diagonal vanished (locked
//     residuals, a zero P block) are decoupled and never selected;
//  3. Cholesky of the scaled M-Gram fused with the first triangular solve of
//     T = L^-1 GH L^-T, then the second solve (the m x m matrices live in registers,
//     thread (wave w, lane l) holding entries (w + 8 r, l + 64 c); one barrier per
//     column); a pivot below drop_tol sends the problem to SVQB instead
//     (eigen-decomposition of the M-Gram, directions with eigenvalue below drop_tol
//     times the largest dropped, never fewer than b kept);
//  4. the lowest b eigenpairs of T: Householder tridiagonalisation (registers, two
//     barriers per column), eigenvalues by multisection on Sturm counts, eigenvectors
//     by inverse iteration on the tridiagonal, Cholesky-QR, back-transformation; if a
//     check fails there (or m < 10, or 2 b > m + 1), cyclic Jacobi on T instead;
//  5. C = diag(d) L^-T Z; CP = C with the X rows zeroed (all zero when directions
//     were dropped: the P block restarts).
// G's M half is overwritten (L, or the SVQB basis, is kept there).
//
// Latencies that shaped the code (one MI355X workgroup of 512 threads): a barrier
// with an LDS round trip ~75 ns, a dependent LDS load ~40 ns, an FP64 division
// ~44 ns, a wave sum by ds_bpermute shuffles ~240 ns against ~65 ns by DPP.
#include "qds_common.h"

#include <cmath>
#include <mutex>

namespace {

constexpr int kRRThreads = 512;
constexpr int kRRMaxM = 96;
constexpr int kRRMaxSweeps = 40;
constexpr int kStatusHead = 8;

struct RRArgs {
  int m, b, k;
  int have_norms, freeze;
  double tol, drop_tol;
  int64_t g_stride, c_stride;
};


constexpr int kWaves = kRRThreads / 64;                     // 8
constexpr int kRowSlots = (kRRMaxM + kWaves - 1) / kWaves;  // rows of a register tile per thread
constexpr int kColSlots = (kRRMaxM + 63) / 64;              // columns of a register tile per thread
constexpr int kGroupRows = (kRRMaxM + 15) / 16;             // rows per lane of a 16-lane column group

// ---- cross-lane sums and maxima (DPP within 16-lane rows, readlane across them) --------
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

struct SumOp {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
  __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
struct MinOp {
  __device__ double operator()(double a, double b) const { return fmin(a, b); }
};

// reduction over each aligned group of 16 lanes (quad swaps, half-row and row mirrors),
// the same bits in every lane of the group; all 64 lanes must be active
template <class Op>
__device__ __forceinline__ double row16_reduce(double v, Op op) {
  v = op(v, dpp_mov<0xB1>(v));    // quad_perm [1, 0, 3, 2]
  v = op(v, dpp_mov<0x4E>(v));    // quad_perm [2, 3, 0, 1]
  v = op(v, dpp_mov<0x141>(v));   // row_half_mirror
  v = op(v, dpp_mov<0x140>(v));   // row_mirror
  return v;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// reduction over the wave, the same value in every lane; all 64 lanes must be active
template <class Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
  v = row16_reduce(v, op);
  return op(op(readlane_d(v, 0), readlane_d(v, 16)), op(readlane_d(v, 32), readlane_d(v, 48)));
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_reduce(v, SumOp());
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < kWaves; ++i) s += red[i];
  return s;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_reduce(v, MaxOp());
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < kWaves; ++i) s = fmax(s, red[i]);
  return s;
}

// 1/x from the hardware estimate (relative error ~4e-8) and Newton steps, each of
// which squares the error (one step: ~1.5e-15)
__device__ __forceinline__ double rcp_newton1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rcp_newton2(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}


// Players of round r of the circle method (position 0 fixed, the others rotate):
// pair i of round r is (player(i), player(m-1-i)); every unordered pair of the m
// indices meets exactly once in rounds 0 .. m-2.
__device__ __forceinline__ int player(int m, int r, int j) { return j == 0 ? 0 : 1 + (j - 1 + r) % (m - 1); }

// Rotation J = [[c, s], [-s, c]] on (p, q) with (J^T A J)_pq = 0.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double& c, double& s) {
  const double tau = (aqq - app) / (2.0 * apq);
  double t;
  if (fabs(tau) > 1e150) t = 0.5 / tau;
  else t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  c = 1.0 / sqrt(1.0 + t * t);
  s = t * c;
}

// Cyclic Jacobi on the symmetric A (m x m, stride ld), V <- V J (V initialised by
// the caller).  Returns the number of sweeps.
__device__ int jacobi(double* A, double* V, int m, int ld, double* cs, int* pq, int* flag, double* red) {
  const int np = m / 2;
  const int nblk = np * (np + 1) / 2;
  int sweeps = 0;
  for (int sw = 0; sw < kRRMaxSweeps; ++sw) {
    double off = 0.0, dia = 0.0;
    for (int u = threadIdx.x; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      const double a = A[i * ld + j];
      if (i == j) dia += a * a;
      else off += a * a;
    }
    off = block_sum(off, red);
    dia = block_sum(dia, red);
    if (off == 0.0 || off <= 1e-32 * dia) break;
    const double abs_floor = 1e-34 * sqrt(off + dia);
    if (threadIdx.x == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      for (int t = threadIdx.x; t < np; t += kRRThreads) {
        const int a = player(m, r, t), b = player(m, r, m - 1 - t);
        const int p = a < b ? a : b, q = a < b ? b : a;
        pq[2 * t] = p;
        pq[2 * t + 1] = q;
        const double app = A[p * ld + p], aqq = A[q * ld + q], apq = A[p * ld + q];
        double c = 1.0, s = 0.0;
        if (fabs(apq) > abs_floor && fabs(apq) > 1e-16 * sqrt(fabs(app * aqq))) {
          jacobi_rotation(app, aqq, apq, c, s);
          *flag = 1;
        }
        cs[2 * t] = c;
        cs[2 * t + 1] = s;
      }
      __syncthreads();
      // unordered pairs of rotation pairs (P <= Q): rows P x columns Q and their
      // transpose belong to one thread
      for (int u = threadIdx.x; u < nblk; u += kRRThreads) {
        int P = 0, rem = u;
        while (rem >= np - P) { rem -= np - P; ++P; }
        const int Q = P + rem;
        const double s1 = cs[2 * P + 1], s2 = cs[2 * Q + 1];
        if (s1 == 0.0 && s2 == 0.0) continue;
        const double c1 = cs[2 * P], c2 = cs[2 * Q];
        const int p1 = pq[2 * P], q1 = pq[2 * P + 1], p2 = pq[2 * Q], q2 = pq[2 * Q + 1];
        const double a00 = A[p1 * ld + p2], a01 = A[p1 * ld + q2], a10 = A[q1 * ld + p2], a11 = A[q1 * ld + q2];
        // J_P^T a: rows
        const double r00 = c1 * a00 - s1 * a10, r01 = c1 * a01 - s1 * a11;
        const double r10 = s1 * a00 + c1 * a10, r11 = s1 * a01 + c1 * a11;
        // (.) J_Q: columns
        double n00 = c2 * r00 - s2 * r01, n01 = s2 * r00 + c2 * r01;
        double n10 = c2 * r10 - s2 * r11, n11 = s2 * r10 + c2 * r11;
        if (P == Q) {
          n01 = 0.0;
          n10 = 0.0;
        }
        A[p1 * ld + p2] = n00;
        A[p1 * ld + q2] = n01;
        A[q1 * ld + p2] = n10;
        A[q1 * ld + q2] = n11;
        if (P != Q) {
          A[p2 * ld + p1] = n00;
          A[q2 * ld + p1] = n01;
          A[p2 * ld + q1] = n10;
          A[q2 * ld + q1] = n11;
        }
      }
      for (int u = threadIdx.x; u < m * np; u += kRRThreads) {
        const int i = u / np, Q = u - i * np;
        const double s2 = cs[2 * Q + 1];
        if (s2 == 0.0) continue;
        const double c2 = cs[2 * Q];
        const int p2 = pq[2 * Q], q2 = pq[2 * Q + 1];
        const double vp = V[i * ld + p2], vq = V[i * ld + q2];
        V[i * ld + p2] = c2 * vp - s2 * vq;
        V[i * ld + q2] = s2 * vp + c2 * vq;
      }
      __syncthreads();
    }
    ++sweeps;
    const int rotated = *flag;
    __syncthreads();
    if (!rotated) break;
  }
  return sweeps;
}

__device__ void set_identity(double* V, int m, int ld) {
  for (int u = threadIdx.x; u < m * m; u += kRRThreads) {
    const int i = u / m, j = u - i * m;
    V[i * ld + j] = i == j ? 1.0 : 0.0;
  }
}

// sel[r] = index of the r-th smallest of val[0..m) among valid ones (ties by index).
__device__ void select_lowest(const double* val, const int* valid, int m, int b, int* sel) {
  for (int i = threadIdx.x; i < m; i += kRRThreads) {
    if (!valid[i]) continue;
    int rank = 0;
    for (int j = 0; j < m; ++j)
      if (valid[j] && (val[j] < val[i] || (val[j] == val[i] && j < i))) ++rank;
    if (rank < b) sel[rank] = i;
  }
}


// B <- L^-1 B for the lower-triangular L (diagonal included), stride ld: one
// thread per column, rows in order (L's entries are read by every thread at once)
__device__ void lower_solve_columns(const double* L, double* B, int m, int ld) {
  for (int j = threadIdx.x; j < m; j += kRRThreads) {
    for (int i = 0; i < m; ++i) {
      double s = B[i * ld + j];
      for (int k = 0; k < i; ++k) s -= L[i * ld + k] * B[k * ld + j];
      B[i * ld + j] = s / L[i * ld + i];
    }
  }
}

// V <- L^-1 V L^-T for symmetric V (two column-parallel solves around a transpose),
// symmetrised
__device__ void congruence_lower(const double* L, double* V, int m, int ld) {
  lower_solve_columns(L, V, m, ld);
  __syncthreads();
  for (int u = threadIdx.x; u < m * m; u += kRRThreads) {
    const int i = u / m, j = u - i * m;
    if (i < j) {
      const double x = V[i * ld + j];
      V[i * ld + j] = V[j * ld + i];
      V[j * ld + i] = x;
    }
  }
  __syncthreads();
  lower_solve_columns(L, V, m, ld);
  __syncthreads();
  for (int u = threadIdx.x; u < m * m; u += kRRThreads) {
    const int i = u / m, j = u - i * m;
    if (i < j) {
      const double sv = 0.5 * (V[i * ld + j] + V[j * ld + i]);
      V[i * ld + j] = sv;
      V[j * ld + i] = sv;
    }
  }
  __syncthreads();
}


// ---- register tiles ----------------------------------------------------------------
// Thread (wave w, lane l) of the workgroup holds entries (w + 8 r, l + 64 c) of an
// m x m matrix, r < R, c < C (zero outside the matrix; R and C are the smallest that
// cover m, so that no step spends instructions on empty rows).  Every index into a
// tile is a compile-time constant after unrolling, so tiles stay in registers.
template <int R, int C>
using TileT = double[R][C];

template <int C>
__device__ __forceinline__ double col_slot(const double (&x)[C], int j) {
  if constexpr (C == 1) return x[0];
  else return j < 64 ? x[0] : x[1];
}

__device__ __forceinline__ double readfirstlane_d(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// Column step k of the fused Cholesky / first solve (KS = k / 64, QS = (k + 1) / 64:
// the column slots of k and k + 1, fixed at compile time so that no register is
// selected at run time).  The multiplier of each row comes from the lane holding
// column k by readlane (a scalar), zero for rows not below k; every update is an
// unconditional FMA.  Row k + 1 of U and column k + 1 of A are final afterwards and
// are published.
template <int KS, int QS, int R, int C>
__device__ __forceinline__ void chol_step(int k, int m, int w, int lane, double rp, const double* cb, const double* rb,
                                          double* cbn, double* rbn, TileT<R, C>& Ar, TileT<R, C>& Tr) {
  double akj[C], ukj[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int j = lane + 64 * c;
    akj[c] = (j > k && j < m) ? cb[j] * rp : 0.0;
    ukj[c] = j < m ? rb[j] * rp : 0.0;
  }
  const int q = k + 1;
#pragma unroll
  for (int r = 0; r < R; ++r) {   // rows beyond m hold zeros and stay zero
    const int i = w + kWaves * r;
    const double aik = readlane_d(Ar[r][KS], k & 63);
    const double li = i > k ? aik : 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      Ar[r][c] = fma(-li, akj[c], Ar[r][c]);
      Tr[r][c] = fma(-li, ukj[c], Tr[r][c]);
    }
    if (i == q && q < m) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int j = lane + 64 * c;
        if (j < m) rbn[j] = Tr[r][c];
      }
    }
  }
  if (lane == (q & 63)) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = w + kWaves * r;
      if (i >= q && i < m) cbn[i] = Ar[r][QS];
    }
  }
}

// Steps 3 and 4 on the Cholesky path.
//  * Cholesky of the scaled M-Gram in right-looking outer-product form (column k keeps
//    L_ik L_kk) fused with U = L1^-1 B, B the scaled H-Gram and L1 the unit lower
//    factor (l_ik = A_ik / A_kk): one barrier per column; column k + 1 of A and row
//    k + 1 of U are published to LDS (double-buffered, in V) once final;
//  * T = D^-1/2 L1^-1 (U^T D^-1/2), D = diag(pivots): the same substitution on the
//    scaled transpose (multipliers from LDS), then symmetrised.
// On success T is in Tr, L = A diag(D^-1/2) in G's M half (zero above the diagonal),
// D^1/2 in sq, and 0 is returned; 1 (G untouched) when a pivot is not above drop_tol.
// rdiag (m) is scratch.
template <int R, int C>
__device__ __forceinline__ int chol_congruence(double* G, int m, int ldg, const double* d, const int* valid,
                                               double drop_tol, double* A, double* V, int ld, double* sq,
                                               double* rdiag, TileT<R, C>& Tr) {
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  TileT<R, C> Ar;
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      double av = 0.0, hv = 0.0;
      if (i < m && j < m) {
        if (valid[i] && valid[j]) {
          const double dd = d[i] * d[j];
          av = 0.5 * (G[i * ldg + m + j] + G[j * ldg + m + i]) * dd;
          hv = 0.5 * (G[i * ldg + j] + G[j * ldg + i]) * dd;
        } else if (i == j) {
          av = 1.0;
        }
      }
      Ar[r][c] = av;
      Tr[r][c] = hv;
    }
  }
  double* colbuf = V;          // [2][m]: column k of A
  double* rowbuf = V + 2 * m;  // [2][m]: row k of U
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = w + kWaves * r;
      if (i < m) colbuf[i] = Ar[r][0];
    }
  }
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = lane + 64 * c;
      if (j < m) rowbuf[j] = Tr[0][c];
    }
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const int par = k & 1, np = par ^ 1;
    const double piv = colbuf[par * m + k];
    if (!(piv > drop_tol)) return 1;   // uniform: every thread reads the same value
    const double rp = readfirstlane_d(1.0 / piv);
    const double *cb = colbuf + par * m, *rb = rowbuf + par * m;
    double *cbn = colbuf + np * m, *rbn = rowbuf + np * m;
    if constexpr (C == 1) {
      chol_step<0, 0>(k, m, w, lane, rp, cb, rb, cbn, rbn, Ar, Tr);
    } else {
      if (k < 63) chol_step<0, 0>(k, m, w, lane, rp, cb, rb, cbn, rbn, Ar, Tr);
      else if (k == 63) chol_step<0, 1>(k, m, w, lane, rp, cb, rb, cbn, rbn, Ar, Tr);
      else chol_step<1, 1>(k, m, w, lane, rp, cb, rb, cbn, rbn, Ar, Tr);
    }
    __syncthreads();
  }
  // pivots' roots and reciprocals; A (strictly lower, zero elsewhere) and U to LDS
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      if (i < m && j < m) {
        A[i * ld + j] = j < i ? Ar[r][c] : 0.0;
        if (j == i) {
          sq[i] = sqrt(Ar[r][c]);
          rdiag[i] = 1.0 / Ar[r][c];
        }
        V[i * ld + j] = Tr[r][c];
      }
    }
  }
  __syncthreads();
  for (int i = w; i < m; i += kWaves)
    for (int j = lane; j < m; j += 64) G[i * ldg + m + j] = j < i ? A[i * ld + j] / sq[j] : (j == i ? sq[i] : 0.0);
  // C = U^T D^-1/2, then C <- L1^-1 C
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      Tr[r][c] = (i < m && j < m) ? V[j * ld + i] / sq[j] : 0.0;
    }
  }
  __syncthreads();   // V becomes the row buffer
  rowbuf = V;
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = lane + 64 * c;
      if (j < m) rowbuf[j] = Tr[0][c];
    }
  }
  __syncthreads();
  for (int k = 0; k + 1 < m; ++k) {
    const int par = k & 1, np = par ^ 1;
    const double rp = rdiag[k];
    double ckj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = lane + 64 * c;
      ckj[c] = j < m ? rowbuf[par * m + j] * rp : 0.0;
    }
    const int q = k + 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = w + kWaves * r;
      const double li = i < m ? A[i * ld + k] : 0.0;   // zero unless k < i < m
#pragma unroll
      for (int c = 0; c < C; ++c) Tr[r][c] = fma(-li, ckj[c], Tr[r][c]);
      if (i == q) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int j = lane + 64 * c;
          if (j < m) rowbuf[np * m + j] = Tr[r][c];
        }
      }
    }
    __syncthreads();
  }
  // T = D^-1/2 C, symmetrised (0.5 (a + b) is the same bits from either side)
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      if (i < m && j < m) V[i * ld + j] = Tr[r][c] / sq[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      Tr[r][c] = (i < m && j < m) ? 0.5 * (V[i * ld + j] + V[j * ld + i]) : 0.0;
    }
  }
  __syncthreads();   // V is free again
  return 0;
}

struct TriWork {
  double* td;    // diagonal (n)
  double* te;    // off-diagonal (n; te[n-1] = 0)
  double* te2;   // te^2 (n)
  double* tb;    // reflector factors beta (n)
  double* pp;    // scratch (n)
  double* bnd;   // [0] lower bound, [1] upper bound, [2] pivmin, [3] |T|
  int* clst;     // cluster starts (nsel + 1)
  int* flag;     // nonzero: a check failed
};

// Publication of column q of the tile T for the Householder step that reflects it:
// x[i] = T[i][q] below the diagonal, zero on and above it; td[q] = T[q][q].
template <int QS, int R, int C>
__device__ __forceinline__ void publish_column(const TileT<R, C>& Tr, int q, int n, int w, int lane, double* x, double* td) {
  if (lane == (q & 63)) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = w + kWaves * r;
      if (i < n) {
        x[i] = i > q ? Tr[r][QS] : 0.0;
        if (i == q) td[q] = Tr[r][QS];
      }
    }
  }
}

// Householder reduction of the symmetric T (register tile, n x n) to the tridiagonal
// (td, te).  Step k reflects rows and columns k+1.. by H_k = I - beta_k v v^T,
// v = x - alpha e_1 for x = T[k+1..n)[k], alpha = -sign(x_0) |x|; v goes to column k of
// V (rows k+1..n), beta_k to tb[k] (0: no reflection).  Per step: every wave forms v
// from the published column (DPP sums); each wave's partial products T v over its rows
// go to LDS (barrier); every wave forms p = beta T v and w = p - (beta p.v / 2) v for
// all columns and updates its tile, T <- T - v w^T - w v^T (unconditional FMAs: v and
// w vanish on the rows and columns not below k); column k + 1 is published (barrier).
template <int R, int C>
__device__ __forceinline__ void tridiagonalize(TileT<R, C>& Tr, int n, double* V, int ld, double* scratch, TriWork tw) {
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* xk = scratch;            // [2][n]: column k below the diagonal
  double* part = scratch + 2 * n;  // [kWaves][n]: partial products
  publish_column<0>(Tr, 0, n, w, lane, xk, tw.td);
  __syncthreads();
  for (int k = 0; k + 2 < n; ++k) {
    const double* x = xk + (k & 1) * n;
    double* xn = xk + ((k + 1) & 1) * n;
    const double x0 = x[k + 1];
    double s = 0.0;
    for (int i = k + 2 + lane; i < n; i += 64) s += x[i] * x[i];
    const double sqs = wave_reduce(s, SumOp());
    double beta = 0.0, alpha = x0, v0 = x0;
    if (sqs != 0.0) {   // uniform
      const double nrm = sqrt(x0 * x0 + sqs);
      alpha = x0 > 0.0 ? -nrm : nrm;
      v0 = x0 - alpha;
      beta = 2.0 / (v0 * v0 + sqs);
    }
    if (tid == 0) {
      tw.te[k] = alpha;
      tw.tb[k] = beta;
    }
    if (beta != 0.0) {   // uniform
      const double dv = v0 - x0;
      if (w == 0)
        for (int i = k + 1 + lane; i < n; i += 64) V[i * ld + k] = i == k + 1 ? v0 : x[i];
      double vi[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = w + kWaves * r;
        vi[r] = i < n ? x[i] : 0.0;
        if (i == k + 1) vi[r] += dv;
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) acc = fma(Tr[r][c], vi[r], acc);
        const int j = lane + 64 * c;
        if (j < n) part[w * n + j] = acc;
      }
      __syncthreads();
      double pj[C], vj[C], wj[C];
      double pv = 0.0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int j = lane + 64 * c;
        pj[c] = 0.0;
        vj[c] = 0.0;
        if (j > k && j < n) {
          double acc = 0.0;
#pragma unroll
          for (int q = 0; q < kWaves; ++q) acc += part[q * n + j];
          pj[c] = beta * acc;
          vj[c] = j == k + 1 ? v0 : x[j];
        }
        pv = fma(pj[c], vj[c], pv);
      }
      const double K = 0.5 * beta * wave_reduce(pv, SumOp());
#pragma unroll
      for (int c = 0; c < C; ++c) wj[c] = pj[c] - K * vj[c];
#pragma unroll
      for (int r = 0; r < R; ++r) {   // w_i (zero unless k < i < n) from the lane of column i
        const int i = w + kWaves * r;
        const double wi = readlane_d(wj[(C == 2 && r >= 64 / kWaves) ? 1 : 0], i & 63);
#pragma unroll
        for (int c = 0; c < C; ++c) Tr[r][c] = fma(-vi[r], wj[c], fma(-wi, vj[c], Tr[r][c]));
      }
    }
    if constexpr (C == 1) {
      publish_column<0>(Tr, k + 1, n, w, lane, xn, tw.td);
    } else {
      if (k + 1 < 64) publish_column<0>(Tr, k + 1, n, w, lane, xn, tw.td);
      else publish_column<1>(Tr, k + 1, n, w, lane, xn, tw.td);
    }
    __syncthreads();
  }
  // the last off-diagonal and diagonal entry
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int i = w + kWaves * r, j = lane + 64 * c;
      if (i == n - 1 && j == n - 2) tw.te[n - 2] = Tr[r][c];
      if (i == n - 1 && j == n - 1) tw.td[n - 1] = Tr[r][c];
    }
  }
  if (tid == 0) {
    tw.te[n - 1] = 0.0;
    tw.tb[n - 2] = 0.0;
    tw.tb[n - 1] = 0.0;
  }
  __syncthreads();
}

// Number of eigenvalues of the tridiagonal (td, te) below x1 and below x2 (two
// interleaved Sturm sequences, pivot guard of the bisection in LAPACK's dlaebz,
// reciprocals from the hardware estimate and one Newton step).
__device__ __forceinline__ void sturm_count2(const double* td, const double* te2, int n, double x1, double x2,
                                             double pivmin, int& c1, int& c2) {
  double q1 = td[0] - x1, q2 = td[0] - x2;
  int a1 = q1 < 0.0, a2 = q2 < 0.0;
  for (int i = 1; i < n; ++i) {
    const double di = td[i], e2 = te2[i - 1];
    if (fabs(q1) < pivmin) q1 = -pivmin;
    if (fabs(q2) < pivmin) q2 = -pivmin;
    q1 = fma(-e2, rcp_newton1(q1), di - x1);
    q2 = fma(-e2, rcp_newton1(q2), di - x2);
    a1 += q1 < 0.0;
    a2 += q2 < 0.0;
  }
  c1 = a1;
  c2 = a2;
}

// 16 bits to the even positions of 32
__device__ __forceinline__ unsigned spread16(unsigned x) {
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}

__device__ __forceinline__ double start_value(int i, int j) {
  unsigned h = (unsigned)(i + 1) * 2654435761u ^ (unsigned)(j + 7) * 2246822519u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  return (double)(h & 0xFFFFFu) / 1048576.0 - 0.5;
}

// Lowest nsel eigenpairs of the symmetric T (register tile, n x n):
//  1. tridiagonalisation (above);
//  2. eigenvalues by multisection: the 16 lanes of group j evaluate Sturm counts at 32
//     interior points of eigenvalue j's interval (5 bits per step), to ~eps |T|;
//  3. eigenvectors of the tridiagonal by inverse iteration, one thread per cluster of
//     eigenvalues closer than 1e3 eps |T| (Gram-Schmidt inside a cluster): two solves
//     from a fixed pseudo-random start, each an LU with partial pivoting whose forward
//     substitution runs inside the factorisation, the recurrences carried in registers
//     (U's three diagonals and the right-hand side in four scratch columns of A);
//  4. Cholesky-QR of the nsel vectors (restores orthogonality across clusters to
//     rounding), then back-transformation by the reflectors: 16 lanes per column, DPP
//     sums, no barriers.
// Eigenvalues ascending in theta[0..nsel), vectors in columns 0..nsel of A (stride ld).
// Returns nonzero (uniformly) when a check fails: non-finite values, a Gram matrix
// that is not positive definite, too little scratch.
template <int R, int C>
__device__ __forceinline__ int tridiag_lowest(TileT<R, C>& Tr, double* V, double* A, int n, int ld, int nsel, double* theta, TriWork tw) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double eps = 2.220446049250313e-16;
  if (2 * nsel > ld || n < 10) return 1;   // Cholesky-QR needs nsel free columns; scratch 10 n <= n ld
  if (tid == 0) *tw.flag = 0;
  tridiagonalize(Tr, n, V, ld, A, tw);
  // ---- bounds (Gershgorin), pivot guard ----------------------------------------------
  if (wave == 0) {   // uniform: all 64 lanes of wave 0
    double lo = 1e308, hi = -1e308, e2max = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double em = i > 0 ? fabs(tw.te[i - 1]) : 0.0, ep = fabs(tw.te[i]);
      lo = fmin(lo, tw.td[i] - em - ep);
      hi = fmax(hi, tw.td[i] + em + ep);
      tw.te2[i] = tw.te[i] * tw.te[i];
      e2max = fmax(e2max, tw.te2[i]);
    }
    lo = wave_reduce(lo, MinOp());
    hi = wave_reduce(hi, MaxOp());
    e2max = wave_reduce(e2max, MaxOp());
    if (lane == 0) {
      const double tn = fmax(fabs(lo), fabs(hi));
      const double pivmin = 2.2250738585072014e-308 * fmax(1.0, e2max);
      tw.bnd[0] = lo - 2.0 * eps * tn - 2.0 * pivmin;
      tw.bnd[1] = hi + 2.0 * eps * tn + 2.0 * pivmin;
      tw.bnd[2] = pivmin;
      tw.bnd[3] = tn;
      if (!isfinite(lo) || !isfinite(hi)) *tw.flag = 1;
    }
  }
  __syncthreads();
  if (*tw.flag) return 1;
  const double pivmin = tw.bnd[2], tnorm = tw.bnd[3];
  // ---- 2. multisection ------------------------------------------------------------------
  {
    const int sub = tid & 15, shift = lane & 48;
    const double atol = 2.0 * eps * tnorm + 4.0 * pivmin;
    for (int g0 = 0; g0 < nsel; g0 += kRRThreads / 16) {
      const int g = g0 + (tid >> 4);
      const bool active = g < nsel;
      double lo = tw.bnd[0], hi = tw.bnd[1];
      for (int it = 0; it < 64; ++it) {
        const double width = hi - lo;
        const bool done = !active || width <= fmax(atol, 4.0 * eps * fmax(fabs(lo), fabs(hi)));
        if (__ballot(!done) == 0ull) break;   // the whole wave is done
        const double step = width / 33.0;
        int c1 = 0, c2 = 0;
        if (!done) sturm_count2(tw.td, tw.te2, n, lo + (2 * sub + 1) * step, lo + (2 * sub + 2) * step, pivmin, c1, c2);
        const unsigned long long m1 = __ballot(!done && c1 > g), m2 = __ballot(!done && c2 > g);
        const unsigned gm = spread16((unsigned)((m1 >> shift) & 0xFFFFull)) |
                            (spread16((unsigned)((m2 >> shift) & 0xFFFFull)) << 1);
        if (!done) {
          if (gm == 0u) {
            lo = lo + 32.0 * step;
          } else {
            const int f = __ffs(gm) - 1;   // first of the 32 points whose count exceeds g
            const double nhi = lo + (f + 1) * step;
            lo = f == 0 ? lo : lo + f * step;
            hi = nhi;
          }
        }
      }
      if (active && sub == 0) theta[g] = 0.5 * (lo + hi);
    }
  }
  __syncthreads();
  // ---- clusters --------------------------------------------------------------------
  if (tid == 0) {
    int nc = 0;
    const double gaptol = 1e3 * eps * tnorm;
    for (int j = 0; j < nsel; ++j)
      if (j == 0 || theta[j] - theta[j - 1] > gaptol) tw.clst[nc++] = j;
    tw.clst[nc] = nsel;
    tw.clst[nsel] = nc;   // the count rides in the last slot (nc <= nsel)
  }
  __syncthreads();
  const int ncl = tw.clst[nsel];
  const int slots = (ld - nsel) / 4;
  if (slots < 1) return 1;
  // ---- 3. inverse iteration ------------------------------------------------------------
  // start vectors (fixed pseudo-random) into the z columns of the slots, all threads
  for (int c0 = 0; c0 < ncl; c0 += slots) {
    const int nc = min(slots, ncl - c0);
    for (int u = tid; u < nc * n; u += kRRThreads) {
      const int t = u / n, i = u - t * n;
      A[i * ld + nsel + 4 * t + 3] = start_value(i, tw.clst[c0 + t]);
    }
    __syncthreads();
    if (tid < nc) {
      double* sc = A + nsel + 4 * tid;   // columns: 0 1/U_ii, 1 U_i,i+1, 2 U_i,i+2, 3 z
      const double ptiny = eps * tnorm + pivmin;
      const int c = c0 + tid;
      const int j0 = tw.clst[c], j1 = c + 1 < ncl ? tw.clst[c + 1] : nsel;
      for (int j = j0; j < j1; ++j) {
        const double lam = theta[j];
        double scale = 1.0;
        bool bad = false;
        if (j > j0)
          for (int i = 0; i < n; ++i) sc[i * ld + 3] = start_value(i, j);
        for (int pass = 0; pass < 2; ++pass) {
          // T - lam I = P L U (partial pivoting, row i + 1 swapped up when its entry in
          // column i is the larger); L^-1 P applied to the right-hand side as the rows
          // are eliminated; pivots of U below ptiny moved to +-ptiny
          double dcur = tw.td[0] - lam, ucur = tw.te[0], zcur = sc[3] * scale;
          for (int i = 0; i + 1 < n; ++i) {
            const double li = tw.te[i], dn = tw.td[i + 1] - lam, un = tw.te[i + 1];
            const double zn = sc[(i + 1) * ld + 3] * scale;
            const bool swap = fabs(dcur) < fabs(li);
            double u0 = swap ? li : dcur;
            if (fabs(u0) < ptiny) u0 = u0 < 0.0 ? -ptiny : ptiny;
            const double ru = rcp_newton2(u0);
            const double f = (swap ? dcur : li) * ru;
            sc[i * ld + 0] = ru;
            sc[i * ld + 1] = swap ? dn : ucur;
            sc[i * ld + 2] = swap ? un : 0.0;
            sc[i * ld + 3] = swap ? zn : zcur;
            const double dnext = swap ? fma(-f, dn, ucur) : fma(-f, ucur, dn);
            const double unext = swap ? -f * un : un;
            const double znext = swap ? fma(-f, zn, zcur) : fma(-f, zcur, zn);
            dcur = dnext;
            ucur = unext;
            zcur = znext;
          }
          if (fabs(dcur) < ptiny) dcur = dcur < 0.0 ? -ptiny : ptiny;
          sc[(n - 1) * ld + 0] = rcp_newton2(dcur);
          sc[(n - 1) * ld + 1] = 0.0;
          sc[(n - 1) * ld + 2] = 0.0;
          sc[(n - 1) * ld + 3] = zcur;
          double z1 = 0.0, z2 = 0.0, zmax = 0.0;
          for (int i = n - 1; i >= 0; --i) {
            const double zi = (sc[i * ld + 3] - sc[i * ld + 1] * z1 - sc[i * ld + 2] * z2) * sc[i * ld + 0];
            sc[i * ld + 3] = zi;
            zmax = fmax(zmax, fabs(zi));
            z2 = z1;
            z1 = zi;
          }
          if (j > j0) {   // Gram-Schmidt against the cluster's earlier vectors
            for (int q = j0; q < j; ++q) {
              double dt = 0.0;
              for (int i = 0; i < n; ++i) dt += sc[i * ld + 3] * A[i * ld + q];
              for (int i = 0; i < n; ++i) sc[i * ld + 3] -= dt * A[i * ld + q];
            }
            zmax = 0.0;
            for (int i = 0; i < n; ++i) zmax = fmax(zmax, fabs(sc[i * ld + 3]));
          }
          if (!(zmax > 0.0) || !isfinite(zmax)) {
            bad = true;
            break;
          }
          scale = 1.0 / zmax;
        }
        if (bad) {
          *tw.flag = 1;
          break;
        }
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) {
          const double zi = sc[i * ld + 3] * scale;
          s2 += zi * zi;
        }
        const double inv = scale / sqrt(s2);
        for (int i = 0; i < n; ++i) A[i * ld + j] = sc[i * ld + 3] * inv;
      }
    }
    __syncthreads();
  }
  if (*tw.flag) return 1;
  // ---- 4. Cholesky-QR: Z <- Z R^-1 with Z^T Z = R^T R ------------------------------------
  double* W = A + nsel;   // nsel x nsel at rows 0..nsel, columns nsel.. (the scratch is free)
  for (int u = tid; u < nsel * nsel; u += kRRThreads) {
    const int r = u / nsel, q = u - r * nsel;
    if (q <= r) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += A[i * ld + r] * A[i * ld + q];
      W[r * ld + q] = s;
    }
  }
  __syncthreads();
  for (int kk = 0; kk < nsel; ++kk) {   // outer-product form: column kk keeps R_kk R_k,i
    const double piv = W[kk * ld + kk];
    if (!(piv > 0.0) || !isfinite(piv)) return 1;   // uniform
    const double rp = 1.0 / piv;
    for (int i = kk + 1 + wave; i < nsel; i += kWaves) {
      const double li = W[i * ld + kk] * rp;
      for (int jj = kk + 1 + lane; jj <= i; jj += 64) W[i * ld + jj] -= li * W[jj * ld + kk];
    }
    __syncthreads();
  }
  for (int i = tid; i < nsel; i += kRRThreads) tw.pp[i] = sqrt(W[i * ld + i]);
  __syncthreads();
  for (int i = wave; i < nsel; i += kWaves)   // W <- R^T
    for (int jj = lane; jj <= i; jj += 64) W[i * ld + jj] /= tw.pp[jj];
  __syncthreads();
  for (int i = tid; i < n; i += kRRThreads) {   // row i: y R = z
    for (int q = 0; q < nsel; ++q) {
      double s = A[i * ld + q];
      for (int p = 0; p < q; ++p) s -= A[i * ld + p] * W[q * ld + p];
      A[i * ld + q] = s / W[q * ld + q];
    }
  }
  __syncthreads();
  // ---- back-transformation Z <- H_0 ... H_{n-3} Z ---------------------------------------
  for (int q0 = 0; q0 < nsel; q0 += kRRThreads / 16) {
    const int q = q0 + (tid >> 4), s = tid & 15;
    const bool act = q < nsel;
    double z[kGroupRows];
#pragma unroll
    for (int t = 0; t < kGroupRows; ++t) {
      const int i = s + 16 * t;
      z[t] = (act && i < n) ? A[i * ld + q] : 0.0;
    }
    for (int k = n - 3; k >= 0; --k) {
      const double beta = tw.tb[k];
      if (beta == 0.0) continue;   // uniform
      double dt = 0.0;
#pragma unroll
      for (int t = 0; t < kGroupRows; ++t) {
        const int i = s + 16 * t;
        if (i > k && i < n) dt += V[i * ld + k] * z[t];
      }
      dt = beta * row16_reduce(dt, SumOp());
#pragma unroll
      for (int t = 0; t < kGroupRows; ++t) {
        const int i = s + 16 * t;
        if (i > k && i < n) z[t] -= dt * V[i * ld + k];
      }
    }
#pragma unroll
    for (int t = 0; t < kGroupRows; ++t) {
      const int i = s + 16 * t;
      if (act && i < n) A[i * ld + q] = z[t];
    }
  }
  __syncthreads();
  return 0;
}

// C = diag(d) L^-T Z for the b columns of Z (columns 0..b of A) on the tridiagonal path;
// L (lower, from G's M half) is staged in V.  16 lanes per column, the solved entry
// passed round the group by a lane shuffle.  CP = C with the X rows zeroed.
__device__ void direct_coefficients(const double* G, int m, int ldg, int b, const double* d, const double* A,
                                    double* V, int ld, double* C, double* CP) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int i = w; i < m; i += kWaves)
    for (int j = lane; j <= i; j += 64) V[i * ld + j] = G[i * ldg + m + j];
  __syncthreads();
  for (int q0 = 0; q0 < b; q0 += kRRThreads / 16) {
    const int q = q0 + (tid >> 4), s = tid & 15;
    const bool act = q < b;
    double z[kGroupRows];
#pragma unroll
    for (int t = 0; t < kGroupRows; ++t) {
      const int i = s + 16 * t;
      z[t] = (act && i < m) ? A[i * ld + q] : 0.0;
    }
    for (int kk = m - 1; kk >= 0; --kk) {
      const int own = kk & 15, tk = kk >> 4;
      double zk = 0.0;
#pragma unroll
      for (int t = 0; t < kGroupRows; ++t)
        if (t == tk) zk = z[t];
      const double xk = __shfl(zk, (lane & 48) | own, 64) * rcp_newton2(V[kk * ld + kk]);
#pragma unroll
      for (int t = 0; t < kGroupRows; ++t) {
        const int i = s + 16 * t;
        if (i < kk) z[t] -= V[kk * ld + i] * xk;
        if (i == kk) z[t] = xk;
      }
    }
#pragma unroll
    for (int t = 0; t < kGroupRows; ++t) {
      const int i = s + 16 * t;
      if (act && i < m) {
        const double cv = d[i] * z[t];
        C[i * b + q] = cv;
        if (CP) CP[i * b + q] = i < b ? 0.0 : cv;
      }
    }
  }
}

template <int TR, int TC>
__global__ void __launch_bounds__(kRRThreads)
rr_kernel(RRArgs a, double* __restrict__ Gall, const double* __restrict__ norms_all, double* __restrict__ lam_all,
          double* __restrict__ Call, double* __restrict__ CPall, double* __restrict__ colscale_all,
          double* __restrict__ status_all) {
  extern __shared__ double sm[];
  const int m = a.m, b = a.b, ld = m + 1;
  const int pid = blockIdx.x;
  double* G = Gall + (int64_t)pid * a.g_stride;
  double* C = Call + (int64_t)pid * a.c_stride;
  double* CP = CPall ? CPall + (int64_t)pid * a.c_stride : nullptr;
  double* lam = lam_all + (int64_t)pid * b;
  double* status = status_all + (int64_t)pid * (kStatusHead + b);
  double* colscale = colscale_all ? colscale_all + (int64_t)pid * b : nullptr;
  double* A = sm;                       // m x ld
  double* V = A + m * ld;               // m x ld
  double* d = V + m * ld;               // m
  double* cs = d + m;                   // m
  double* theta = cs + m;               // m
  double* red = theta + m;              // kRRThreads / 64
  double* res = red + kRRThreads / 64;  // b
  double* tri = res + b;                // 5 m + 8: tridiagonal eigensolver
  int* ip = reinterpret_cast<int*>(tri + 5 * m + 8);
  int* pq = ip;                         // m
  int* valid = pq + m;                  // m
  int* sel = valid + m;                 // b
  int* flags = sel + b;                 // 4: [0] jacobi rotated, [1] cholesky failed, [2] tridiagonal failed
  int* clst = flags + 4;                // b + 1
  const int tid = threadIdx.x;
  const int ldg = 2 * m;

  // ---- 1. residual test, locking mask, convergence -----------------------------
  int converged = 0;
  if (a.have_norms) {
    const double* nrm = norms_all + (int64_t)pid * 2 * b;
    double sc = 0.0;
    for (int j = tid; j < a.k; j += kRRThreads) sc = fmax(sc, fabs(lam[j]));
    sc = fmax(block_max(sc, red), 1e-300);
    double worst = 0.0;
    for (int j = tid; j < b; j += kRRThreads) {
      const double rj = sqrt(fmax(nrm[j], 0.0)) / (fmax(fabs(lam[j]), sc) * sqrt(fmax(nrm[b + j], 1e-300)));
      res[j] = rj;
      if (j < a.k) worst = fmax(worst, rj);
    }
    worst = block_max(worst, red);
    converged = worst < a.tol;
    for (int j = tid; j < b; j += kRRThreads) {
      status[kStatusHead + j] = res[j];
      if (colscale) colscale[j] = res[j] >= a.tol ? 1.0 : 0.0;
    }
    if (tid == 0) {
      status[0] = worst;
      status[1] = converged ? 1.0 : 0.0;
    }
  } else {
    for (int j = tid; j < b; j += kRRThreads) {
      status[kStatusHead + j] = 0.0;
      if (colscale) colscale[j] = 1.0;
    }
    if (tid == 0) {
      status[0] = 0.0;
      status[1] = 0.0;
    }
  }
  if (converged && a.freeze) {
    // keep the converged block: X <- X, P <- 0, Ritz values unchanged
    for (int u = tid; u < m * b; u += kRRThreads) {
      const int i = u / b, j = u - i * b;
      C[u] = i == j ? 1.0 : 0.0;
      if (CP) CP[u] = 0.0;
    }
    if (tid == 0) {
      status[2] = 0.0;
      status[3] = 2.0;
      status[4] = 0.0;
      status[5] = (double)m;
      status[6] = 0.0;
      status[7] = 0.0;
    }
    return;
  }

  // ---- 2. scaling by the M-Gram diagonal; vanished columns ----------------------
  double dmax = 0.0;
  for (int i = tid; i < m; i += kRRThreads) dmax = fmax(dmax, G[i * ldg + m + i]);
  dmax = block_max(dmax, red);
  int nvan = 0;
  for (int i = tid; i < m; i += kRRThreads) {
    const double gi = G[i * ldg + m + i];
    const int ok = gi > 1e-30 * dmax;
    valid[i] = ok;
    d[i] = ok ? 1.0 / sqrt(gi) : 0.0;
    nvan += !ok;
  }
  nvan = (int)block_sum((double)nvan, red);
  // ---- 3-4. Cholesky and T = L^-1 GH L^-T on register tiles ------------------------
  TileT<TR, TC> Tr;
  const int chol_failed = chol_congruence(G, m, ldg, d, valid, a.drop_tol, A, V, ld, cs, theta, Tr);
  int dropped = 0, sweeps = 0, direct = 0;
  const double* coef_basis = nullptr;    // global m x m: L (Cholesky path) or W (SVQB)

  if (!chol_failed) {
    // vanished directions (zero rows and columns of T) get a diagonal above the
    // spectrum of the rest, 2 ||T||_F + 1, which keeps them out of the lowest b
    double f2 = 0.0;
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int c = 0; c < TC; ++c) f2 += Tr[r][c] * Tr[r][c];
    const double big = 2.0 * sqrt(block_sum(f2, red)) + 1.0;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
#pragma unroll
      for (int c = 0; c < TC; ++c) {
        const int i = __builtin_amdgcn_readfirstlane(tid >> 6) + kWaves * r, j = (tid & 63) + 64 * c;
        if (i == j && i < m && !valid[i]) Tr[r][c] = big;
      }
    }
    // ---- 4. the b lowest eigenpairs of T ---------------------------------------------
    TriWork tw{tri, tri + m, tri + 2 * m, tri + 3 * m, tri + 4 * m, tri + 5 * m, clst, flags + 2};
    const int tri_failed = tridiag_lowest(Tr, V, A, m, ld, b, theta, tw);
    if (!tri_failed) {
      direct = 1;
    } else {
      // cyclic Jacobi on T formed again (L from G's M half, the scaled H-Gram from
      // its H half); vanished directions are excluded by `valid` at the selection
      for (int u = tid; u < m * m; u += kRRThreads) {
        const int i = u / m, j = u - i * m;
        A[i * ld + j] = j <= i ? G[i * ldg + m + j] : 0.0;
        const double gh = 0.5 * (G[i * ldg + j] + G[j * ldg + i]) * d[i] * d[j];
        V[i * ld + j] = (valid[i] && valid[j]) ? gh : 0.0;
      }
      __syncthreads();
      congruence_lower(A, V, m, ld);
      set_identity(A, m, ld);
      __syncthreads();
      sweeps = jacobi(V, A, m, ld, cs, pq, flags, red);
      for (int i = tid; i < m; i += kRRThreads) theta[i] = V[i * ld + i];
      __syncthreads();
    }
    coef_basis = G + m;   // L, lower triangular, row stride ldg
  } else {
    // ---- SVQB: eigen-decomposition of the scaled M-Gram ----------------------------
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      const double gm = 0.5 * (G[i * ldg + m + j] + G[j * ldg + m + i]) * d[i] * d[j];
      A[i * ld + j] = (valid[i] && valid[j]) ? gm : (i == j ? 1.0 : 0.0);
    }
    set_identity(V, m, ld);
    __syncthreads();
    sweeps = jacobi(A, V, m, ld, cs, pq, flags, red);
    double mumax = 0.0;
    for (int i = tid; i < m; i += kRRThreads) {
      theta[i] = A[i * ld + i];
      if (valid[i]) mumax = fmax(mumax, theta[i]);
    }
    mumax = block_max(mumax, red);
    // keep mu > drop_tol mu_max among the non-vanished directions (an eigenvector of
    // the decoupled identity rows of a vanished column is that column itself)
    int n_ok = 0;
    for (int i = tid; i < m; i += kRRThreads) n_ok += valid[i] && theta[i] > a.drop_tol * mumax;
    n_ok = (int)block_sum((double)n_ok, red);
    // keep flag in pq (free until the next Jacobi): the b largest if too few pass
    for (int i = tid; i < m; i += kRRThreads) {
      int keep;
      if (n_ok >= b) {
        keep = valid[i] && theta[i] > a.drop_tol * mumax;
      } else {
        int rank = 0;
        for (int j = 0; j < m; ++j)
          if (valid[j] && (theta[j] > theta[i] || (theta[j] == theta[i] && j < i))) ++rank;
        keep = valid[i] && rank < b;
      }
      cs[i] = keep ? 1.0 / sqrt(fmax(theta[i], 1e-300)) : 0.0;
    }
    dropped = (m - nvan) - n_ok;
    __syncthreads();
    // W = V diag(w) into G's M half
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      G[i * ldg + m + j] = V[i * ld + j] * cs[j];
    }
    __syncthreads();
    // T1 = GHs W (A), then B = W^T T1 (V); GHs from G's H half, scaled
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      double s = 0.0;
      if (valid[i])
        for (int l = 0; l < m; ++l)
          if (valid[l]) s += 0.5 * (G[i * ldg + l] + G[l * ldg + i]) * d[i] * d[l] * G[l * ldg + m + j];
      A[i * ld + j] = s;
    }
    __syncthreads();
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      double s = 0.0;
      for (int l = 0; l < m; ++l) s += G[l * ldg + m + i] * A[l * ld + j];
      V[i * ld + j] = s;
    }
    __syncthreads();
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      if (i < j) {
        const double s = 0.5 * (V[i * ld + j] + V[j * ld + i]);
        V[i * ld + j] = s;
        V[j * ld + i] = s;
      }
    }
    for (int i = tid; i < m; i += kRRThreads) valid[i] = cs[i] != 0.0;
    set_identity(A, m, ld);
    __syncthreads();
    sweeps += jacobi(V, A, m, ld, cs, pq, flags, red);
    for (int i = tid; i < m; i += kRRThreads) theta[i] = V[i * ld + i];
    __syncthreads();
    coef_basis = G + m;   // W, row stride ldg
  }

  // ---- 5. coefficients of the b lowest pairs ---------------------------------------
  if (direct) {
    direct_coefficients(G, m, ldg, b, d, A, V, ld, C, CP);
    for (int r = tid; r < b; r += kRRThreads) lam[r] = theta[r];
  } else {
    // (the X columns never vanish, so at least b indices are valid; the default keeps
    // every index in range regardless)
    for (int r = tid; r < b; r += kRRThreads) sel[r] = r;
    __syncthreads();
    select_lowest(theta, valid, m, b, sel);
    __syncthreads();
    // selected eigenvectors (columns of A) -> V[:, 0..b)
    for (int u = tid; u < m * b; u += kRRThreads) {
      const int i = u / b, r = u - i * b;
      V[i * ld + r] = A[i * ld + sel[r]];
    }
    __syncthreads();
    if (!chol_failed) {
      // Cs = L^-T V_sel: backward substitution (unscaled rows, scaled last), L from G
      for (int kk = m - 1; kk >= 0; --kk) {
        const double lkk = coef_basis[kk * ldg + kk];
        for (int u = tid; u < kk * b; u += kRRThreads) {
          const int i = u / b, r = u - i * b;
          V[i * ld + r] -= (coef_basis[kk * ldg + i] / lkk) * V[kk * ld + r];
        }
        __syncthreads();
      }
      for (int u = tid; u < m * b; u += kRRThreads) {
        const int i = u / b, r = u - i * b;
        C[u] = d[i] * V[i * ld + r] / coef_basis[i * ldg + i];
      }
    } else {
      for (int u = tid; u < m * b; u += kRRThreads) {
        const int i = u / b, r = u - i * b;
        double s = 0.0;
        for (int l = 0; l < m; ++l) s += coef_basis[i * ldg + l] * V[l * ld + r];
        C[u] = d[i] * s;
      }
    }
    __syncthreads();
    if (CP) {
      // rows of C are written by this thread in the loops above with the same u
      for (int u = tid; u < m * b; u += kRRThreads) {
        const int i = u / b;
        CP[u] = (dropped > 0 || i < b) ? 0.0 : C[u];
      }
    }
    for (int r = tid; r < b; r += kRRThreads) lam[r] = theta[sel[r]];
  }
  if (tid == 0) {
    status[2] = (double)dropped;
    status[3] = chol_failed ? 1.0 : 0.0;
    status[4] = (double)sweeps;
    status[5] = (double)(m - nvan);
    status[6] = direct ? 1.0 : 0.0;
    status[7] = 0.0;
  }
}

size_t rr_lds_bytes(int m, int b) {
  return (size_t)(2 * m * (m + 1) + 3 * m + kRRThreads / 64 + b + 5 * m + 8) * sizeof(double) +
         (size_t)(2 * m + b + 4 + b + 1) * sizeof(int);
}

}  // namespace

QDS_API int32_t qds_rr_status_len(int32_t b) { return kStatusHead + b; }

QDS_API int qds_rayleigh_ritz(int64_t batch, int32_t m, int32_t b, int32_t k, double* G, int64_t g_stride,
                              const double* norms, double* lam, double* C, double* CP, int64_t c_stride,
                              double* colscale, double* status, double tol, double drop_tol, int32_t freeze,
                              void* stream) {
  QDS_CHECK_ARG(batch >= 1 && G && lam && C && status, "qds_rayleigh_ritz: null pointer or empty batch");
  QDS_CHECK_ARG(m >= 2 && m <= kRRMaxM && m % 2 == 0, "qds_rayleigh_ritz: m = %d must be even and at most %d", m,
                kRRMaxM);
  QDS_CHECK_ARG(b >= 1 && b <= m && k >= 1 && k <= b, "qds_rayleigh_ritz: need 1 <= k <= b <= m (k %d, b %d, m %d)",
                k, b, m);
  QDS_CHECK_ARG(g_stride >= (int64_t)m * 2 * m && c_stride >= (int64_t)m * b, "qds_rayleigh_ritz: strides too small");
  RRArgs a;
  a.m = m;
  a.b = b;
  a.k = k;
  a.have_norms = norms != nullptr;
  a.freeze = freeze;
  a.tol = tol;
  a.drop_tol = drop_tol;
  a.g_stride = g_stride;
  a.c_stride = c_stride;
  const size_t shm = rr_lds_bytes(m, b);
  // register tiles cover m with 6 x 1, 9 x 2 or 12 x 2 slots per thread (m <= 48,
  // <= 72, <= 96); the LDS limit is raised once per instantiation, to the largest
  // problem, so that later launches (also under graph capture) are plain launches
  static std::once_flag once;
  static hipError_t attr_rc = hipSuccess;
  std::call_once(once, [] {
    const int lim = (int)rr_lds_bytes(kRRMaxM, kRRMaxM);
    const void* fns[3] = {reinterpret_cast<const void*>(rr_kernel<6, 1>), reinterpret_cast<const void*>(rr_kernel<9, 2>),
                          reinterpret_cast<const void*>(rr_kernel<12, 2>)};
    for (const void* f : fns) {
      const hipError_t rc = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
      if (rc != hipSuccess) attr_rc = rc;
    }
  });
  QDS_HIP(attr_rc);
  hipStream_t st = qds_stream(stream);
  if (m <= 48)
    hipLaunchKernelGGL((rr_kernel<6, 1>), dim3((unsigned)batch), dim3(kRRThreads), shm, st, a, G, norms, lam, C, CP,
                       colscale, status);
  else if (m <= 72)
    hipLaunchKernelGGL((rr_kernel<9, 2>), dim3((unsigned)batch), dim3(kRRThreads), shm, st, a, G, norms, lam, C, CP,
                       colscale, status);
  else
    hipLaunchKernelGGL((rr_kernel<12, 2>), dim3((unsigned)batch), dim3(kRRThreads), shm, st, a, G, norms, lam, C, CP,
                       colscale, status);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
