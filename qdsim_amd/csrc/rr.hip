// Rayleigh-Ritz of the block eigensolver on the GPU (FP64), one workgroup per
// problem of a batch.
//
// Reference: the dense steps inside scipy's lobpcg / ARPACK
// (qdsim_cython/solvers/eigen_solver.pyx:225-266): the projected pencil
// (S^T H S, S^T M S) of the current basis S = [X | W | P] is solved for its lowest
// b eigenpairs.  The projected matrices are m x m with m <= 96, so a whole problem
// fits one workgroup's LDS (A and V: 2 m (m + 1) doubles, 149 KB at m = 96), and
// the solve never leaves the device: the block iteration runs without a host
// round trip and can be captured in a HIP graph.
//
// Per problem, with G = [GH | GM] (m x 2m, row-major) from the Gram kernel:
//  1. residual test of the current block from its column norms (when given):
//     res_j = ||r_j|| / (max(|lam_j|, s) ||M x_j||), s = max_{j<k} |lam_j|; the
//     locking mask (res_j >= tol) scales the next residual; converged when
//     max_{j<k} res_j < tol, and with `freeze` a converged problem keeps its block
//     (C = [I; 0]) from then on;
//  2. scale by the M-Gram diagonal; columns whose diagonal vanished (locked
//     residuals, a zero P block) are decoupled and never selected;
//  3. Cholesky of the scaled M-Gram; a pivot below drop_tol sends the problem to
//     SVQB instead (eigen-decomposition of the M-Gram, directions with eigenvalue
//     below drop_tol times the largest dropped, never fewer than b kept);
//  4. T = L^-1 GH L^-T by two forward substitutions (column-parallel);
//  5. cyclic Jacobi on T: round-robin ordering (m/2 disjoint rotations per round),
//     each round one rotation-parameter pass and one pass that applies all
//     rotations to disjoint 2 x 2 blocks of T (symmetric, both triangles written)
//     and to the rows of V; a sweep that rotates nothing, or an off-diagonal norm
//     below 1e-16 of the diagonal's, ends it;
//  6. the b lowest eigenvalues by rank (ties by index), C = diag(d) L^-T V_sel by
//     backward substitution; CP = C with the X rows zeroed (all zero when
//     directions were dropped: the P block restarts).
// G's M half is overwritten (L, or the SVQB basis, is kept there).
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

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < kRRThreads / 64; ++i) s += red[i];
  return s;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < kRRThreads / 64; ++i) s = fmax(s, red[i]);
  return s;
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
  int* ip = reinterpret_cast<int*>(res + b);
  int* pq = ip;                         // m
  int* valid = pq + m;                  // m
  int* sel = valid + m;                 // b
  int* flags = sel + b;                 // 4: [0] jacobi rotated, [1] cholesky failed
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
  // A = scaled M-Gram (vanished: identity rows), V = scaled H-Gram (vanished: zero)
  for (int u = tid; u < m * m; u += kRRThreads) {
    const int i = u / m, j = u - i * m;
    const double dd = d[i] * d[j];
    const double gm = 0.5 * (G[i * ldg + m + j] + G[j * ldg + m + i]) * dd;
    const double gh = 0.5 * (G[i * ldg + j] + G[j * ldg + i]) * dd;
    A[i * ld + j] = (valid[i] && valid[j]) ? gm : (i == j ? 1.0 : 0.0);
    V[i * ld + j] = (valid[i] && valid[j]) ? gh : 0.0;
  }
  if (tid == 0) flags[1] = 0;
  __syncthreads();

  // ---- 3. Cholesky (outer product on the lower triangle, unscaled columns) ----------
  for (int kk = 0; kk < m; ++kk) {
    const double piv = A[kk * ld + kk];
    if (!(piv > a.drop_tol)) {   // uniform: every thread reads the same value
      if (tid == 0) flags[1] = 1;
      break;
    }
    const int nt = m - kk - 1;
    for (int u = tid; u < nt * nt; u += kRRThreads) {
      const int i = kk + 1 + u / nt, j = kk + 1 + u % nt;
      if (j <= i) A[i * ld + j] -= A[i * ld + kk] * A[j * ld + kk] / piv;
    }
    __syncthreads();
  }
  __syncthreads();
  const int chol_failed = flags[1];
  int dropped = 0, sweeps = 0;
  const double* coef_basis = nullptr;    // global m x m: L (Cholesky path) or W (SVQB)

  if (!chol_failed) {
    // L[i][j] = A[i][j] / sqrt(A[j][j]) for i >= j; keep L in G's M half
    for (int i = tid; i < m; i += kRRThreads) theta[i] = sqrt(A[i * ld + i]);
    __syncthreads();
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      const double l = j <= i ? A[i * ld + j] / theta[j] : 0.0;
      G[i * ldg + m + j] = l;
    }
    for (int u = tid; u < m * m; u += kRRThreads) {
      const int i = u / m, j = u - i * m;
      if (j <= i) A[i * ld + j] /= theta[j];
    }
    __syncthreads();
    // ---- 4. T = L^-1 GH L^-T: forward substitutions on V (unscaled rows, scaled last)
    for (int pass = 0; pass < 2; ++pass) {
      for (int kk = 0; kk < m; ++kk) {
        const double lkk = A[kk * ld + kk];
        const int nt = m - kk - 1;
        for (int u = tid; u < nt * m; u += kRRThreads) {
          const int i = kk + 1 + u / m, j = u % m;
          V[i * ld + j] -= (A[i * ld + kk] / lkk) * V[kk * ld + j];
        }
        __syncthreads();
      }
      for (int u = tid; u < m * m; u += kRRThreads) {
        const int i = u / m, j = u - i * m;
        V[i * ld + j] /= A[i * ld + i];
      }
      __syncthreads();
      // transpose (the second pass then gives (L^-1 (L^-1 GH)^T)^T = T, symmetric)
      for (int u = tid; u < m * m; u += kRRThreads) {
        const int i = u / m, j = u - i * m;
        if (i < j) {
          const double x = V[i * ld + j], y = V[j * ld + i];
          if (pass == 0) {
            V[i * ld + j] = y;
            V[j * ld + i] = x;
          } else {
            const double s = 0.5 * (x + y);
            V[i * ld + j] = s;
            V[j * ld + i] = s;
          }
        }
      }
      __syncthreads();
    }
    // ---- 5. Jacobi on T (in V), eigenvectors in A --------------------------------
    set_identity(A, m, ld);
    __syncthreads();
    sweeps = jacobi(V, A, m, ld, cs, pq, flags, red);
    for (int i = tid; i < m; i += kRRThreads) theta[i] = V[i * ld + i];
    __syncthreads();
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

  // ---- 6. the b lowest pairs, coefficients ---------------------------------------
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
  if (tid == 0) {
    status[2] = (double)dropped;
    status[3] = chol_failed ? 1.0 : 0.0;
    status[4] = (double)sweeps;
    status[5] = (double)(m - nvan);
    status[6] = 0.0;
    status[7] = 0.0;
  }
}

size_t rr_lds_bytes(int m, int b) {
  return (size_t)(2 * m * (m + 1) + 3 * m + kRRThreads / 64 + b) * sizeof(double) +
         (size_t)(2 * m + b + 4) * sizeof(int);
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
  // the LDS limit is raised once, to the largest problem, so that later launches
  // (also under graph capture) are plain kernel launches
  static std::once_flag once;
  static hipError_t attr_rc = hipSuccess;
  std::call_once(once, [] {
    attr_rc = hipFuncSetAttribute(reinterpret_cast<const void*>(rr_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)rr_lds_bytes(kRRMaxM, kRRMaxM));
  });
  QDS_HIP(attr_rc);
  hipLaunchKernelGGL(rr_kernel, dim3((unsigned)batch), dim3(kRRThreads), shm, qds_stream(stream), a, G, norms, lam, C,
                     CP, colscale, status);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
