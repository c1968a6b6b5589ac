// Rayleigh-Ritz step of the block eigensolver on the host CPU.
//
// Reference operation: the dense projected eigensolve inside scipy's lobpcg and
// ARPACK (qdsim_cython/solvers/eigen_solver.pyx:225-266), which the reference runs
// through LAPACK.  Here the projected pencil of one iteration, G = [S^T H S | S^T M S]
// (m x 2m, m <= 192, FP64, from the device Gram kernel), is solved in the calling
// thread without any library: the solve sits between two GPU launches of every
// iteration while the GPU waits, so its latency is the cost.
//
//   1. columns whose M-Gram diagonal vanished (locked residuals, an empty block)
//      are left out; the rest is scaled to unit M-Gram diagonal;
//   2. Cholesky factor L of the scaled M-Gram, with the reciprocal 1-norm
//      condition number estimated by Hager's method on solves with L; a failed
//      pivot or an estimate below drop_tol returns QDS_HOST_RR_ILLCOND (the caller
//      then uses an eigenvalue-based basis selection instead);
//   3. T = L^-1 H L^-T by two triangular solves (no explicit inverse);
//   4. Householder reduction of T to tridiagonal form (Golub & Van Loan, Alg. 8.3.1);
//   5. the lowest b eigenvalues by bisection on Sturm counts to 1e-12 of the
//      spectrum's span, the b intervals advanced together so that their division
//      chains overlap;
//   6. their eigenvectors by inverse iteration on the tridiagonal matrix (one LU
//      with partial pivoting per eigenvalue, three solves), vectors of eigenvalues
//      closer than 1e-3 ||T|| orthogonalised against each other; the eigenvalues
//      as the vectors' Rayleigh quotients (rounding-level accurate); then the
//      Householder back-transformation, all b vectors per pass;
//   7. C = diag(d) L^-T Z.
#include "qds_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kIllcond = 1;   // the Cholesky path does not apply
constexpr double kEps = 2.220446049250313e-16;

// x . y with four partial sums (fixed order: deterministic, and vectorised)
inline double dot(int n, const double* __restrict__ x, const double* __restrict__ y) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    s0 += x[i] * y[i];
    s1 += x[i + 1] * y[i + 1];
    s2 += x[i + 2] * y[i + 2];
    s3 += x[i + 3] * y[i + 3];
  }
  for (; i < n; ++i) s0 += x[i] * y[i];
  return (s0 + s1) + (s2 + s3);
}

// y -= a x over n contiguous entries (distinct arrays)
inline void axpy_sub(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int i = 0; i < n; ++i) y[i] -= a * x[i];
}

// columns [j0, j0 + W) of row i of the forward substitution: the W running values
// stay in registers over the k loop (W independent FMA chains, one load per term)
template <int W>
inline void lower_solve_cols(int i, int c, int j0, const double* __restrict__ li, double inv,
                             double* __restrict__ Y) {
  double acc[W];
  double* yi = Y + (size_t)i * c + j0;
  for (int q = 0; q < W; ++q) acc[q] = yi[q];
  for (int k = 0; k < i; ++k) {
    const double l = li[k];
    const double* yk = Y + (size_t)k * c + j0;
    for (int q = 0; q < W; ++q) acc[q] -= l * yk[q];
  }
  for (int q = 0; q < W; ++q) yi[q] = acc[q] * inv;
}

// forward substitution L Y = R for a row-major block Y (n x c), in place; every
// entry sees the same operations in the same order as a row-by-row axpy sweep
void lower_solve_rows(int n, int c, const double* L, double* Y) {
  for (int i = 0; i < n; ++i) {
    const double* li = L + (size_t)i * n;
    const double inv = 1.0 / li[i];
    int j0 = 0;
    for (; j0 + 32 <= c; j0 += 32) lower_solve_cols<32>(i, c, j0, li, inv, Y);
    for (; j0 + 8 <= c; j0 += 8) lower_solve_cols<8>(i, c, j0, li, inv, Y);
    for (; j0 < c; ++j0) lower_solve_cols<1>(i, c, j0, li, inv, Y);
  }
}

// backward substitution L^T Y = R for a row-major block Y (n x c), in place
void upper_t_solve_rows(int n, int c, const double* L, double* Y) {
  for (int i = n - 1; i >= 0; --i) {
    double* yi = Y + (size_t)i * c;
    const double inv = 1.0 / L[(size_t)i * n + i];
    for (int j = 0; j < c; ++j) yi[j] *= inv;
    const double* li = L + (size_t)i * n;
    for (int k = 0; k < i; ++k) axpy_sub(c, li[k], yi, Y + (size_t)k * c);
  }
}

// A^-1 x with A = L L^T for one vector; LT = L^T (row-major), so that both
// substitutions run as contiguous axpy updates (vectorised, no reassociation)
void chol_solve(int n, const double* L, const double* LT, double* x) {
  for (int k = 0; k < n; ++k) {            // L y = x: column k of L is row k of LT
    const double yk = x[k] / LT[(size_t)k * n + k];
    x[k] = yk;
    const double* c = LT + (size_t)k * n;
    for (int i = k + 1; i < n; ++i) x[i] -= yk * c[i];
  }
  for (int i = n - 1; i >= 0; --i) {       // L^T z = y: column i of L^T is row i of L
    const double zi = x[i] / L[(size_t)i * n + i];
    x[i] = zi;
    const double* r = L + (size_t)i * n;
    for (int k = 0; k < i; ++k) x[k] -= zi * r[k];
  }
}

// Hager's estimate of ||A^-1||_1 for A = L L^T (A^-T = A^-1), with Higham's
// alternating-sign vector as a second lower bound
double inv_norm1_estimate(int n, const double* L, const double* LT) {
  std::vector<double> x(n, 1.0 / n), y(n), s(n);
  double est = 0.0;
  for (int iter = 0; iter < 5; ++iter) {
    y = x;
    chol_solve(n, L, LT, y.data());
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += std::fabs(y[i]);
    if (iter > 0 && nrm <= est) break;
    est = nrm;
    for (int i = 0; i < n; ++i) s[i] = y[i] >= 0.0 ? 1.0 : -1.0;
    chol_solve(n, L, LT, s.data());
    int jmax = 0;
    double zmax = -1.0, ztx = 0.0;
    for (int i = 0; i < n; ++i) {
      if (std::fabs(s[i]) > zmax) {
        zmax = std::fabs(s[i]);
        jmax = i;
      }
      ztx += s[i] * x[i];
    }
    if (iter > 0 && zmax <= ztx) break;
    std::fill(x.begin(), x.end(), 0.0);
    x[jmax] = 1.0;
  }
  for (int i = 0; i < n; ++i) y[i] = (i % 2 ? -1.0 : 1.0) * (1.0 + (n > 1 ? (double)i / (n - 1) : 0.0));
  chol_solve(n, L, LT, y.data());
  double alt = 0.0;
  for (int i = 0; i < n; ++i) alt += std::fabs(y[i]);
  return std::max(est, 2.0 * alt / (3.0 * n));
}

// Householder reduction of the symmetric T (n x n row-major, both triangles) to
// tridiagonal form: diagonal a, subdiagonal e.  Reflector k is I - beta_k v v^T on
// rows k+1 .. n-1 (v[k+1] = 1), kept in row k of V; T = Q Tri Q^T with
// Q = H_0 H_1 ... H_{n-3}.
void tridiagonalize(int n, double* T, double* a, double* e, std::vector<double>& V, std::vector<double>& beta) {
  V.assign((size_t)n * n, 0.0);
  beta.assign(n, 0.0);
  std::vector<double> v(n), p(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int len = n - k - 1;
    const double x0 = T[(size_t)(k + 1) * n + k];
    double sig = 0.0;
    for (int i = 1; i < len; ++i) {
      const double xi = T[(size_t)(k + 1 + i) * n + k];
      sig += xi * xi;
    }
    if (sig == 0.0) {   // column already in tridiagonal form
      e[k] = x0;
      continue;
    }
    // v = x - ||x|| e_1, scaled to v[0] = 1 (its first entry formed without cancellation)
    const double mu = std::sqrt(x0 * x0 + sig);
    const double v0 = x0 <= 0.0 ? x0 - mu : -sig / (x0 + mu);
    const double bk = 2.0 * v0 * v0 / (sig + v0 * v0);
    v[0] = 1.0;
    for (int i = 1; i < len; ++i) v[i] = T[(size_t)(k + 1 + i) * n + k] / v0;
    e[k] = mu;
    // T22 <- H T22 H = T22 - v w^T - w v^T, p = beta T22 v, w = p - (beta/2)(p.v) v
    // p = beta T22 v as a sum of rows (T22 is symmetric): contiguous axpys
    for (int i = 0; i < len; ++i) p[i] = 0.0;
    for (int j = 0; j < len; ++j) {
      const double* tj = T + (size_t)(k + 1 + j) * n + (k + 1);
      const double vj = v[j];
      for (int i = 0; i < len; ++i) p[i] += tj[i] * vj;
    }
    double pv = 0.0;
    for (int i = 0; i < len; ++i) {
      p[i] *= bk;
      pv += p[i] * v[i];
    }
    const double half = 0.5 * bk * pv;
    for (int i = 0; i < len; ++i) w[i] = p[i] - half * v[i];
    for (int i = 0; i < len; ++i) {
      double* ti = T + (size_t)(k + 1 + i) * n + (k + 1);
      const double vi = v[i], wi = w[i];
      for (int j = 0; j < len; ++j) ti[j] -= vi * w[j] + wi * v[j];
    }
    beta[k] = bk;
    double* vk = V.data() + (size_t)k * n;
    for (int i = 0; i < len; ++i) vk[k + 1 + i] = v[i];
  }
  for (int i = 0; i < n; ++i) a[i] = T[(size_t)i * n + i];
  e[n - 2] = T[(size_t)(n - 1) * n + (n - 2)];
}

// The lowest b eigenvalues of the tridiagonal (a, e) by bisection, to 1e-12 of
// the spectrum's span.  Interval j
// holds eigenvalue j: (number of eigenvalues below L_j) <= j < (number below H_j);
// the count at x is the number of negative pivots of the LDL^T factorisation of
// T - x I.  All b counts run in one pass over the rows.
void lowest_eigenvalues(int n, int b, const double* a, const double* e, double* lam) {
  std::vector<double> e2(n, 0.0);
  double lo = a[0], hi = a[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
    if (i + 1 < n) e2[i] = e[i] * e[i];
  }
  const double span = std::max(std::fabs(lo), std::fabs(hi));
  const double pivmin = std::max(1e-290, kEps * kEps * span * span);
  lo -= 2.0 * kEps * span + pivmin;
  hi += 2.0 * kEps * span + pivmin;
  std::vector<double> L(b, lo), H(b, hi), x(b), q(b);
  std::vector<int> cnt(b);
  // to 1e-12 of the spectrum's span: the shifts of the inverse iteration, whose
  // vectors then give the eigenvalues to rounding as Rayleigh quotients
  const double tol = std::max(2.0 * kEps * span, 1e-12 * span);
  for (int it = 0; it < 80; ++it) {
    double width = 0.0;
    for (int j = 0; j < b; ++j) {
      width = std::max(width, H[j] - L[j]);
      x[j] = 0.5 * (L[j] + H[j]);
      double t = a[0] - x[j];
      t = std::fabs(t) < pivmin ? -pivmin : t;
      q[j] = t;
      cnt[j] = t < 0.0;
    }
    if (width <= tol) break;
    for (int i = 1; i < n; ++i) {
      const double ai = a[i], ei = e2[i - 1];
      for (int j = 0; j < b; ++j) {
        double t = (ai - x[j]) - ei / q[j];
        t = std::fabs(t) < pivmin ? -pivmin : t;
        q[j] = t;
        cnt[j] += t < 0.0;
      }
    }
    for (int j = 0; j < b; ++j) {
      if (cnt[j] > j) H[j] = x[j];
      else L[j] = x[j];
    }
  }
  for (int j = 0; j < b; ++j) lam[j] = 0.5 * (L[j] + H[j]);
}

// LU with partial pivoting of the tridiagonal T - s I (a row may swap with the
// next: U has two superdiagonals u1, u2 and reciprocal pivots r); tiny replaces a
// vanishing pivot
struct TriLU {
  std::vector<double> r, u1, u2, l;
  std::vector<char> sw;
  explicit TriLU(int n) : r(n), u1(n), u2(n), l(n), sw(n) {}

  void factor(int n, const double* a, const double* e, double s, double tiny) {
    double cd = a[0] - s, cu = n > 1 ? e[0] : 0.0;   // current row: columns i, i+1
    for (int i = 0; i + 1 < n; ++i) {
      const double sub = e[i], nd = a[i + 1] - s, nu = i + 2 < n ? e[i + 1] : 0.0;
      if (std::fabs(cd) >= std::fabs(sub)) {
        sw[i] = 0;
        const double piv = std::fabs(cd) < tiny ? (cd < 0.0 ? -tiny : tiny) : cd;
        l[i] = sub / piv;
        r[i] = 1.0 / piv;
        u1[i] = cu;
        u2[i] = 0.0;
        cd = nd - l[i] * cu;
        cu = nu;
      } else {
        sw[i] = 1;
        l[i] = cd / sub;
        r[i] = 1.0 / sub;
        u1[i] = nd;
        u2[i] = nu;
        cd = cu - l[i] * nd;
        cu = -l[i] * nu;
      }
    }
    const double piv = std::fabs(cd) < tiny ? (cd < 0.0 ? -tiny : tiny) : cd;
    r[n - 1] = 1.0 / piv;
    u1[n - 1] = u2[n - 1] = 0.0;
  }

  void solve(int n, double* z) const {
    for (int i = 0; i + 1 < n; ++i) {
      if (sw[i]) std::swap(z[i], z[i + 1]);
      z[i + 1] -= l[i] * z[i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double v = z[i];
      if (i + 1 < n) v -= u1[i] * z[i + 1];
      if (i + 2 < n) v -= u2[i] * z[i + 2];
      z[i] = v * r[i];
    }
  }
};

// Inverse iteration for eigenvalues with no neighbour within the cluster distance:
// no orthogonalisation, so the vectors are independent and nb of them advance
// together (rows outer, vectors inner: the division and substitution chains of
// different vectors overlap).  Every vector sees the operations of TriLU::factor,
// TriLU::solve, dot and the Rayleigh quotient of the one-vector path, in the same
// order.  Z: nb vectors interleaved (entry i of vector t at Z[i * nb + t]), already
// holding the start vectors; s: shifts; rq: Rayleigh quotients out.
int inverse_iteration_batch(int n, int nb, const double* a, const double* e, const double* s, double tiny,
                            double* Z, double* rq) {
  std::vector<double> r((size_t)n * nb), u1((size_t)n * nb), u2((size_t)n * nb), l((size_t)n * nb);
  std::vector<char> sw((size_t)n * nb);
  std::vector<double> cd(nb), cu(nb);
  for (int t = 0; t < nb; ++t) {
    cd[t] = a[0] - s[t];
    cu[t] = n > 1 ? e[0] : 0.0;
  }
  for (int i = 0; i + 1 < n; ++i) {
    const double sub = e[i], nu = i + 2 < n ? e[i + 1] : 0.0;
    for (int t = 0; t < nb; ++t) {
      const size_t it = (size_t)i * nb + t;
      const double nd = a[i + 1] - s[t];
      if (std::fabs(cd[t]) >= std::fabs(sub)) {
        sw[it] = 0;
        const double piv = std::fabs(cd[t]) < tiny ? (cd[t] < 0.0 ? -tiny : tiny) : cd[t];
        l[it] = sub / piv;
        r[it] = 1.0 / piv;
        u1[it] = cu[t];
        u2[it] = 0.0;
        cd[t] = nd - l[it] * cu[t];
        cu[t] = nu;
      } else {
        sw[it] = 1;
        l[it] = cd[t] / sub;
        r[it] = 1.0 / sub;
        u1[it] = nd;
        u2[it] = nu;
        cd[t] = cu[t] - l[it] * nd;
        cu[t] = -l[it] * nu;
      }
    }
  }
  for (int t = 0; t < nb; ++t) {
    const size_t it = (size_t)(n - 1) * nb + t;
    const double piv = std::fabs(cd[t]) < tiny ? (cd[t] < 0.0 ? -tiny : tiny) : cd[t];
    r[it] = 1.0 / piv;
    u1[it] = u2[it] = 0.0;
  }
  std::vector<double> p0(nb), p1(nb), p2(nb), p3(nb);
  const int n4 = n / 4 * 4;
  for (int iter = 0; iter < 3; ++iter) {
    for (int i = 0; i + 1 < n; ++i) {
      double* zi = Z + (size_t)i * nb;
      double* zn = zi + nb;
      for (int t = 0; t < nb; ++t) {
        const size_t it = (size_t)i * nb + t;
        if (sw[it]) std::swap(zi[t], zn[t]);
        zn[t] -= l[it] * zi[t];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      double* zi = Z + (size_t)i * nb;
      for (int t = 0; t < nb; ++t) {
        const size_t it = (size_t)i * nb + t;
        double v = zi[t];
        if (i + 1 < n) v -= u1[it] * zi[nb + t];
        if (i + 2 < n) v -= u2[it] * zi[2 * nb + t];
        zi[t] = v * r[it];
      }
    }
    // norms with dot's four partial sums
    std::fill(p0.begin(), p0.end(), 0.0);
    std::fill(p1.begin(), p1.end(), 0.0);
    std::fill(p2.begin(), p2.end(), 0.0);
    std::fill(p3.begin(), p3.end(), 0.0);
    for (int i = 0; i < n4; i += 4)
      for (int t = 0; t < nb; ++t) {
        const double z0 = Z[(size_t)i * nb + t], z1 = Z[(size_t)(i + 1) * nb + t];
        const double z2 = Z[(size_t)(i + 2) * nb + t], z3 = Z[(size_t)(i + 3) * nb + t];
        p0[t] += z0 * z0;
        p1[t] += z1 * z1;
        p2[t] += z2 * z2;
        p3[t] += z3 * z3;
      }
    for (int i = n4; i < n; ++i)
      for (int t = 0; t < nb; ++t) {
        const double zi = Z[(size_t)i * nb + t];
        p0[t] += zi * zi;
      }
    for (int t = 0; t < nb; ++t) {
      const double nrm = std::sqrt((p0[t] + p1[t]) + (p2[t] + p3[t]));
      if (!(nrm > 0.0) || !std::isfinite(nrm)) return kIllcond;
      p0[t] = 1.0 / nrm;
    }
    for (int i = 0; i < n; ++i)
      for (int t = 0; t < nb; ++t) Z[(size_t)i * nb + t] *= p0[t];
  }
  for (int t = 0; t < nb; ++t) {
    const double zl = Z[(size_t)(n - 1) * nb + t];
    rq[t] = a[n - 1] * zl * zl;
  }
  for (int i = 0; i + 1 < n; ++i)
    for (int t = 0; t < nb; ++t) {
      const double zi = Z[(size_t)i * nb + t], zn = Z[(size_t)(i + 1) * nb + t];
      rq[t] += zi * (a[i] * zi + 2.0 * e[i] * zn);
    }
  return QDS_OK;
}

}  // namespace

// Lowest b Ritz pairs of the pencil (G[:, :m], G[:, m:2m]) (m x 2m row-major FP64),
// as in the file comment.  theta: b values, ascending; C: m x b row-major (rows of
// vanished columns zero), so that S C are the Ritz vectors of the basis S.
// Returns QDS_OK, 1 when the Cholesky path does not apply (a vanishing or
// negative pivot, reciprocal condition estimate below drop_tol, fewer than b
// usable columns), or QDS_EINVAL.
QDS_API int qds_host_rayleigh_ritz(int32_t m, int32_t b, const double* G, double drop_tol, double* theta,
                                   double* C) {
  QDS_CHECK_ARG(G && theta && C && m >= 1 && m <= 192 && b >= 1 && b <= m, "qds_host_rayleigh_ritz: bad arguments");
  const int m2 = 2 * m;
  // 1. vanished columns and the scaling to unit M-Gram diagonal
  double dmax = 0.0;
  for (int i = 0; i < m; ++i) dmax = std::max(dmax, G[(size_t)i * m2 + m + i]);
  if (!(dmax > 0.0)) return kIllcond;
  std::vector<int> idx;
  idx.reserve(m);
  for (int i = 0; i < m; ++i)
    if (G[(size_t)i * m2 + m + i] > 1e-30 * dmax) idx.push_back(i);
  const int n = (int)idx.size();
  if (n < b) return kIllcond;
  std::vector<double> d(n), A((size_t)n * n), Hs((size_t)n * n);
  for (int i = 0; i < n; ++i) d[i] = 1.0 / std::sqrt(G[(size_t)idx[i] * m2 + m + idx[i]]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t ij = (size_t)idx[i] * m2 + idx[j], ji = (size_t)idx[j] * m2 + idx[i];
      const double dd = 0.5 * d[i] * d[j];
      A[(size_t)i * n + j] = (G[ij + m] + G[ji + m]) * dd;
      Hs[(size_t)i * n + j] = (G[ij] + G[ji]) * dd;
    }
  double anorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(A[(size_t)i * n + j]);
    anorm = std::max(anorm, s);
  }
  // 2. Cholesky factor (lower, in place) and the condition estimate
  // right-looking: column j is final once the rank-1 updates of columns < j are
  // in; the updates run over contiguous rows of the trailing lower triangle
  std::vector<double>& L = A;
  std::vector<double> col(n);
  for (int j = 0; j < n; ++j) {
    double* lj = L.data() + (size_t)j * n;
    const double s = lj[j];
    if (!(s > 0.0)) return kIllcond;
    const double ljj = std::sqrt(s);
    lj[j] = ljj;
    const double inv = 1.0 / ljj;
    for (int i = j + 1; i < n; ++i) col[i] = (L[(size_t)i * n + j] *= inv);
    for (int i = j + 1; i < n; ++i) {
      double* li = L.data() + (size_t)i * n;
      const double lij = col[i];
      for (int c = j + 1; c <= i; ++c) li[c] -= lij * col[c];
    }
    for (int k = j + 1; k < n; ++k) lj[k] = 0.0;
  }
  std::vector<double> LT((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) LT[(size_t)j * n + i] = L[(size_t)i * n + j];
  const double rcond = 1.0 / (anorm * inv_norm1_estimate(n, L.data(), LT.data()));
  if (!(rcond >= drop_tol)) return kIllcond;
  // 3. T = L^-1 Hs L^-T: Y = L^-1 Hs, then T = L^-1 Y^T, symmetrised
  lower_solve_rows(n, n, L.data(), Hs.data());
  std::vector<double> T((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) T[(size_t)i * n + j] = Hs[(size_t)j * n + i];
  lower_solve_rows(n, n, L.data(), T.data());
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double v = 0.5 * (T[(size_t)i * n + j] + T[(size_t)j * n + i]);
      T[(size_t)i * n + j] = T[(size_t)j * n + i] = v;
    }
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += std::fabs(T[(size_t)i * n + j]);
    tnorm = std::max(tnorm, s);
  }
  if (!(tnorm < 1e300)) return kIllcond;
  std::vector<double> lam(b), Z((size_t)b * n);   // row j of Z: eigenvector j of T
  std::vector<double> V, beta;                      // Householder reflectors of step 4
  if (n == 1) {
    lam[0] = T[0];
    Z[0] = 1.0;
  } else {
    // 4.-5. tridiagonal form, lowest b eigenvalues
    std::vector<double> a(n), e(n, 0.0);
    tridiagonalize(n, T.data(), a.data(), e.data(), V, beta);
    lowest_eigenvalues(n, b, a.data(), e.data(), lam.data());
    // 6. inverse iteration; a cluster is a run of eigenvalues closer than 1e-3 ||T||
    TriLU lu(n);
    const double tiny = std::max(kEps * tnorm, 1e-300);
    const double close = 1e-3 * tnorm;
    // eigenvalues farther than close (with a margin far above the bisection
    // tolerance, so that the one-vector loop below would also find them alone)
    // from both neighbours go through the batched path first
    const double margin = 1e-9 * tnorm;
    std::vector<char> alone(b, 0);
    std::vector<int> J;
    for (int j = 0; j < b; ++j) {
      const bool lgap = j == 0 || lam[j] - lam[j - 1] > close + margin;
      const bool rgap = j + 1 == b || lam[j + 1] - lam[j] > close + margin;
      if (lgap && rgap) {
        alone[j] = 1;
        J.push_back(j);
      }
    }
    if (!J.empty()) {
      const int nb = (int)J.size();
      std::vector<double> Zb((size_t)n * nb), sb(nb), rqb(nb);
      for (int t = 0; t < nb; ++t) {
        const int j = J[t];
        sb[t] = lam[j];
        uint32_t seed = 2654435761u * (uint32_t)(j + 1);
        for (int i = 0; i < n; ++i) {
          seed = seed * 1664525u + 1013904223u;
          Zb[(size_t)i * nb + t] = 0.5 + (double)(seed >> 8) / 16777216.0;
        }
      }
      if (inverse_iteration_batch(n, nb, a.data(), e.data(), sb.data(), tiny, Zb.data(), rqb.data()) != QDS_OK)
        return kIllcond;
      for (int t = 0; t < nb; ++t) {
        double* z = Z.data() + (size_t)J[t] * n;
        for (int i = 0; i < n; ++i) z[i] = Zb[(size_t)i * nb + t];
        lam[J[t]] = rqb[t];
      }
    }
    int cstart = 0;
    double prev = 0.0;
    for (int j = 0; j < b; ++j) {
      double* z = Z.data() + (size_t)j * n;
      if (alone[j]) {   // a cluster of its own, done above
        cstart = j;
        prev = 0.0;
        continue;
      }
      if (j > 0 && lam[j] - lam[j - 1] > close) cstart = j;
      double s = lam[j];
      // coincident eigenvalues of a cluster get shifts a few ulps of ||T|| apart
      if (j > cstart && s - prev < 10.0 * kEps * tnorm) s = prev + 10.0 * kEps * tnorm;
      prev = s;
      uint32_t seed = 2654435761u * (uint32_t)(j + 1);
      for (int i = 0; i < n; ++i) {
        seed = seed * 1664525u + 1013904223u;
        z[i] = 0.5 + (double)(seed >> 8) / 16777216.0;
      }
      lu.factor(n, a.data(), e.data(), s, tiny);
      for (int iter = 0; iter < 3; ++iter) {
        lu.solve(n, z);
        for (int q = cstart; q < j; ++q) {
          const double* zq = Z.data() + (size_t)q * n;
          axpy_sub(n, dot(n, zq, z), zq, z);
        }
        const double nrm = std::sqrt(dot(n, z, z));
        if (!(nrm > 0.0) || !std::isfinite(nrm)) return kIllcond;
        const double inv = 1.0 / nrm;
        for (int i = 0; i < n; ++i) z[i] *= inv;
      }
      // the eigenvalue as the Rayleigh quotient of the unit vector
      double rq = a[n - 1] * z[n - 1] * z[n - 1];
      for (int i = 0; i + 1 < n; ++i) rq += z[i] * (a[i] * z[i] + 2.0 * e[i] * z[i + 1]);
      lam[j] = rq;
    }
    // ascending order (a Rayleigh quotient may pass its neighbour's inside a cluster)
    for (int j = 1; j < b; ++j)
      for (int q = j; q > 0 && lam[q] < lam[q - 1]; --q) {
        std::swap(lam[q], lam[q - 1]);
        std::swap_ranges(Z.begin() + (size_t)q * n, Z.begin() + (size_t)(q + 1) * n, Z.begin() + (size_t)(q - 1) * n);
      }
  }
  // Y = Z^T (n x b, row-major), back-transformed by the reflectors, the last one
  // first: per reflector one contiguous pass computing the b dot products and one
  // applying the rank-1 update
  std::vector<double> Y((size_t)n * b);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < b; ++j) Y[(size_t)i * b + j] = Z[(size_t)j * n + i];
  if (n > 1) {
    std::vector<double> w(b);
    for (int k = n - 3; k >= 0; --k) {
      if (beta[k] == 0.0) continue;
      const double* vk = V.data() + (size_t)k * n;
      std::fill(w.begin(), w.end(), 0.0);
      for (int i = k + 1; i < n; ++i) {
        const double vi = vk[i];
        const double* yi = Y.data() + (size_t)i * b;
        for (int j = 0; j < b; ++j) w[j] += vi * yi[j];
      }
      for (int j = 0; j < b; ++j) w[j] *= beta[k];
      for (int i = k + 1; i < n; ++i) axpy_sub(b, vk[i], w.data(), Y.data() + (size_t)i * b);
    }
  }
  // 7. C = diag(d) L^-T Z, scattered into the kept rows
  upper_t_solve_rows(n, b, L.data(), Y.data());
  std::memset(C, 0, sizeof(double) * (size_t)m * b);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < b; ++j) C[(size_t)idx[i] * b + j] = d[i] * Y[(size_t)i * b + j];
  for (int j = 0; j < b; ++j) {
    if (!std::isfinite(lam[j])) return kIllcond;
    theta[j] = lam[j];
  }
  return QDS_OK;
}
