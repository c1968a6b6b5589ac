// Poisson--drift--diffusion on a P1 triangle mesh (gfx950): the node and row
// kernels of one Gummel step.
//
// Reference: FullPoissonDriftDiffusionSolver (backend/src/full_poisson_dd_solver.cpp).
// The reference evaluates every quantity in per-node / per-element loops on the
// host (the field and the currents even search all elements for each node,
// :524-583, :1068-1159), assembles COO triplets and factorises.  Here:
//
//  * node kernels (one thread per node): initial carriers (:216-260), carrier
//    statistics (:656-718), ohmic contact values (:778-839), the charge density
//    (:267-288), the potential
//    damping and the non-finite revert + norms of the Gummel test (:1168-1173,
//    :1383-1408);
//  * node-average kernels (one thread per node over its elements, ascending
//    element id as the reference's loop visits them): the field -grad phi
//    (:524-583) and the currents q mu n grad phi_n / q mu p grad phi_p (:1068-1159);
//  * row kernels (one thread per CSR row): each row's values of the continuity
//    matrices of electrons and holes and their loads (:847-995), or of the
//    Poisson matrix and its load (:296-418), computed from the row node's
//    elements and written once into the P1 pattern, with the Dirichlet row
//    replacement (:418-471, :997-1035) applied in the same pass.  With
//    ldlt_lower = 1 a row holds the matrix Eigen's SimplicialLDLT (UpLo = Lower)
//    factors for the row-replaced A: S(i, j) = A(i, j) for j <= i and A(j, i)
//    for j > i; row i reads the (j, i) entries from the same elements, so no
//    transpose pass is needed.
//
// All arrays are FP64 in HBM, node-major; fields and currents are [n][2].
#include "qds_common.h"
#include "../../include/qdsim_hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 256;
constexpr int kMatStride = 8;   // N_c, N_v, E_g, mu_n, mu_p, eps_r, n_i, unused

__device__ __forceinline__ double std_max(double a, double b) { return a < b ? b : a; }

// std::exp(eta) / (1 + 0.27 exp(eta)) for eta <= 0, eta^1.5 / (1.5 + 0.27 eta^1.5)
// otherwise (:690-697)
__device__ __forceinline__ double fd_half(double eta) {
  if (eta <= 0.0) {
    const double e = exp(eta);
    return e / (1.0 + 0.27 * e);
  }
  const double s = pow(eta, 1.5);
  return s / (1.5 + 0.27 * s);
}

struct Tri {
  int a[3];
  double x[3], y[3];
};

__device__ __forceinline__ Tri load_tri(const qds_dd_mesh& m, int64_t e) {
  Tri t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.a[k] = m.conn[3 * e + k];
    t.x[k] = m.x[t.a[k]];
    t.y[k] = m.y[t.a[k]];
  }
  return t;
}

// gradients from J^-T applied to the reference gradients (-1,-1), (1,0), (0,1)
// and the area 0.5 |det J| (:322-345, :881-905)
__device__ __forceinline__ double tri_grad_jac(const Tri& t, double gx[3], double gy[3]) {
  const double j00 = t.x[1] - t.x[0], j01 = t.x[2] - t.x[0];
  const double j10 = t.y[1] - t.y[0], j11 = t.y[2] - t.y[0];
  const double det = j00 * j11 - j01 * j10;
  const double i00 = j11 / det, i01 = -j01 / det, i10 = -j10 / det, i11 = j00 / det;   // J^-1
  // grad N = (J^-1)^T r
  gx[0] = -i00 - i10;
  gy[0] = -i01 - i11;
  gx[1] = i00;
  gy[1] = i01;
  gx[2] = i10;
  gy[2] = i11;
  return 0.5 * fabs((t.x[1] - t.x[0]) * (t.y[2] - t.y[0]) - (t.x[2] - t.x[0]) * (t.y[1] - t.y[0]));
}

// gradients (y2 - y3) / 2A ... of the field and current loops (:544-551, :1101-1108)
__device__ __forceinline__ void tri_grad_area(const Tri& t, double gx[3], double gy[3]) {
  const double area = 0.5 * fabs((t.x[1] - t.x[0]) * (t.y[2] - t.y[0]) - (t.x[2] - t.x[0]) * (t.y[1] - t.y[0]));
  const double a2 = 2.0 * area;
  gx[0] = (t.y[1] - t.y[2]) / a2;
  gy[0] = (t.x[2] - t.x[1]) / a2;
  gx[1] = (t.y[2] - t.y[0]) / a2;
  gy[1] = (t.x[0] - t.x[2]) / a2;
  gx[2] = (t.y[0] - t.y[1]) / a2;
  gy[2] = (t.x[1] - t.x[0]) / a2;
}

__device__ __forceinline__ const double* mat_of(const qds_dd_mesh& m, int64_t node) {
  return m.mat + (int64_t)m.mat_idx[node] * kMatStride;
}

// ---- node kernels -------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
carriers_kernel(qds_dd_mesh m, int fermi_dirac, double kT, double n_min, const double* __restrict__ phi,
                const double* __restrict__ phi_n, const double* __restrict__ phi_p, double* __restrict__ n,
                double* __restrict__ p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double* mt = mat_of(m, i);
    const double eta_n = (phi_n[i] - phi[i]) / kT;
    const double eta_p = (phi[i] - phi_p[i] - mt[2]) / kT;
    double ni, pi;
    if (fermi_dirac) {
      ni = mt[0] * fd_half(eta_n);
      pi = mt[1] * fd_half(eta_p);
    } else {
      ni = mt[0] * exp(eta_n);
      pi = mt[1] * exp(eta_p);
    }
    n[i] = std_max(ni, n_min);
    p[i] = std_max(pi, n_min);
  }
}

__global__ void __launch_bounds__(kBlock)
contacts_kernel(qds_dd_mesh m, double V_p, double V_n, double kT, double n_contact, double* __restrict__ phi,
                double* __restrict__ phi_n, double* __restrict__ phi_p, double* __restrict__ n,
                double* __restrict__ p, double* __restrict__ fn, double* __restrict__ fp) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = m.contact[i];
    if (c == 0) continue;
    const double* mt = mat_of(m, i);
    const double nin = mt[6];
    if (c == 1) {   // x < tol: p-type contact, N_a
      phi[i] = V_p;
      phi_n[i] = V_p + kT * log(nin * nin / n_contact / mt[0]);
      phi_p[i] = V_p - mt[2];
      n[i] = nin * nin / n_contact;
      p[i] = n_contact;
    } else {        // x > Lx - tol: n-type contact, N_d
      phi[i] = V_n;
      phi_n[i] = V_n;
      phi_p[i] = V_n - mt[2] - kT * log(n_contact / mt[1]);
      n[i] = n_contact;
      p[i] = nin * nin / n_contact;
    }
    // the replaced continuity rows carry the new contact quasi-Fermi levels
    if (fn) fn[i] = phi_n[i];
    if (fp) fp[i] = phi_p[i];
  }
}

// initial carriers from the doping (:216-260): majority = |N|, minority = n_i^2 / |N|,
// n = p = n_i where N = 0, floor n_min; phi_n = kT ln(n / N_c), phi_p = -kT ln(p / N_v) - E_g
__global__ void __launch_bounds__(kBlock)
initial_kernel(qds_dd_mesh m, double kT, double n_min, const double* __restrict__ doping, double* __restrict__ n,
               double* __restrict__ p, double* __restrict__ phi_n, double* __restrict__ phi_p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double* mt = mat_of(m, i);
    const double N = doping[i], nin = mt[6];
    double ni, pi;
    if (N > 0) {
      ni = N;
      pi = nin * nin / N;
    } else if (N < 0) {
      pi = -N;
      ni = nin * nin / pi;
    } else {
      ni = nin;
      pi = nin;
    }
    ni = std_max(ni, n_min);
    pi = std_max(pi, n_min);
    n[i] = ni;
    p[i] = pi;
    phi_n[i] = kT * log(ni / mt[0]);
    phi_p[i] = -kT * log(pi / mt[1]) - mt[2];
  }
}

__global__ void __launch_bounds__(kBlock)
charge_kernel(int64_t nn, double scale, const double* __restrict__ n, const double* __restrict__ p,
              const double* __restrict__ doping, double* __restrict__ rho) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x)
    rho[i] = scale * (p[i] - n[i] + doping[i]);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  return s;
}

// mode 0: phi = phi_old + d (phi - phi_old)
// mode 1: non-finite phi -> phi_old; partial sums of phi^2, (phi - phi_old)^2, #non-finite
__global__ void __launch_bounds__(kBlock)
potential_kernel(int64_t nn, int mode, double d, const double* __restrict__ old, double* __restrict__ phi,
                 double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
    const double o = old[i];
    double v = phi[i];
    if (mode == 0) {
      phi[i] = o + d * (v - o);
    } else {
      if (!isfinite(v)) {
        v = o;
        phi[i] = o;
        s2 += 1.0;
      }
      s0 += v * v;
      s1 += (v - o) * (v - o);
    }
  }
  if (mode == 0) return;
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = s0;
    partial[3 * blockIdx.x + 1] = s1;
    partial[3 * blockIdx.x + 2] = s2;
  }
}

__global__ void __launch_bounds__(kBlock)
potential_final(int nblk, const double* __restrict__ partial, double* __restrict__ sums) {
  __shared__ double sh[kBlock / 64];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
    s0 += partial[3 * k];
    s1 += partial[3 * k + 1];
    s2 += partial[3 * k + 2];
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    sums[0] = s0;
    sums[1] = s1;
    sums[2] = s2;
  }
}

// ---- node averages over the node's elements ---------------------------------------------

// E[i] = -(mean over the elements of i of grad phi_e)
__global__ void __launch_bounds__(kBlock)
field_kernel(qds_dd_mesh m, const double* __restrict__ phi, double* __restrict__ E) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = m.ne_ptr[i], en = m.ne_ptr[i + 1];
    double ex = 0.0, ey = 0.0;
    for (int32_t k = b; k < en; ++k) {
      const Tri t = load_tri(m, m.ne_elem[k]);
      double gx[3], gy[3];
      tri_grad_area(t, gx, gy);
      const double u0 = phi[t.a[0]], u1 = phi[t.a[1]], u2 = phi[t.a[2]];
      ex += -(u0 * gx[0] + u1 * gx[1] + u2 * gx[2]);
      ey += -(u0 * gy[0] + u1 * gy[1] + u2 * gy[2]);
    }
    if (en > b) {
      ex /= (double)(en - b);
      ey /= (double)(en - b);
    }
    E[2 * i] = ex;
    E[2 * i + 1] = ey;
  }
}

// Jn[i] = mean over the elements of q mu_n,c n_c grad phi_n (centroid means of the
// nodal mu and n), Jp likewise with mu_p, p, phi_p
__global__ void __launch_bounds__(kBlock)
currents_kernel(qds_dd_mesh m, double q, const double* __restrict__ phi_n, const double* __restrict__ phi_p,
                const double* __restrict__ n, const double* __restrict__ p, const double* __restrict__ mu_n,
                const double* __restrict__ mu_p, double* __restrict__ Jn, double* __restrict__ Jp) {
  const double third = 1.0 / 3.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = m.ne_ptr[i], en = m.ne_ptr[i + 1];
    double jnx = 0.0, jny = 0.0, jpx = 0.0, jpy = 0.0;
    for (int32_t k = b; k < en; ++k) {
      const Tri t = load_tri(m, m.ne_elem[k]);
      double gx[3], gy[3];
      tri_grad_area(t, gx, gy);
      double nc = 0.0, pc = 0.0, mnc = 0.0, mpc = 0.0, dnx = 0.0, dny = 0.0, dpx = 0.0, dpy = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int v = t.a[a];
        nc += third * n[v];
        pc += third * p[v];
        mnc += third * mu_n[v];
        mpc += third * mu_p[v];
        dnx += phi_n[v] * gx[a];
        dny += phi_n[v] * gy[a];
        dpx += phi_p[v] * gx[a];
        dpy += phi_p[v] * gy[a];
      }
      const double wn = q * mnc * nc, wp = q * mpc * pc;
      jnx += wn * dnx;
      jny += wn * dny;
      jpx += wp * dpx;
      jpy += wp * dpy;
    }
    if (en > b) {
      const double c = (double)(en - b);
      jnx /= c;
      jny /= c;
      jpx /= c;
      jpy /= c;
    }
    Jn[2 * i] = jnx;
    Jn[2 * i + 1] = jny;
    Jp[2 * i] = jpx;
    Jp[2 * i + 1] = jpy;
  }
}

// ---- row kernels ------------------------------------------------------------------------

__device__ __forceinline__ int32_t find_slot(const int32_t* __restrict__ idx, int32_t b, int32_t en, int32_t col) {
  for (int32_t k = b; k < en; ++k)
    if (idx[k] == col) return k;
  return -1;
}

// SRH rate of :72-82 (n_i = 1.5e6 cm^-3, tau_n = tau_p = 1e-9 s)
__device__ __forceinline__ double srh(double n, double p) {
  const double ni = 1.5e6, tn = 1e-9, tp = 1e-9;
  return (n * p - ni * ni) / (tp * (n + ni) + tn * (p + ni));
}

// Continuity rows of electrons (Kn, fn) and holes (Kp, fp):
//   K_e(i, j) = D_c grad N_i . grad N_j A +/- mu_c E_c . grad N_j A / 3,  D_c = mu_c kT
//   f_e(i)    = G_c A / 3
// with centroid means of E, mu and G (:896-990); Dirichlet rows carry phi_n / phi_p.
__global__ void __launch_bounds__(kBlock)
continuity_rows(qds_dd_mesh m, int ldlt_lower, double kT, const double* __restrict__ E,
                const double* __restrict__ mu_n, const double* __restrict__ mu_p, const double* __restrict__ G,
                const double* __restrict__ n, const double* __restrict__ p, const double* __restrict__ phi_n,
                const double* __restrict__ phi_p, double* __restrict__ Kn, double* __restrict__ Kp,
                double* __restrict__ fn, double* __restrict__ fp) {
  const double third = 1.0 / 3.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r0 = m.indptr[i], r1 = m.indptr[i + 1];
    for (int32_t k = r0; k < r1; ++k) {
      Kn[k] = 0.0;
      Kp[k] = 0.0;
    }
    const bool bci = m.contact[i] != 0;
    double load = 0.0;
    const int32_t b = m.ne_ptr[i], en = m.ne_ptr[i + 1];
    for (int32_t s = b; s < en; ++s) {
      const Tri t = load_tri(m, m.ne_elem[s]);
      double gx[3], gy[3];
      const double area = tri_grad_jac(t, gx, gy);
      double ecx = 0.0, ecy = 0.0, mnc = 0.0, mpc = 0.0, gc = 0.0;
      int li = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int v = t.a[a];
        if (v == i) li = a;
        ecx += third * E[2 * v];
        ecy += third * E[2 * v + 1];
        mnc += third * mu_n[v];
        mpc += third * mu_p[v];
        gc += third * (G ? G[v] : srh(n[v], p[v]));
      }
      const double dn = mnc * kT, dp = mpc * kT;
      load += gc * area / 3.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int32_t j = t.a[a];
        int r, c;          // element entry (r, c) that lands in slot (i, j)
        if (ldlt_lower && j > i) {
          if (m.contact[j] != 0) continue;   // row j was replaced: A(j, i) = 0
          r = a;
          c = li;
        } else {
          if (bci) continue;                 // row i is replaced below
          r = li;
          c = a;
        }
        const double gg = (gx[r] * gx[c] + gy[r] * gy[c]) * area;
        const double drift = (ecx * gx[c] + ecy * gy[c]) * area / 3.0;
        const int32_t k = find_slot(m.indices, r0, r1, j);
        if (k < 0) continue;   // not in the pattern: the caller's pattern is not this mesh's
        Kn[k] += dn * gg + mnc * drift;
        Kp[k] += dp * gg - mpc * drift;
      }
    }
    if (bci) {
      const int32_t k = find_slot(m.indices, r0, r1, (int32_t)i);
      if (k >= 0) {
        Kn[k] = 1.0;
        Kp[k] = 1.0;
      }
      fn[i] = phi_n[i];
      fp[i] = phi_p[i];
    } else {
      fn[i] = load;
      fp[i] = load;
    }
  }
}

// Poisson rows: K_e(i, j) = eps_e grad N_i . grad N_j A (eps_e = eps_r eps0 of the
// centroid's material), f_i -= sum_q rho_q lambda_i(q) (1/6) A over the three points
// (1/6, 1/6, 2/3) and permutations, rho = q (p - n + N) (:296-418); Dirichlet rows
// V_p (contact 1) / V_n (contact 2).
__global__ void __launch_bounds__(kBlock)
poisson_rows(qds_dd_mesh m, int ldlt_lower, double eps0, double q, const double* __restrict__ n,
             const double* __restrict__ p, const double* __restrict__ doping, double V_p, double V_n,
             double* __restrict__ K, double* __restrict__ f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.n_nodes;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r0 = m.indptr[i], r1 = m.indptr[i + 1];
    for (int32_t k = r0; k < r1; ++k) K[k] = 0.0;
    const int ci = m.contact[i];
    double load = 0.0;
    const int32_t b = m.ne_ptr[i], en = m.ne_ptr[i + 1];
    for (int32_t s = b; s < en; ++s) {
      const int64_t e = m.ne_elem[s];
      const Tri t = load_tri(m, e);
      double gx[3], gy[3];
      const double area = tri_grad_jac(t, gx, gy);
      const double eps = m.mat[(int64_t)m.emat_idx[e] * kMatStride + 5] * eps0;
      int li = 0;
      double rho[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int v = t.a[a];
        if (v == i) li = a;
        rho[a] = q * (p[v] - n[v] + doping[v]);
      }
      // three-point rule with weights 1/6 on the physical area
#pragma unroll
      for (int qp = 0; qp < 3; ++qp) {
        const double l0 = qp == 2 ? 2.0 / 3.0 : 1.0 / 6.0;
        const double l1 = qp == 1 ? 2.0 / 3.0 : 1.0 / 6.0;
        const double l2 = qp == 0 ? 2.0 / 3.0 : 1.0 / 6.0;
        const double rq = rho[0] * l0 + rho[1] * l1 + rho[2] * l2;
        const double li_v = li == 0 ? l0 : (li == 1 ? l1 : l2);
        load -= rq * li_v * (1.0 / 6.0) * area;
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int32_t j = t.a[a];
        int r, c;
        if (ldlt_lower && j > i) {
          if (m.contact[j] != 0) continue;
          r = a;
          c = li;
        } else {
          if (ci) continue;
          r = li;
          c = a;
        }
        const int32_t k = find_slot(m.indices, r0, r1, j);
        if (k >= 0) K[k] += eps * (gx[r] * gx[c] + gy[r] * gy[c]) * area;
      }
    }
    if (ci) {
      const int32_t k = find_slot(m.indices, r0, r1, (int32_t)i);
      if (k >= 0) K[k] = 1.0;
      f[i] = ci == 1 ? V_p : V_n;
    } else {
      f[i] = load;
    }
  }
}

inline dim3 node_grid(int64_t n) { return dim3(qds_blocks(n, kBlock, 1 << 16)); }

}  // namespace

#define DD_CHECK_MESH(m, name)                                                                    \
  QDS_CHECK_ARG((m) && (m)->n_nodes > 0 && (m)->conn && (m)->x && (m)->y && (m)->ne_ptr &&         \
                    (m)->ne_elem && (m)->mat_idx && (m)->mat && (m)->contact,                      \
                name ": incomplete mesh description")

QDS_API int qds_dd_carriers(const qds_dd_mesh* m, int32_t fermi_dirac, double kT, double n_min, const double* phi,
                            const double* phi_n, const double* phi_p, double* n, double* p, void* stream) {
  DD_CHECK_MESH(m, "qds_dd_carriers");
  QDS_CHECK_ARG(phi && phi_n && phi_p && n && p, "qds_dd_carriers: null pointer");
  hipLaunchKernelGGL(carriers_kernel, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, fermi_dirac,
                     kT, n_min, phi, phi_n, phi_p, n, p);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_contacts(const qds_dd_mesh* m, double V_p, double V_n, double kT, double n_contact, double* phi,
                            double* phi_n, double* phi_p, double* n, double* p, double* fn, double* fp,
                            void* stream) {
  DD_CHECK_MESH(m, "qds_dd_contacts");
  QDS_CHECK_ARG(phi && phi_n && phi_p && n && p, "qds_dd_contacts: null pointer");
  hipLaunchKernelGGL(contacts_kernel, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, V_p, V_n, kT,
                     n_contact, phi, phi_n, phi_p, n, p, fn, fp);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_initial(const qds_dd_mesh* m, double kT, double n_min, const double* doping, double* n, double* p,
                           double* phi_n, double* phi_p, void* stream) {
  DD_CHECK_MESH(m, "qds_dd_initial");
  QDS_CHECK_ARG(doping && n && p && phi_n && phi_p, "qds_dd_initial: null pointer");
  hipLaunchKernelGGL(initial_kernel, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, kT, n_min,
                     doping, n, p, phi_n, phi_p);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_charge(int64_t n_nodes, double scale, const double* n, const double* p, const double* doping,
                          double* rho, void* stream) {
  QDS_CHECK_ARG(n && p && doping && rho && n_nodes >= 0, "qds_dd_charge: bad argument");
  if (n_nodes == 0) return QDS_OK;
  hipLaunchKernelGGL(charge_kernel, node_grid(n_nodes), dim3(kBlock), 0, qds_stream(stream), n_nodes, scale, n, p,
                     doping, rho);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int64_t qds_dd_workspace(void) { return (3 * kRedBlocks + 8) * (int64_t)sizeof(double); }

QDS_API int qds_dd_potential_update(int64_t n_nodes, int32_t mode, double damping, const double* phi_old, double* phi,
                                    double* sums, void* workspace, int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(phi_old && phi && n_nodes > 0 && (mode == 0 || mode == 1), "qds_dd_potential_update: bad argument");
  hipStream_t st = qds_stream(stream);
  if (mode == 0) {
    hipLaunchKernelGGL(potential_kernel, node_grid(n_nodes), dim3(kBlock), 0, st, n_nodes, 0, damping, phi_old, phi,
                       (double*)nullptr);
    QDS_LAUNCH_CHECK();
    return QDS_OK;
  }
  QDS_CHECK_ARG(sums && workspace && ws_bytes >= qds_dd_workspace(), "qds_dd_potential_update: workspace");
  double* partial = static_cast<double*>(workspace);
  const unsigned nb = qds_blocks(n_nodes, kBlock, kRedBlocks);
  hipLaunchKernelGGL(potential_kernel, dim3(nb), dim3(kBlock), 0, st, n_nodes, 1, damping, phi_old, phi, partial);
  hipLaunchKernelGGL(potential_final, dim3(1), dim3(kBlock), 0, st, (int)nb, partial, sums);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_field(const qds_dd_mesh* m, const double* phi, double* E, void* stream) {
  DD_CHECK_MESH(m, "qds_dd_field");
  QDS_CHECK_ARG(phi && E, "qds_dd_field: null pointer");
  hipLaunchKernelGGL(field_kernel, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, phi, E);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_currents(const qds_dd_mesh* m, double q, const double* phi_n, const double* phi_p, const double* n,
                            const double* p, const double* mu_n, const double* mu_p, double* Jn, double* Jp,
                            void* stream) {
  DD_CHECK_MESH(m, "qds_dd_currents");
  QDS_CHECK_ARG(phi_n && phi_p && n && p && mu_n && mu_p && Jn && Jp, "qds_dd_currents: null pointer");
  hipLaunchKernelGGL(currents_kernel, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, q, phi_n, phi_p,
                     n, p, mu_n, mu_p, Jn, Jp);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_continuity(const qds_dd_mesh* m, int32_t ldlt_lower, double kT, const double* E, const double* mu_n,
                              const double* mu_p, const double* G, const double* n, const double* p,
                              const double* phi_n, const double* phi_p, double* Kn, double* Kp, double* fn,
                              double* fp, void* stream) {
  DD_CHECK_MESH(m, "qds_dd_continuity");
  QDS_CHECK_ARG(m->indptr && m->indices, "qds_dd_continuity: no CSR pattern");
  QDS_CHECK_ARG(E && mu_n && mu_p && (G || (n && p)) && phi_n && phi_p && Kn && Kp && fn && fp,
                "qds_dd_continuity: null pointer");
  hipLaunchKernelGGL(continuity_rows, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, ldlt_lower, kT,
                     E, mu_n, mu_p, G, n, p, phi_n, phi_p, Kn, Kp, fn, fp);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int qds_dd_poisson_system(const qds_dd_mesh* m, int32_t ldlt_lower, double eps0, double q, const double* n,
                                  const double* p, const double* doping, double V_p, double V_n, double* K, double* f,
                                  void* stream) {
  DD_CHECK_MESH(m, "qds_dd_poisson_system");
  QDS_CHECK_ARG(m->indptr && m->indices && m->emat_idx, "qds_dd_poisson_system: no CSR pattern / element materials");
  QDS_CHECK_ARG(n && p && doping && K && f, "qds_dd_poisson_system: null pointer");
  hipLaunchKernelGGL(poisson_rows, node_grid(m->n_nodes), dim3(kBlock), 0, qds_stream(stream), *m, ldlt_lower, eps0, q,
                     n, p, doping, V_p, V_n, K, f);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
