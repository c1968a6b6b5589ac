// Block product Y_a = A_a X, Y_b = A_b X for operators in the lattice (stencil)
// form: the free-DOF Hamiltonian and mass matrix of a structured order-p mesh,
// whose rows (interior lattice points) couple only to points at most p lattice
// steps away, with one offset list per class (I mod p, J mod p).
//
// Reference operation: the sparse products scipy performs inside ARPACK / lobpcg
// (qdsim_cython/solvers/eigen_solver.pyx:225-235, 254-266) on the CSR matrices of
// fem_solver.pyx:142-145.  The CSR kernel (qds_spmm) issues, per nonzero and row
// group, an index load, two value loads and a dependent 16-byte X gather; on
// gfx950 its address unit is busy ~95% of the kernel time (rocprofv3 TA_TA_BUSY)
// at ~45 vector memory instructions per wave.  Here:
//   * a workgroup owns one block of the lattice: one cell row (p lattice rows)
//     times `tic` cells (m value tiles of the stencil layout).  The X rows of the
//     block and a p-point halo, and the block's values of both matrices (one
//     contiguous run per matrix in HBM), are staged in LDS with 16-byte loads
//     that are all issued before the first LDS write; each X row is fetched ~3.6
//     times per product instead of once per nonzero, and the re-fetches hit L2;
//   * there are no column indices: entry k of a class-c row reads the LDS point at
//     a per-class offset, a wave-uniform scalar;
//   * each wave task holds rows of one class only, so the trip count is uniform;
//   * LDS layouts make every wave read contiguous: an LDS row stores the points
//     of each residue I mod p together, so the rows of a task (stride p on the
//     lattice) are adjacent, and the value layout puts the rows of a class next to
//     each other in groups of four entries.
// Entries are summed in ascending column order with fused multiply-adds, the order
// of the CSR kernel; the padding and boundary entries add exact zeros, so both
// kernels give bitwise equal products.
//
// Optional residual epilogue (qds_lattice_spmm_residual): with Y_a = H X and
// Y_b = M X in registers, each row also stores R = (Y_a - lambda Y_b) diag(colscale)
// and adds ||r_j||^2 and ||(M X)_j||^2 of its row to per-lane sums; the workgroup
// reduces them in a fixed order into one partial row per lattice block, and two
// small kernels sum the partials (fixed order: deterministic).  The block iteration
// otherwise reads H X and M X back from HBM for exactly this.
#include "qds_common.h"
#include "../../include/qdsim_hip.h"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTasks = 24;            // wave tasks per wave and block
constexpr int kBatch = 12;               // 16-byte loads in flight per thread while staging
constexpr int kVB = 4;                   // value chunks per thread in the common staging path
constexpr size_t kLdsBudget = 52 * 1024; // LDS per workgroup: 3 workgroups per CU
constexpr size_t kLdsMax = 64 * 1024;

// residual epilogue arguments (R == nullptr: plain product)
struct ResEpilogue {
  const double* lam;       // b eigenvalue estimates
  const double* colscale;  // b column scales or nullptr (1)
  void* R;                 // residual block, element size r_es
  int64_t ldr;
  int32_t r_es;
  int32_t r_vec;           // R and ldr allow one vector store per lane and row
  double* part;            // per lattice block: b sums of r^2, then b sums of (M x)^2
};

struct LatticeArgs {
  int64_t nx, ny, cells_x;
  int64_t n_blocks;     // cells_y * bx_per_row
  int64_t tiles_x;      // value tiles per cell row
  int32_t bx_per_row;   // lattice blocks per cell row
  int32_t tile;         // cells per value tile
  int32_t m;            // value tiles per lattice block
  int32_t tic;          // cells per lattice block (m * tile)
  int32_t b, lpr, rpw, nch;  // block width, lanes per row, rows per wave, chunks per class row
  int32_t lw;           // LDS row width in points (p * tic + 2p)
  int32_t ps;           // LDS point stride in elements (= b)
  int32_t row_chunks, xchunks;  // 16-byte chunks per LDS row / of the X block
  int32_t vlds;         // first 16-byte chunk of the value runs in LDS
  int32_t tile_chunks;  // 16-byte chunks of one value tile
  int32_t vchunks;      // 16-byte chunks of one matrix's value run (m tiles)
  int64_t tile_slots;   // values per tile
  int32_t len[QDS_LATTICE_MAX_CLASS];
  int32_t base[QDS_LATTICE_MAX_CLASS];  // class offset within a value tile (elements)
  int16_t off[QDS_LATTICE_MAX_CLASS][QDS_LATTICE_MAX_ENTRY];  // LDS point offset of each entry
  int8_t tasks[kWaves][kMaxTasks];  // task = class * nch + chunk, -1 terminated
  // batch of problems on one stencil (blockIdx.y): element strides between the
  // problems' value arrays, X and Y blocks (0: shared)
  int64_t sva, svb, sx, sya, syb;
  int32_t batch;
};

template <typename T, int N>
struct Vec;
template <> struct Vec<float, 4> { using type = float4; };
template <> struct Vec<float, 2> { using type = float2; };
template <> struct Vec<float, 1> { using type = float; };
template <> struct Vec<double, 2> { using type = double2; };
template <> struct Vec<double, 1> { using type = double; };

template <typename T, int N>
__device__ __forceinline__ void ld(const T* p, T (&v)[N]) {
  using V = typename Vec<T, N>::type;
  const V t = *reinterpret_cast<const V*>(p);
  const T* tp = reinterpret_cast<const T*>(&t);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = tp[i];
}

template <typename T, int N>
__device__ __forceinline__ void st(T* p, const T (&v)[N]) {
  using V = typename Vec<T, N>::type;
  V t;
  T* tp = reinterpret_cast<T*>(&t);
#pragma unroll
  for (int i = 0; i < N; ++i) tp[i] = v[i];
  *reinterpret_cast<V*>(p) = t;
}

// N values of type S stored as type D: vector stores of at most 16 bytes (VEC
// aligned p) or one store per value
template <typename D, typename S, int N>
__device__ __forceinline__ void st_cvt(D* p, const S (&v)[N], bool vec) {
  D t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = (D)v[i];
  if (!vec) {
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = t[i];
  } else if constexpr (N * sizeof(D) <= 16) {
    st<D, N>(p, t);
  } else {
    constexpr int H = N / 2;
    D a[H], b[H];
#pragma unroll
    for (int i = 0; i < H; ++i) { a[i] = t[i]; b[i] = t[H + i]; }
    st<D, H>(p, a);
    st<D, H>(p + H, b);
  }
}

// four consecutive values (one 16-byte LDS read for float, two for double)
template <typename T>
__device__ __forceinline__ void ld4(const T* p, T (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    ld<T, 4>(p, v);
  } else {
    T a[2], b[2];
    ld<T, 2>(p, a);
    ld<T, 2>(p + 2, b);
    v[0] = a[0]; v[1] = a[1]; v[2] = b[0]; v[3] = b[1];
  }
}

// Entries of one task row, LEN known at compile time (fully unrolled; summation
// order as in the generic loop: entry by entry).  The task's LDS offsets are read
// into scalar registers before any LDS read: scalar loads and LDS reads share one
// wait counter, so an offset fetched inside the loop drained every LDS read in
// flight.  Entries then go in groups of 8 whose value and X reads are all issued
// before their FMAs (the compiler otherwise kept two X reads in flight and exposed
// the LDS latency at every entry).
template <typename T, int VEC, bool FUSED, int LEN>
__device__ __forceinline__ void task_entries(const int16_t* __restrict__ off, const T* xl, int ps,
                                             const T* lva, const T* lvb, int64_t vrow, int vstep,
                                             T (&acc_a)[VEC], T (&acc_b)[VEC]) {
  constexpr int G = 8;
  int o[LEN];
#pragma unroll
  for (int k = 0; k < LEN; ++k) o[k] = __builtin_amdgcn_readfirstlane((int)off[k]);
#pragma unroll
  for (int k0 = 0; k0 < LEN; k0 += G) {
    const int g = LEN - k0 < G ? LEN - k0 : G;   // a multiple of 4
    T av[G], bv[G];
    T x[G][VEC];
#pragma unroll
    for (int q = 0; q < G / 4; ++q) {
      if (q * 4 >= g) break;
      const int64_t vk = vrow + (int64_t)((k0 >> 2) + q) * vstep;
      T t4[4];
      ld4<T>(lva + vk, t4);
#pragma unroll
      for (int e = 0; e < 4; ++e) av[4 * q + e] = t4[e];
      if (FUSED) {
        ld4<T>(lvb + vk, t4);
#pragma unroll
        for (int e = 0; e < 4; ++e) bv[4 * q + e] = t4[e];
      }
    }
#pragma unroll
    for (int e = 0; e < G; ++e) {
      if (e >= g) break;
      ld<T, VEC>(xl + o[k0 + e] * ps, x[e]);
    }
#pragma unroll
    for (int e = 0; e < G; ++e) {
      if (e >= g) break;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        acc_a[v] = fma(av[e], x[e][v], acc_a[v]);
        if (FUSED) acc_b[v] = fma(bv[e], x[e][v], acc_b[v]);
      }
    }
  }
}

// Workgroups go round-robin to the 8 XCDs; block g % 8 * (G / 8) + g / 8 makes every
// XCD sweep a contiguous range of lattice blocks in row-major order, so the halo
// rows a block shares with the cell row above were fetched into the same L2
// shortly before.
__device__ __forceinline__ int64_t xcd_block(int64_t g, int64_t G) { return (g % 8) * (G / 8) + g / 8; }

// VLDS: the block's value runs are staged in LDS with the X rows (true) or read by
// the compute phase straight from HBM (false: every value is used once, by the
// lanes of one row, so staging buys nothing but LDS footprint and a barrier's
// worth of latency; without it more workgroups fit on a CU).
// RES: the residual epilogue is compiled in (its registers would otherwise cost the
// plain product occupancy and scheduling freedom).
template <typename T, int VEC, int P, bool FUSED, bool VLDS, bool RES>
__global__ void __launch_bounds__(kThreads)
lattice_spmm_kernel(const LatticeArgs a, const T* __restrict__ va, const T* __restrict__ vb,
                    const T* __restrict__ X, int64_t ldx, int chunked, T* __restrict__ Ya,
                    int64_t ldya, T* __restrict__ Yb, int64_t ldyb, const ResEpilogue res) {
  constexpr int CE = 16 / sizeof(T);  // elements per 16-byte chunk
  constexpr int NV = FUSED ? 2 : 1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  float4* lds4 = reinterpret_cast<float4*>(lds_raw);
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  if (blk >= a.n_blocks) return;  // uniform over the workgroup
  {
    const int64_t set = blockIdx.y;
    va += set * a.sva;
    vb += set * a.svb;
    X += set * a.sx;
    Ya += set * a.sya;
    if (FUSED) Yb += set * a.syb;
  }
  const int64_t cr = blk / a.bx_per_row;
  const int64_t t0 = (blk - cr * a.bx_per_row) * a.m;  // first value tile of the block
  const int64_t cc0 = t0 * a.tile;                    // first cell of the block
  const int64_t J0 = P * cr - P, I0 = P * cc0 - P;    // full-lattice origin of the LDS block
  const int lw = a.lw, ps = a.ps;
  const int wr = lw / P;  // points per residue class I mod P in an LDS row
  const int ncell = (int)min((int64_t)a.tic, a.cells_x - cc0);  // cells of this block
  // value runs: m tiles of each matrix, contiguous in HBM; tiles past the row end
  // are not loaded (their cells have no rows)
  const int vvalid = (int)min((int64_t)a.m, a.tiles_x - t0) * a.tile_chunks;
  const float4* vsrc_a = reinterpret_cast<const float4*>(va + (cr * a.tiles_x + t0) * a.tile_slots);
  const float4* vsrc_b = reinterpret_cast<const float4*>((FUSED ? vb : va) + (cr * a.tiles_x + t0) * a.tile_slots);

  // ---- stage the block in LDS: X rows of the block and its p-point halo (zeros
  // outside the interior), then the value runs.  Every thread issues all of its
  // 16-byte loads before the first LDS write, so one memory latency covers the
  // whole block.
  if (chunked && a.row_chunks <= kThreads && (!VLDS || NV * a.vchunks <= kThreads * kVB)) {
    // common case: each thread owns one 16-byte column slot of the X rows
    // (decoded once, then one load per row) and at most kVB value chunks
    const int cpp = ps / CE;
    const int e = threadIdx.x;
    const bool xact = e < a.row_chunks;
    const int pt = xact ? e / cpp : 0;
    const int part = e - pt * cpp;
    const int64_t I = I0 + pt;
    const bool iin = xact && I >= 1 && I <= a.nx;
    const T* xsrc = X + (I - 1) * ldx + part * CE;
    const int xdst = ((pt % P) * wr + pt / P) * cpp + part;
    float4 xv[3 * P];
#pragma unroll
    for (int lr = 0; lr < 3 * P; ++lr) {
      const int64_t J = J0 + lr;
      xv[lr] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (iin && J >= 1 && J <= a.ny) xv[lr] = *reinterpret_cast<const float4*>(xsrc + (J - 1) * a.nx * ldx);
    }
    float4 vv[kVB];
    if constexpr (VLDS) {
#pragma unroll
      for (int i = 0; i < kVB; ++i) {
        const int f2 = i * kThreads + (int)threadIdx.x;
        vv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        const bool second = f2 >= a.vchunks;
        const int f = second ? f2 - a.vchunks : f2;
        if (f2 < NV * a.vchunks && f < vvalid) vv[i] = second ? vsrc_b[f] : vsrc_a[f];
      }
    }
    if (xact) {
#pragma unroll
      for (int lr = 0; lr < 3 * P; ++lr) lds4[lr * a.row_chunks + xdst] = xv[lr];
    }
    if constexpr (VLDS) {
#pragma unroll
      for (int i = 0; i < kVB; ++i) {
        const int f2 = i * kThreads + (int)threadIdx.x;
        if (f2 < NV * a.vchunks) lds4[a.vlds + f2] = vv[i];
      }
    }
  } else {
    const int xchunks = chunked ? a.xchunks : 0;
    const int total = xchunks + (VLDS ? NV * a.vchunks : 0);
    for (int e0 = 0; e0 < total; e0 += kThreads * kBatch) {
      float4 v[kBatch];
      int dst[kBatch];
#pragma unroll
      for (int i = 0; i < kBatch; ++i) {
        const int e = e0 + i * kThreads + (int)threadIdx.x;
        v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        dst[i] = -1;
        if (e >= total) continue;
        if (e < xchunks) {
          const int cpp = ps / CE;
          const int lr = e / a.row_chunks;
          const int rem = e - lr * a.row_chunks;
          const int pt = rem / cpp;
          const int part = rem - pt * cpp;
          const int64_t J = J0 + lr, I = I0 + pt;
          if (J >= 1 && J <= a.ny && I >= 1 && I <= a.nx)
            v[i] = *reinterpret_cast<const float4*>(X + ((J - 1) * a.nx + (I - 1)) * ldx + part * CE);
          dst[i] = (lr * lw + (pt % P) * wr + pt / P) * cpp + part;
        } else {
          const int f2 = e - xchunks;
          const bool second = f2 >= a.vchunks;
          const int f = second ? f2 - a.vchunks : f2;
          if (f < vvalid) v[i] = second ? vsrc_b[f] : vsrc_a[f];
          dst[i] = a.vlds + f2;
        }
      }
#pragma unroll
      for (int i = 0; i < kBatch; ++i)
        if (dst[i] >= 0) lds4[dst[i]] = v[i];
    }
    if (!chunked) {
      const int row_elems = lw * ps;
      for (int lr = 0; lr < 3 * P; ++lr) {
        const int64_t J = J0 + lr;
        const bool jin = J >= 1 && J <= a.ny;
        T* dstx = lds + (int64_t)lr * lw * ps;
        for (int e = threadIdx.x; e < row_elems; e += kThreads) {
          const int pt = e / ps;
          const int col = e - pt * ps;
          const int64_t I = I0 + pt;
          T x = T(0);
          if (jin && I >= 1 && I <= a.nx) x = X[((J - 1) * a.nx + (I - 1)) * ldx + col];
          dstx[((pt % P) * wr + pt / P) * ps + col] = x;
        }
      }
    }
  }
  __syncthreads();

  // ---- wave tasks: rpw rows of one class, lpr lanes (VEC columns each) per row,
  // operands from LDS only
  const T* lva = VLDS ? lds + (int64_t)a.vlds * CE : reinterpret_cast<const T*>(vsrc_a);
  const T* lvb = VLDS ? lva + (int64_t)a.vchunks * CE : reinterpret_cast<const T*>(vsrc_b);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int g = lane / a.lpr;
  const int li = lane - g * a.lpr;
  const bool lane_ok = g < a.rpw && li * VEC < a.b;
  const int c0 = lane_ok ? li * VEC : 0;
  const int gg = g < a.rpw ? g : 0;
  constexpr bool resid = FUSED && RES;
  T rlam[VEC], rcs[VEC];
  double rsum[VEC], msum[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    rlam[v] = resid ? (T)res.lam[c0 + v] : T(0);
    rcs[v] = resid && res.colscale ? (T)res.colscale[c0 + v] : T(1);
    rsum[v] = 0.0;
    msum[v] = 0.0;
  }
  for (int s = 0; s < kMaxTasks; ++s) {
    const int task = a.tasks[wave][s];
    if (task < 0) break;
    const int c = task / a.nch;
    const int h = task - c * a.nch;
    const int ci = c % P, cj = c / P;
    const int u = h * a.rpw + gg;
    const int64_t I = P * (cc0 + u) + ci;
    const int64_t J = P * cr + cj;
    const bool ok = lane_ok && u < ncell && I >= 1 && I <= a.nx && J >= 1 && J <= a.ny;
    // lanes without a row read a valid row of the block and store nothing
    const int uu = u < ncell ? u : ncell - 1;
    const int len = a.len[c];
    const int seg = uu / a.tile;
    const int64_t vrow = seg * a.tile_slots + a.base[c] + (uu - seg * a.tile) * 4;
    const int vstep = a.tile * 4;  // next group of four entries
    const T* xl = lds + (int64_t)((P + cj) * lw + ci * wr + 1 + uu) * ps + c0;
    T acc_a[VEC], acc_b[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) { acc_a[v] = T(0); acc_b[v] = T(0); }
    // the stencil lengths of P1-P3 meshes get fully unrolled bodies: their LDS
    // offsets are then loaded once per task (scalar loads) instead of inside a loop
    switch (len) {
      case 8: task_entries<T, VEC, FUSED, 8>(a.off[c], xl, ps, lva, lvb, vrow, vstep, acc_a, acc_b); break;
      case 12: task_entries<T, VEC, FUSED, 12>(a.off[c], xl, ps, lva, lvb, vrow, vstep, acc_a, acc_b); break;
      case 16: task_entries<T, VEC, FUSED, 16>(a.off[c], xl, ps, lva, lvb, vrow, vstep, acc_a, acc_b); break;
      case 20: task_entries<T, VEC, FUSED, 20>(a.off[c], xl, ps, lva, lvb, vrow, vstep, acc_a, acc_b); break;
      case 40: task_entries<T, VEC, FUSED, 40>(a.off[c], xl, ps, lva, lvb, vrow, vstep, acc_a, acc_b); break;
      default:
        for (int k = 0; k < len; k += 4) {
          T av[4], bv[4];
          const int64_t vk = vrow + (int64_t)(k >> 2) * vstep;
          ld4<T>(lva + vk, av);
          if (FUSED) ld4<T>(lvb + vk, bv);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            T x[VEC];
            ld<T, VEC>(xl + (int)a.off[c][k + q] * ps, x);
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
              acc_a[v] = fma(av[q], x[v], acc_a[v]);
              if (FUSED) acc_b[v] = fma(bv[q], x[v], acc_b[v]);
            }
          }
        }
    }
    if (ok) {
      const int64_t row = (J - 1) * a.nx + (I - 1);
      st<T, VEC>(Ya + row * ldya + c0, acc_a);
      if (FUSED) st<T, VEC>(Yb + row * ldyb + c0, acc_b);
      if (resid) {
        // the formula of the residual-norms kernel: v = hx - lambda mx, R = v cs
        T rv[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const T m = acc_b[v];
          rv[v] = acc_a[v] - rlam[v] * m;
          rsum[v] += (double)rv[v] * (double)rv[v];
          msum[v] += (double)m * (double)m;
        }
        T rs[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) rs[v] = rv[v] * rcs[v];
        if (res.r_es == 4)
          st_cvt<float, T, VEC>(reinterpret_cast<float*>(res.R) + row * res.ldr + c0, rs, res.r_vec != 0);
        else
          st_cvt<double, T, VEC>(reinterpret_cast<double*>(res.R) + row * res.ldr + c0, rs, res.r_vec != 0);
      }
    }
  }
  if (resid) {
    // every wave is done with the staged block: reuse LDS for the reduction
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds_raw);
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      red[threadIdx.x * 2 * VEC + v] = rsum[v];
      red[threadIdx.x * 2 * VEC + VEC + v] = msum[v];
    }
    __syncthreads();
    // the kWaves * rpw row lanes of each column: every thread sums one share
    // (fixed assignment), then one thread per column adds the shares in order
    const int nrow = kWaves * a.rpw;
    const int shares = min(kThreads / (2 * a.b), nrow);
    double share = 0.0;
    const int t2 = threadIdx.x % (2 * a.b), sh = threadIdx.x / (2 * a.b);
    if (sh < shares) {
      const int col = t2 < a.b ? t2 : t2 - a.b;
      const int half = t2 < a.b ? 0 : VEC;
      const int tli = col / VEC, tv = col - tli * VEC;
      for (int rr = sh; rr < nrow; rr += shares) {
        const int w = rr / a.rpw, q = rr - w * a.rpw;
        share += red[(w * 64 + q * a.lpr + tli) * 2 * VEC + half + tv];
      }
    }
    __syncthreads();
    red[threadIdx.x] = share;
    __syncthreads();
    if (threadIdx.x < 2 * a.b) {
      double s0 = 0.0;
      for (int q = 0; q < shares; ++q) s0 += red[q * 2 * a.b + threadIdx.x];
      res.part[blk * 2 * a.b + threadIdx.x] = s0;
    }
  }
}

// Sums of the per-block residual partials (rows of width w = 2b) in a fixed
// order: grid kResGroups blocks, block i sums rows i, i + kResGroups, ... per
// column group of threads; then one block sums the kResGroups rows in order.
constexpr int kResGroups = 256;
__global__ void __launch_bounds__(kThreads)
res_reduce_rows(int64_t nrows, int w, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[kThreads];
  const int c = threadIdx.x % w;
  const int g = threadIdx.x / w;
  const int groups = kThreads / w;
  double a = 0.0;
  if (g < groups) {
    // four independent sums so that the loads of a thread overlap
    const int64_t step = (int64_t)groups * gridDim.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t r = (int64_t)blockIdx.x + (int64_t)g * gridDim.x;
    for (; r + 3 * step < nrows; r += 4 * step) {
      a0 += part[r * w + c];
      a1 += part[(r + step) * w + c];
      a2 += part[(r + 2 * step) * w + c];
      a3 += part[(r + 3 * step) * w + c];
    }
    for (; r < nrows; r += step) a0 += part[r * w + c];
    a = (a0 + a1) + (a2 + a3);
  }
  red[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x < w) {
    double s = 0.0;
    for (int q = 0; q < groups; ++q) s += red[q * w + threadIdx.x];
    out[(int64_t)blockIdx.x * w + threadIdx.x] = s;
  }
}

bool aligned(const void* p, int bytes) { return (reinterpret_cast<uintptr_t>(p) % bytes) == 0; }

template <typename T, int VEC, int P, bool FUSED, bool VLDS, bool RES = false>
int launch(const LatticeArgs& a, size_t lds, const void* va, const void* vb, const void* X,
           int64_t ldx, int chunked, void* Ya, int64_t ldya, void* Yb, int64_t ldyb, const ResEpilogue& res,
           hipStream_t st) {
  const int64_t grid = (a.n_blocks + 7) / 8 * 8;
  hipLaunchKernelGGL((lattice_spmm_kernel<T, VEC, P, FUSED, VLDS, RES>), dim3((unsigned)grid, (unsigned)a.batch), dim3(kThreads), lds,
                     st, a, (const T*)va, (const T*)vb, (const T*)X, ldx, chunked, (T*)Ya, ldya,
                     (T*)Yb, ldyb, res);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

template <typename T, int VEC>
int dispatch(int p, bool fused, bool vlds, const LatticeArgs& a, size_t lds, const void* va, const void* vb,
             const void* X, int64_t ldx, int chunked, void* Ya, int64_t ldya, void* Yb, int64_t ldyb,
             const ResEpilogue& res, hipStream_t st) {
#define QDS_LAT_CASE(PP)                                                                                       \
  case PP:                                                                                                   \
    if (fused && res.R != nullptr)                                                                           \
      return vlds ? launch<T, VEC, PP, true, true, true>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st) \
                  : launch<T, VEC, PP, true, false, true>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st); \
    if (vlds)                                                                                                \
      return fused ? launch<T, VEC, PP, true, true>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st)   \
                   : launch<T, VEC, PP, false, true>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st); \
    return fused ? launch<T, VEC, PP, true, false>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st)    \
                 : launch<T, VEC, PP, false, false>(a, lds, va, vb, X, ldx, chunked, Ya, ldya, Yb, ldyb, res, st);
  switch (p) {
    QDS_LAT_CASE(1)
    QDS_LAT_CASE(2)
    QDS_LAT_CASE(3)
    default:
      QDS_CHECK_ARG(false, "qds_lattice_spmm: order %d not in 1..3", p);
  }
#undef QDS_LAT_CASE
  return QDS_EINVAL;
}

}  // namespace

namespace {
struct BatchStrides {
  int32_t batch;
  int64_t sva, svb, sx, sya, syb;
};

struct Tuning {
  bool vlds;        // values staged in LDS
  size_t budget;    // LDS bytes per workgroup the block is sized for
};
constexpr Tuning kDefaultTuning{true, kLdsBudget};

// res (optional): the residual epilogue; norms gets the 2b sums (r^2 then (M x)^2)
// and part1 holds kResGroups rows of 2b partial sums.
int lattice_spmm_impl(int32_t dtype, const qds_lattice_stencil* s, int32_t b, const void* vals_a,
                      const void* vals_b, const void* X, int64_t ldx, void* Y_a, int64_t ldy_a,
                      void* Y_b, int64_t ldy_b, const BatchStrides& bs, void* stream,
                      const Tuning& tune = kDefaultTuning, const ResEpilogue* res = nullptr,
                      double* part1 = nullptr, double* norms = nullptr) {
  QDS_CHECK_ARG(s && vals_a && X && Y_a, "qds_lattice_spmm: null pointer");
  QDS_CHECK_ARG(!vals_b || Y_b, "qds_lattice_spmm: vals_b without Y_b");
  QDS_CHECK_ARG(dtype == qds::F32 || dtype == qds::F64, "qds_lattice_spmm: dtype must be F32 or F64");
  QDS_CHECK_ARG(b >= 1 && ldx >= b && ldy_a >= b && (!vals_b || ldy_b >= b), "qds_lattice_spmm: bad block shape");
  const int p = s->order;
  QDS_CHECK_ARG(p >= 1 && p <= 3 && s->n_class == p * p && s->nx >= 1 && s->ny >= 1 &&
                    s->cells_x == s->nx / p + 1 && s->cells_y == s->ny / p + 1 && s->tile >= 1 &&
                    s->tiles_x == (s->cells_x + s->tile - 1) / s->tile,
                "qds_lattice_spmm: inconsistent stencil description");
  const bool fused = vals_b != nullptr;
  const int es = qds::dtype_size(dtype);
  int64_t sum_len = 0;
  for (int c = 0; c < s->n_class; ++c) {
    QDS_CHECK_ARG(s->len[c] % 4 == 0 && s->len[c] <= QDS_LATTICE_MAX_ENTRY && s->base[c] == s->tile * sum_len,
                  "qds_lattice_spmm: inconsistent stencil description");
    sum_len += s->len[c];
  }
  QDS_CHECK_ARG(s->tile_slots == s->tile * sum_len && s->n_slots == s->cells_y * s->tiles_x * s->tile_slots,
                "qds_lattice_spmm: inconsistent stencil description");
  const size_t nval = fused ? 2 : 1;
  const size_t tile_vbytes = (size_t)s->tile_slots * es;
  // blocks wider than one LDS block of a single value tile allows: column slices
  const size_t xrow1 = (size_t)3 * p * (size_t)(p * s->tile + 2 * p) * es;  // X bytes per column, one tile
  QDS_CHECK_ARG(nval * tile_vbytes + 4 * xrow1 <= kLdsMax, "qds_lattice_spmm: stencil too large for LDS");
  const int bmax = (int)((kLdsMax - nval * tile_vbytes) / xrow1) / 4 * 4;
  if (b > bmax) {
    QDS_CHECK_ARG(res == nullptr, "qds_lattice_spmm_residual: block width %d above the single-pass limit %d", b,
                  bmax);
    for (int j = 0; j < b; j += bmax) {
      const int w = std::min(bmax, b - j);
      const int rc = lattice_spmm_impl(dtype, s, w, vals_a, vals_b, (const char*)X + (size_t)j * es, ldx,
                                       (char*)Y_a + (size_t)j * es, ldy_a,
                                       Y_b ? (char*)Y_b + (size_t)j * es : nullptr, ldy_b, bs, stream, tune);
      if (rc != QDS_OK) return rc;
    }
    return QDS_OK;
  }
  // lanes per row: VEC columns each, the widest VEC dividing b whose stores stay
  // aligned
  auto vec_ok = [&](int vec) {
    const int bytes = vec * es;
    if (b % vec) return false;
    bool ok = aligned(Y_a, bytes) && (ldy_a * es) % bytes == 0;
    if (fused) ok = ok && aligned(Y_b, bytes) && (ldy_b * es) % bytes == 0;
    return ok;
  };
  int vec = 1;
  if (dtype == qds::F32) vec = vec_ok(4) ? 4 : vec_ok(2) ? 2 : 1;
  else vec = vec_ok(2) ? 2 : 1;
  QDS_CHECK_ARG(aligned(vals_a, 16) && (!fused || aligned(vals_b, 16)),
                "qds_lattice_spmm: value arrays must be 16-byte aligned");
  QDS_CHECK_ARG(bs.batch >= 1 && bs.batch <= 65535, "qds_lattice_spmm: batch %d out of range", bs.batch);
  if (bs.batch > 1)
    QDS_CHECK_ARG((bs.sva * es) % 16 == 0 && (bs.svb * es) % 16 == 0 && (bs.sx * es) % 16 == 0 &&
                      (bs.sya * es) % 16 == 0 && (bs.syb * es) % 16 == 0,
                  "qds_lattice_spmm: batch strides must keep 16-byte alignment");
  LatticeArgs a{};
  a.batch = bs.batch;
  a.sva = bs.sva;
  a.svb = bs.svb;
  a.sx = bs.sx;
  a.sya = bs.sya;
  a.syb = bs.syb;
  a.nx = s->nx;
  a.ny = s->ny;
  a.cells_x = s->cells_x;
  a.tiles_x = s->tiles_x;
  a.tile = s->tile;
  a.tile_slots = s->tile_slots;
  a.b = b;
  a.lpr = (b + vec - 1) / vec;
  QDS_CHECK_ARG(a.lpr <= 64, "qds_lattice_spmm: block width %d too large", b);
  a.rpw = 64 / a.lpr;
  a.ps = b;
  const size_t point_bytes = (size_t)b * es;
  // value tiles per lattice block: enough cells to fill a wave's rows, within the
  // LDS budget
  auto x_bytes = [&](int tic) { return ((size_t)3 * p * (size_t)(p * tic + 2 * p) * point_bytes + 15) / 16 * 16; };
  auto block_bytes = [&](int m) { return x_bytes(m * s->tile) + (tune.vlds ? nval * m * tile_vbytes : 0); };
  // smallest m whose block gives every row slot of a wave a cell and every wave a task
  int m = (a.rpw + s->tile - 1) / s->tile;
  while (p * p * ((m * s->tile + a.rpw - 1) / a.rpw) < kWaves) ++m;
  while (m > 1 && block_bytes(m) > tune.budget) --m;
  QDS_CHECK_ARG(block_bytes(m) <= kLdsMax, "qds_lattice_spmm: block width %d too large for LDS", b);
  const int tic = m * s->tile;
  a.m = m;
  a.tic = tic;
  a.nch = (tic + a.rpw - 1) / a.rpw;
  a.lw = p * tic + 2 * p;
  a.bx_per_row = (int32_t)((s->tiles_x + m - 1) / m);
  a.n_blocks = (int64_t)a.bx_per_row * s->cells_y;
  const int chunked = (point_bytes % 16 == 0) && aligned(X, 16) && ((ldx * es) % 16 == 0);
  a.row_chunks = chunked ? (int)(a.lw * point_bytes / 16) : 0;
  a.xchunks = 3 * p * a.row_chunks;
  a.vlds = (int)(x_bytes(tic) / 16);
  a.tile_chunks = (int)(tile_vbytes / 16);
  a.vchunks = m * a.tile_chunks;
  for (int c = 0; c < s->n_class; ++c) {
    a.len[c] = s->len[c];
    a.base[c] = (int32_t)s->base[c];
    for (int k = 0; k < s->len[c]; ++k) {
      QDS_CHECK_ARG(std::abs((int)s->di[c][k]) <= p && std::abs((int)s->dj[c][k]) <= p,
                    "qds_lattice_spmm: stencil offset beyond the halo");
      // LDS rows hold the points of each residue I mod p together (p runs of
      // tic + 2 points): a row of class column ci at run position q has its
      // neighbour (di, dj) in run (ci + di) mod p at position q + floor((ci + di) / p)
      const int ci = c % p, t = ci + s->di[c][k];
      const int run = ((t % p) + p) % p, shift = (t - run) / p;
      a.off[c][k] = (int16_t)(s->dj[c][k] * a.lw + (run - ci) * (tic + 2) + shift);
    }
  }
  // longest-task-first assignment of the (class, chunk) tasks to the waves
  const int n_tasks = s->n_class * a.nch;
  std::vector<int> order(n_tasks);
  for (int t = 0; t < n_tasks; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return a.len[x / a.nch] > a.len[y / a.nch]; });
  int load[kWaves] = {0}, cnt[kWaves] = {0};
  for (int w = 0; w < kWaves; ++w)
    for (int t = 0; t < kMaxTasks; ++t) a.tasks[w][t] = -1;
  for (int t : order) {
    int w = 0;
    for (int v = 1; v < kWaves; ++v)
      if (load[v] < load[w]) w = v;
    QDS_CHECK_ARG(cnt[w] < kMaxTasks, "qds_lattice_spmm: too many tasks per block");
    a.tasks[w][cnt[w]++] = (int8_t)t;
    load[w] += a.len[t / a.nch];
  }
  const ResEpilogue none{nullptr, nullptr, nullptr, 0, 0, 0, nullptr};
  const ResEpilogue& rs = res ? *res : none;
  size_t lds = block_bytes(m);
  if (res) {
    QDS_CHECK_ARG(fused && bs.batch == 1, "qds_lattice_spmm_residual: needs both matrices and batch 1");
    lds = std::max(lds, (size_t)kThreads * 2 * vec * sizeof(double));
  }
  hipStream_t st = qds_stream(stream);
  int rc;
  if (dtype == qds::F32)
    rc = vec == 4 ? dispatch<float, 4>(p, fused, tune.vlds, a, lds, vals_a, vals_b, X, ldx, chunked, Y_a, ldy_a, Y_b, ldy_b, rs, st)
       : vec == 2 ? dispatch<float, 2>(p, fused, tune.vlds, a, lds, vals_a, vals_b, X, ldx, chunked, Y_a, ldy_a, Y_b, ldy_b, rs, st)
                  : dispatch<float, 1>(p, fused, tune.vlds, a, lds, vals_a, vals_b, X, ldx, chunked, Y_a, ldy_a, Y_b, ldy_b, rs, st);
  else
    rc = vec == 2 ? dispatch<double, 2>(p, fused, tune.vlds, a, lds, vals_a, vals_b, X, ldx, chunked, Y_a, ldy_a, Y_b, ldy_b, rs, st)
                  : dispatch<double, 1>(p, fused, tune.vlds, a, lds, vals_a, vals_b, X, ldx, chunked, Y_a, ldy_a, Y_b, ldy_b, rs, st);
  if (rc != QDS_OK || !res) return rc;
  const int w = 2 * b;
  hipLaunchKernelGGL(res_reduce_rows, dim3(kResGroups), dim3(kThreads), 0, st, a.n_blocks, w, res->part, part1);
  QDS_LAUNCH_CHECK();
  hipLaunchKernelGGL(res_reduce_rows, dim3(1), dim3(kThreads), 0, st, (int64_t)kResGroups, w, part1, norms);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
}  // namespace

QDS_API int qds_lattice_spmm(int32_t dtype, const qds_lattice_stencil* s, int32_t b, const void* vals_a,
                             const void* vals_b, const void* X, int64_t ldx, void* Y_a, int64_t ldy_a,
                             void* Y_b, int64_t ldy_b, void* stream) {
  const BatchStrides one{1, 0, 0, 0, 0, 0};
  return lattice_spmm_impl(dtype, s, b, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, one, stream);
}

// The same product for `batch` problems on one stencil in one launch: problem i
// reads vals_a + i sva, vals_b + i svb (svb = 0: one shared matrix, e.g. M), X + i sx
// and writes Y_a + i sya, Y_b + i syb (strides in elements).
// The batched product with an explicit block shape (see qds_lattice_tuning): what
// the package's autotuner times and then uses.
QDS_API int qds_lattice_spmm_tuned(int32_t dtype, const qds_lattice_stencil* s, int32_t b, int32_t batch,
                                   const void* vals_a, int64_t sva, const void* vals_b, int64_t svb,
                                   const void* X, int64_t ldx, int64_t sx, void* Y_a, int64_t ldy_a,
                                   int64_t sya, void* Y_b, int64_t ldy_b, int64_t syb,
                                   const qds_lattice_tuning* tune, void* stream) {
  QDS_CHECK_ARG(tune, "qds_lattice_spmm_tuned: null tuning");
  QDS_CHECK_ARG(tune->lds_budget >= 4096 && tune->lds_budget <= (int32_t)kLdsMax,
                "qds_lattice_spmm_tuned: lds_budget %d outside [4096, %d]", tune->lds_budget, (int)kLdsMax);
  const BatchStrides bs{batch, sva, svb, sx, sya, syb};
  const Tuning t{tune->values_in_lds != 0, (size_t)tune->lds_budget};
  return lattice_spmm_impl(dtype, s, b, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, bs, stream, t);
}

// Workspace of qds_lattice_spmm_residual: one partial row of 2b sums per lattice
// block (bounded by one block per value tile) and kResGroups rows for the
// second reduction stage.
QDS_API int64_t qds_lattice_spmm_residual_workspace(const qds_lattice_stencil* s, int32_t b) {
  if (!s || b < 1) return 0;
  return ((int64_t)s->cells_y * s->tiles_x + kResGroups) * 2 * b * (int64_t)sizeof(double);
}

// The fused product Y_a = H X, Y_b = M X of qds_lattice_spmm_tuned (batch 1) with
// the residual step of the block iteration in its epilogue:
// R = (Y_a - Y_b diag(lam)) diag(colscale) stored as r_dtype (F32 or the product's
// type), norms[j] = ||(Y_a - Y_b diag(lam))_j||^2 and norms[b + j] = ||(Y_b)_j||^2
// (FP64, device).  R's values equal qds_residual_norms_to's for the same products;
// the norms are the same sums in another order.  colscale may be NULL.
QDS_API int qds_lattice_spmm_residual(int32_t dtype, const qds_lattice_stencil* s, int32_t b,
                                      const void* vals_a, const void* vals_b, const void* X, int64_t ldx,
                                      void* Y_a, int64_t ldy_a, void* Y_b, int64_t ldy_b, const double* lam,
                                      const double* colscale, int32_t r_dtype, void* R, int64_t ldr,
                                      double* norms, void* workspace, int64_t ws_bytes,
                                      const qds_lattice_tuning* tune, void* stream) {
  QDS_CHECK_ARG(tune && vals_b && Y_b && lam && R && norms && workspace, "qds_lattice_spmm_residual: null pointer");
  QDS_CHECK_ARG(tune->lds_budget >= 4096 && tune->lds_budget <= (int32_t)kLdsMax,
                "qds_lattice_spmm_residual: lds_budget %d outside [4096, %d]", tune->lds_budget, (int)kLdsMax);
  QDS_CHECK_ARG(r_dtype == qds::F32 || r_dtype == dtype, "qds_lattice_spmm_residual: R must be F32 or the product's type");
  QDS_CHECK_ARG(ws_bytes >= qds_lattice_spmm_residual_workspace(s, b), "qds_lattice_spmm_residual: workspace too small");
  QDS_CHECK_ARG(ldr >= b && b <= 128, "qds_lattice_spmm_residual: b %d must be at most 128 and ldr %lld at least b",
                b, (long long)ldr);
  double* part = static_cast<double*>(workspace);
  double* part1 = part + (int64_t)s->cells_y * s->tiles_x * 2 * b;
  const int32_t r_es = r_dtype == qds::F32 ? 4 : 8;
  // a lane stores VEC consecutive values of a row (VEC = 4 for FP32 products, 2 for
  // FP64): vector stores when R, ldr and the lane's first column keep them aligned
  const int64_t rvec_elems = dtype == qds::F32 ? 4 : 2;
  const int32_t r_vec = ((uintptr_t)R % (uintptr_t)(rvec_elems * r_es) == 0) && ldr % rvec_elems == 0;
  const ResEpilogue res{lam, colscale, R, ldr, r_es, r_vec, part};
  const BatchStrides one{1, 0, 0, 0, 0, 0};
  const Tuning t{tune->values_in_lds != 0, (size_t)tune->lds_budget};
  return lattice_spmm_impl(dtype, s, b, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, one, stream, t, &res, part1,
                           norms);
}

QDS_API int qds_lattice_spmm_batched(int32_t dtype, const qds_lattice_stencil* s, int32_t b, int32_t batch,
                                     const void* vals_a, int64_t sva, const void* vals_b, int64_t svb,
                                     const void* X, int64_t ldx, int64_t sx, void* Y_a, int64_t ldy_a,
                                     int64_t sya, void* Y_b, int64_t ldy_b, int64_t syb, void* stream) {
  const BatchStrides bs{batch, sva, svb, sx, sya, syb};
  return lattice_spmm_impl(dtype, s, b, vals_a, vals_b, X, ldx, Y_a, ldy_a, Y_b, ldy_b, bs, stream);
}
