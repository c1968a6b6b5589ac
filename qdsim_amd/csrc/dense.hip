// Tall-skinny dense block kernels for the block eigensolver (gfx950 MFMA).
//
// * qds_gram: C = A^T [B1 | B2] in FP64 with v_mfma_f64_16x16x4f64.  A and B are
//   n x m row-major blocks stored in FP32 or FP64; rows are staged through LDS as
//   FP64, so the Rayleigh-Ritz matrices are accumulated in FP64 whatever the
//   storage precision.  Every workgroup reduces a contiguous slab of rows into a
//   partial m x k tile set; a second kernel sums the partials in a fixed order
//   (deterministic, no atomics).  With sym, only tiles (ti <= tj) of each half are
//   computed and the rest are mirrored.
// * qds_block_update: Y1 = S C1 (and Y2 = S C2) with S n x m, C m x k, on the
//   16x16x4 MFMA of the storage type.  Each lane loads a 16-byte vector of one row
//   of S and the contraction index is permuted accordingly (the same permutation
//   is applied to the rows of C staged in LDS).
// * qds_block_residual, qds_coldot: R = HX - MX diag(lambda); column dot products.
#include "qds_common.h"

#include <mutex>
#include <type_traits>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kRB = 16;       // rows per LDS chunk in the Gram kernel
constexpr int kMaxTiles = 24; // per wave

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_f32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

struct GramShape {
  int m, k1, k2;
  int tm, tk1, tk2;  // tiles
  int mpad, kpad;
  int sym;
};

// Up to four column blocks forming one logical n x cols matrix (each block its own
// allocation and leading dimension).  start[j] is the first logical column of j.
// Device form of a column-block operand: the caller's blocks split into pieces of
// at most 16 vectors' width (kMaxPieces pieces).
constexpr int kMaxPieces = 8;
struct ColBlocks {
  const void* p[kMaxPieces];
  int64_t ld[kMaxPieces];
  int32_t start[kMaxPieces + 1];
  int32_t nb;
};

__device__ __forceinline__ bool tile_needed(const GramShape& g, int ti, int tj) {
  if (!g.sym) return true;
  const int tl = tj < g.tk1 ? tj : tj - g.tk1;
  return ti <= tl;
}

// Vector width of the global loads of the Gram kernel (elements of T per load).
template <typename T, bool VEC> struct GLoad;
template <> struct GLoad<float, true> { static constexpr int W = 4; typedef float4 v; };
template <> struct GLoad<float, false> { static constexpr int W = 1; typedef float v; };
template <> struct GLoad<double, true> { static constexpr int W = 2; typedef double2 v; };
template <> struct GLoad<double, false> { static constexpr int W = 1; typedef double v; };

// grid: (row slabs, output column groups of tpg tiles).  A block streams its slab
// kRB rows at a time: the rows are staged in LDS as FP64, [A | this group's B
// columns] per row, zero padded.  Operand pieces are at most 16 vectors wide, so
// for kRB = 16 rows every thread owns at most one vector per piece; its offsets
// (global, relative to the chunk's first row, and in LDS) are computed once, and
// per chunk only uniform row bases advance.  The next chunk is loaded into
// registers while the MFMAs of the current one run.  Each wave's MFMA loop reads
// all operands of a k-step before its MFMAs, without branches (unused slots compute
// into accumulators that are never stored).
constexpr int kGramPieces = 3 * kMaxPieces;

struct GramStage {
  int n;                           // pieces: A's, then B1's, then B2's
  const void* src[kGramPieces];
  int64_t ld[kGramPieces];
  int w[kGramPieces];              // columns
  int op[kGramPieces];             // 0 A, 1 B1, 2 B2
  int col0[kGramPieces];           // first column of the piece inside its operand
  int es[kGramPieces];             // TYPED kernels: bytes per element of the piece (4 FP32, 8 FP64)
};

// TYPED: every piece has its own storage type (stg.es), read with 16-byte loads
// (4 FP32 or 2 FP64 values per load) and converted to FP64 in LDS; T is unused.
template <typename T, bool VEC, int MAXP, int MAXT, bool TYPED = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
gram_partial(int64_t n, int64_t rows_per_blk, GramShape g, int tpg, int g0, GramStage stg,
             double* __restrict__ partial, int64_t s_stride) {
  using L = GLoad<T, VEC>;
  using V = typename std::conditional<TYPED, uint4, typename L::v>::type;
  constexpr int VW = L::W;
  extern __shared__ double gsm[];
  const int tk = g.tk1 + g.tk2;
  const int tj0 = (g0 + blockIdx.y) * tpg;
  const int tj1 = min(tk, tj0 + tpg);
  const int k1pad = 16 * g.tk1;
  const int c_lo = tj0 * 16, c_hi = tj1 * 16;
  const int ldsw = g.mpad + (c_hi - c_lo);     // LDS row: A (padded), then the group's B columns
  __shared__ int s_tile[kWaves * MAXT];
  __shared__ int s_count[kWaves];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ntl = g.tm * (tj1 - tj0);
  if (lane == 0) {
    int c = 0;
    for (int t = wave; t < ntl; t += kWaves) {
      const int ti = t / (tj1 - tj0), tj = tj0 + (t - ti * (tj1 - tj0));
      if (!tile_needed(g, ti, tj)) continue;
      if (c < MAXT) s_tile[wave * MAXT + c] = (ti << 8) | tj;
      ++c;
    }
    s_count[wave] = c;
  }
  for (int t = threadIdx.x; t < kRB * ldsw; t += kBlock) gsm[t] = 0.0;
  // this thread's vector of every piece: offset from the chunk's first row (global)
  // and position in LDS (-1: none)
  int goff[MAXP], loff[MAXP];
#pragma unroll
  for (int q = 0; q < MAXP; ++q) {
    goff[q] = -1;
    loff[q] = -1;
    if (q >= stg.n) continue;
    const int w = stg.w[q];
    const int vwq = TYPED ? 16 / stg.es[q] : VW;
    const int vpr = w / vwq;
    const int v = threadIdx.x;
    if (v >= kRB * vpr) continue;
    const int rr = v / vpr, cc = (v - rr * vpr) * vwq;
    int lcol;
    if (stg.op[q] == 0) {
      lcol = stg.col0[q] + cc;
    } else {
      const int tcol = (stg.op[q] == 1 ? 0 : k1pad) + stg.col0[q] + cc;   // tile-space column
      if (tcol < c_lo || tcol >= c_hi) continue;                          // vector inside one group
      lcol = g.mpad + tcol - c_lo;
    }
    goff[q] = (int)(rr * stg.ld[q]) + cc;
    loff[q] = rr * ldsw + lcol;
  }
  __syncthreads();
  const int nslots = s_count[wave];
  v4d acc[MAXT];
  int a_off[MAXT], b_off[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    const int code = t < nslots ? s_tile[wave * MAXT + t] : 0;
    a_off[t] = (code >> 8) * 16 + (lane & 15);
    b_off[t] = g.mpad + (code & 255) * 16 - c_lo + (lane & 15);
    if (t >= nslots) b_off[t] = g.mpad + (lane & 15);   // valid LDS, result unused
  }
  const int64_t r_begin = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t r_end = min(n, r_begin + rows_per_blk);
  V pre[MAXP];
  auto load = [&](int64_t r0) {
    const bool full = r0 + kRB <= r_end;
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      pre[q] = V{};
      if (goff[q] < 0) continue;
      if (!full && r0 + loff[q] / ldsw >= r_end) continue;   // tail chunk: rows past the slab stay 0
      if constexpr (TYPED) {
        const char* base = reinterpret_cast<const char*>(stg.src[q]) + (r0 * stg.ld[q] + goff[q]) * stg.es[q];
        pre[q] = *reinterpret_cast<const V*>(base);
      } else {
        const T* base = reinterpret_cast<const T*>(stg.src[q]) + blockIdx.z * s_stride + r0 * stg.ld[q];
        pre[q] = *reinterpret_cast<const V*>(base + goff[q]);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < MAXP; ++q) {
      if (loff[q] < 0) continue;
      if constexpr (TYPED) {
        if (stg.es[q] == 4) {
          const float* v = reinterpret_cast<const float*>(&pre[q]);
#pragma unroll
          for (int e = 0; e < 4; ++e) gsm[loff[q] + e] = (double)v[e];
        } else {
          const double* v = reinterpret_cast<const double*>(&pre[q]);
          gsm[loff[q]] = v[0];
          gsm[loff[q] + 1] = v[1];
        }
      } else {
        const T* v = reinterpret_cast<const T*>(&pre[q]);
#pragma unroll
        for (int e = 0; e < VW; ++e) gsm[loff[q] + e] = (double)v[e];
      }
    }
  };
  if (r_begin < r_end) load(r_begin);
  for (int64_t r0 = r_begin; r0 < r_end; r0 += kRB) {
    __syncthreads();  // the previous chunk's MFMAs are done with LDS
    store();
    __syncthreads();
    if (r0 + kRB < r_end) load(r0 + kRB);   // in flight during the MFMAs below
#pragma unroll
    for (int kk = 0; kk < kRB / 4; ++kk) {
      __builtin_amdgcn_sched_barrier(0);
      const int row = (kk * 4 + (lane >> 4)) * ldsw;
      double av[MAXT], bv[MAXT];
#pragma unroll
      for (int slot = 0; slot < MAXT; ++slot) {
        av[slot] = gsm[row + a_off[slot]];
        bv[slot] = gsm[row + b_off[slot]];
      }
#pragma unroll
      for (int slot = 0; slot < MAXT; ++slot) acc[slot] = mfma_f64(av[slot], bv[slot], acc[slot]);
    }
  }
  double* P = partial + ((int64_t)blockIdx.z * gridDim.x + blockIdx.x) * g.mpad * g.kpad;
#pragma unroll
  for (int slot = 0; slot < MAXT; ++slot) {
    if (slot < nslots) {
      const int code = s_tile[wave * MAXT + slot];
      const int ti = code >> 8, tj = code & 255;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // v_mfma_f64_16x16x4f64 result: lane l, register r holds D[l/16 + 4r][l%16]
        const int i = ti * 16 + (lane >> 4) + 4 * r;
        const int j = tj * 16 + (lane & 15);
        P[i * g.kpad + j] = acc[slot][r];
      }
    }
  }
}

// grid: one block per 16 outputs; 16 thread groups of a block split the partial
// rows (group g sums rows g, g + 16, ...), then combine in a fixed order.
constexpr int kRedOut = 16;
constexpr int kRedGroups = kBlock / kRedOut;
__global__ void __launch_bounds__(kBlock)
gram_reduce(int nblk, GramShape g, const double* __restrict__ partial, double* __restrict__ C) {
  // problem of a batch (blockIdx.y): its slabs' partials and its output
  partial += (int64_t)blockIdx.y * nblk * g.mpad * g.kpad;
  C += (int64_t)blockIdx.y * g.m * (g.k1 + g.k2);
  __shared__ double red[kBlock];
  const int ktot = g.k1 + g.k2;
  const int k1pad = 16 * g.tk1;
  const int o = threadIdx.x % kRedOut;
  const int grp = threadIdx.x / kRedOut;
  const int t = blockIdx.x * kRedOut + o;
  double a = 0.0;
  if (t < g.m * ktot) {
    const int i = t / ktot, j = t - i * ktot;
    int jp = j < g.k1 ? j : k1pad + (j - g.k1);
    int ip = i;
    const int half_off = j < g.k1 ? 0 : k1pad;
    if (g.sym) {
      const int jl = jp - half_off;
      if ((ip >> 4) > (jl >> 4)) {  // mirrored tile
        const int tmp = ip;
        ip = jl;
        jp = tmp + half_off;
      }
    }
    const int64_t stride = (int64_t)g.mpad * g.kpad;
    const double* src = partial + ip * g.kpad + jp;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = grp;
    for (; b + 3 * kRedGroups < nblk; b += 4 * kRedGroups) {
      a0 += src[(int64_t)b * stride];
      a1 += src[(int64_t)(b + kRedGroups) * stride];
      a2 += src[(int64_t)(b + 2 * kRedGroups) * stride];
      a3 += src[(int64_t)(b + 3 * kRedGroups) * stride];
    }
    for (; b < nblk; b += kRedGroups) a0 += src[(int64_t)b * stride];
    a = (a0 + a1) + (a2 + a3);
  }
  red[threadIdx.x] = a;
  __syncthreads();
  if (grp == 0 && t < g.m * ktot) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kRedGroups; ++q) s += red[q * kRedOut + o];
    C[t] = s;
  }
}

// ---------------------------------------------------------------- block update
// Operand layouts of both 16x16x4 MFMAs: A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16].  Result rows differ: FP32 lane l register e holds D[4(l/16)+e][l%16],
// FP64 holds D[l/16 + 4e][l%16].
template <typename T> struct MT;
template <> struct MT<float> {
  static constexpr int VEC = 4;
  typedef float4 vec_t;
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) { return mfma_f32(a, b, c); }
  static __device__ __forceinline__ int drow(int q, int e) { return 4 * q + e; }
};
template <> struct MT<double> {
  static constexpr int VEC = 2;
  typedef double2 vec_t;
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) { return mfma_f64(a, b, c); }
  static __device__ __forceinline__ int drow(int q, int e) { return q + 4 * e; }
};

constexpr int kMaxOutTiles = 4;  // k <= 64
constexpr int kUpdLds = 64 * 1024;

// Y1 = S C1 (and Y2 = S C2), S given as column blocks.  A wave owns a tile of 16
// rows; a block of 4 waves advances 64 rows per pass.  Within a pass the column
// blocks of S are visited in order: the MAXJ 16-byte loads per lane of block j + 1
// are issued before the MFMAs of block j (always MAXJ loads, with clamped
// addresses, so the compiler can wait for exactly the older ones), and every wave
// keeps one block in flight while it computes.  C lives in LDS: all of it when it
// fits (resident), otherwise the rows of the current column block, staged with
// barriers.  OT = ceil(k / 16) output tiles exactly.  Column blocks that end before
// c2_first contribute nothing to Y2 (those rows of C2 are zero) and skip its MFMAs.
// C arrives in FP64 and is staged in the storage type; the MFMA runs in the storage
// type.  (FP64 accumulation for FP32 storage was measured: it lowered the FP32
// residual floor of the eigensolver from ~1e-3 to ~3e-4 but made the update
// MFMA-bound at twice the time, a net loss.)
template <typename T, int NOUT, int OT, int MAXJ>
__global__ void __launch_bounds__(kBlock)
block_update_kernel(int64_t n, int m, int k, ColBlocks S, const double* __restrict__ C1,
                    const double* __restrict__ C2, T* __restrict__ Y1, int64_t ldy1, T* __restrict__ Y2,
                    int64_t ldy2, int resident, int crows, int c2_first, int64_t sS, int64_t cstride, int64_t sY) {
  using M = MT<T>;
  constexpr int VEC = M::VEC;                  // elements of T per 16-byte load
  constexpr int KB = 4 * VEC;  // contraction indices covered by one vector load per lane
  constexpr int KP = OT * 16;
  using Vt = typename MT<T>::vec_t;
  extern __shared__ unsigned char upd_smem[];
  T* sC = reinterpret_cast<T*>(upd_smem);   // [NOUT][crows][KP], zero padded
  // problem of a batch (blockIdx.y): its S blocks, coefficients and outputs
  const int64_t set = blockIdx.y;
  C1 += set * cstride;
  if (C2) C2 += set * cstride;
  Y1 += set * sY;
  if (NOUT == 2) Y2 += set * sY;
  const int64_t s_off = set * sS;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int q = lane >> 4;      // 0..3
  const int rl = lane & 15;
  // copy C rows [r0, r0 + rows) into LDS rows [0, rows), zero elsewhere
  auto stage = [&](int r0, int rows) {
    for (int t = threadIdx.x; t < NOUT * crows * KP; t += kBlock) {
      const int o = t / (crows * KP);
      const int rem = t - o * crows * KP;
      const int i = rem / KP, j = rem - i * KP;
      const double* C = o == 0 ? C1 : C2;
      sC[t] = (i < rows && j < k) ? (T)C[(r0 + i) * k + j] : T(0);
    }
  };
  if (resident) {
    stage(0, m);
    __syncthreads();
  }
  const int64_t rows_per_pass = (int64_t)kWaves * 16;
  const int64_t first = (int64_t)blockIdx.x * rows_per_pass;
  const int64_t stride = (int64_t)gridDim.x * rows_per_pass;
  Vt nxt[MAXJ];
  // MAXJ 16-byte loads of column block j for this lane's row of the tile at `base`;
  // out-of-range rows and columns load a valid address and are zeroed or unused later
  auto fetch = [&](int64_t base, int j) {
    int64_t row = base + wave * 16 + rl;
    row = row < n ? row : n - 1;
    const int w = S.start[j + 1] - S.start[j];
    const T* rp = reinterpret_cast<const T*>(S.p[j]) + s_off + row * S.ld[j];
#pragma unroll
    for (int jj = 0; jj < MAXJ; ++jj) {
      const int c = jj * KB + q * VEC;
      nxt[jj] = *reinterpret_cast<const Vt*>(rp + (c < w ? c : 0));
    }
  };
  if (first < n) fetch(first, 0);
  for (int64_t base = first; base < n; base += stride) {
    typename M::acc_t acc[NOUT][OT];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
      for (int t = 0; t < OT; ++t) acc[o][t] = typename M::acc_t{0, 0, 0, 0};
    for (int j = 0; j < S.nb; ++j) {
      Vt cur[MAXJ];
#pragma unroll
      for (int jj = 0; jj < MAXJ; ++jj) cur[jj] = nxt[jj];
      const int w = S.start[j + 1] - S.start[j];
      const int cj = S.start[j];
      if (!resident) {
        __syncthreads();
        stage(cj, w);
        __syncthreads();
      }
      if (j + 1 < S.nb) fetch(base, j + 1);
      else if (base + stride < n) fetch(base + stride, 0);
      const int crow0 = resident ? cj : 0;
      const bool two = NOUT == 2 && cj + w > c2_first;
#pragma unroll
      for (int jj = 0; jj < MAXJ; ++jj) {
        if (jj * KB >= w) break;
        __builtin_amdgcn_sched_barrier(0);
        // columns past the block end carry zeros
        const T keep = jj * KB + q * VEC < w ? T(1) : T(0);
        const T* sv = reinterpret_cast<const T*>(&cur[jj]);
        T bv[NOUT][VEC][OT];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const int kl = crow0 + jj * KB + q * VEC + e;   // staged C row (zero past the block)
#pragma unroll
          for (int t = 0; t < OT; ++t) {
            bv[0][e][t] = sC[kl * KP + t * 16 + rl];
            if (NOUT == 2) bv[NOUT - 1][e][t] = sC[(crows + kl) * KP + t * 16 + rl];
          }
        }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const T a = sv[e] * keep;
#pragma unroll
          for (int t = 0; t < OT; ++t) acc[0][t] = M::mma(a, bv[0][e][t], acc[0][t]);
        }
        if (two) {
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const T a = sv[e] * keep;
#pragma unroll
            for (int t = 0; t < OT; ++t) acc[NOUT - 1][t] = M::mma(a, bv[NOUT - 1][e][t], acc[NOUT - 1][t]);
          }
        }
      }
    }
    const int64_t r0 = base + wave * 16;
#pragma unroll
    for (int t = 0; t < OT; ++t) {
      const int col = t * 16 + rl;
      if (col >= k) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t orow = r0 + M::drow(q, e);
        if (orow < n) {
          Y1[orow * ldy1 + col] = (T)acc[0][t][e];
          if (NOUT == 2) Y2[orow * ldy2 + col] = (T)acc[NOUT - 1][t][e];
        }
      }
    }
  }
}

template <typename T, int NOUT>
int launch_update(int64_t n, int m, int k, const ColBlocks& S, int c2_first, const double* C1, const double* C2,
                  T* Y1, int64_t ldy1, T* Y2, int64_t ldy2, hipStream_t st, int batch = 1, int64_t sS = 0,
                  int64_t sC = 0, int64_t sY = 0) {
  constexpr int KB = 4 * MT<T>::VEC;
  const int ot = (k + 15) / 16;
  const int kp = ot * 16;
  int wmax = 0;
  for (int j = 0; j < S.nb; ++j) wmax = max(wmax, S.start[j + 1] - S.start[j]);
  QDS_CHECK_ARG(wmax <= 4 * KB, "qds_block_update: column blocks wider than %d", 4 * KB);
  // staged C rows: all m (+ one vector of zero padding) or one column block
  int crows = m + KB;
  int resident = 1;
  if ((size_t)NOUT * crows * kp * sizeof(T) > (size_t)kUpdLds) {
    resident = 0;
    crows = wmax + KB;
  }
  QDS_CHECK_ARG((size_t)NOUT * crows * kp * sizeof(T) <= (size_t)kUpdLds, "qds_block_update: C too large");
  const size_t shm = (size_t)NOUT * crows * kp * sizeof(T);
  const int64_t passes = (n + kWaves * 16 - 1) / (kWaves * 16);
  const unsigned blocks = qds_blocks(passes, 1, batch > 1 ? (4096 + batch - 1) / batch * 4 : 4096);
#define QDS_UPD(OTV, J)                                                                                       \
  hipLaunchKernelGGL((block_update_kernel<T, NOUT, OTV, J>), dim3(blocks, (unsigned)batch), dim3(kBlock), shm, \
                     st, n, m, k, S, C1, C2, Y1, ldy1, Y2, ldy2, resident, crows, c2_first, sS, sC, sY)
#define QDS_UPD_J(OTV) do { if (J <= 2) QDS_UPD(OTV, 2); else if (J == 3) QDS_UPD(OTV, 3); else QDS_UPD(OTV, 4); } while (0)
  const int J = (wmax + KB - 1) / KB;
  switch (ot) {
    case 1: QDS_UPD_J(1); break;
    case 2: QDS_UPD_J(2); break;
    case 3: QDS_UPD_J(3); break;
    default: QDS_UPD_J(4); break;
  }
#undef QDS_UPD_J
#undef QDS_UPD
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// ---------------------------------------------------------------- typed update
// Y1 = S C1 (FP64) and Y2 = S C2 (TY2) for column blocks of their own storage
// type (es[j] bytes per element): the update of the mixed-precision iteration,
// whose basis [X | W | P] keeps X in FP64 and the corrections W, P in FP32.  The
// layout and pipelining follow block_update_kernel.  FP64 blocks run on the FP64
// MFMA with C in FP64; FP32 blocks on the FP32 MFMA (twice the rate; gfx950's FP64
// MFMA bounds this kernel otherwise) with C rounded to FP32, into their own
// accumulators that are added in FP64 at the end: the FP32 blocks' contribution
// (the correction) is then exact to FP32 rounding relative to its own size.  The
// two MFMAs place result rows differently (FP64: lane l, register e holds row
// l/16 + 4e; FP32: row 4(l/16) + e), so an FP32 block's lane loads tile row
// pi(l % 16), pi(i) = i/4 + 4 (i % 4), which puts both results of a row in the same
// lane and register.  C is staged in LDS per precision: rows of FP64 blocks as
// FP64 (cbase[j] = first staged row of block j), rows of FP32 blocks as FP32.
struct TypedBlocks {
  const void* p[4];
  int64_t ld[4];
  int32_t start[5];
  int32_t es[4];
  int32_t cbase[4];
  int32_t nb;
};

// Four waves per SIMD (128 registers, a few spilled outside the MFMA loop): the
// kernel is bound by the loads in flight, one column block ahead per wave; at the
// compiler's own choice (164 registers, three waves) it took 472 us on the
// benchmark basis against 439 us.
template <int OT, int MAXJ, typename TY2, bool TWO>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
block_update_typed_kernel(int64_t n, int m, int k, TypedBlocks S, const double* __restrict__ C1,
                          const double* __restrict__ C2, double* __restrict__ Y1, int64_t ldy1,
                          TY2* __restrict__ Y2, int64_t ldy2, int rows64, int rows32, int c2_first) {
  constexpr int KP = OT * 16;
  constexpr int NOUT = TWO ? 2 : 1;
  extern __shared__ unsigned char upd_smem[];
  double* sC64 = reinterpret_cast<double*>(upd_smem);              // [NOUT][rows64][KP]
  float* sC32 = reinterpret_cast<float*>(sC64 + NOUT * rows64 * KP);  // [NOUT][rows32][KP]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int q = lane >> 4;
  const int rl = lane & 15;
  const int prl = (rl >> 2) + 4 * (rl & 3);   // pi(rl)
  // staged row r of a region -> C row (or -1: zero)
  auto crow_of = [&](int r, int es) {
    for (int j = 0; j < S.nb; ++j) {
      if (S.es[j] != es) continue;
      const int w = S.start[j + 1] - S.start[j];
      if (r >= S.cbase[j] && r < S.cbase[j] + w) return S.start[j] + (r - S.cbase[j]);
    }
    return -1;
  };
  for (int t = threadIdx.x; t < NOUT * rows64 * KP; t += kBlock) {
    const int o = t / (rows64 * KP);
    const int rem = t - o * rows64 * KP;
    const int i = rem / KP, j = rem - i * KP;
    const int cr = crow_of(i, 8);
    const double* C = o == 0 ? C1 : C2;
    sC64[t] = (cr >= 0 && j < k) ? C[cr * k + j] : 0.0;
  }
  for (int t = threadIdx.x; t < NOUT * rows32 * KP; t += kBlock) {
    const int o = t / (rows32 * KP);
    const int rem = t - o * rows32 * KP;
    const int i = rem / KP, j = rem - i * KP;
    const int cr = crow_of(i, 4);
    const double* C = o == 0 ? C1 : C2;
    sC32[t] = (cr >= 0 && j < k) ? (float)C[cr * k + j] : 0.f;
  }
  __syncthreads();
  const int64_t rows_per_pass = (int64_t)kWaves * 16;
  const int64_t first = (int64_t)blockIdx.x * rows_per_pass;
  const int64_t stride = (int64_t)gridDim.x * rows_per_pass;
  uint4 nxt[MAXJ];
  auto fetch = [&](int64_t base, int j) {
    const int es = S.es[j];
    int64_t row = base + wave * 16 + (es == 4 ? prl : rl);
    row = row < n ? row : n - 1;
    const int w = S.start[j + 1] - S.start[j];
    const int vec = 16 / es;
    const char* rp = reinterpret_cast<const char*>(S.p[j]) + row * S.ld[j] * es;
#pragma unroll
    for (int jj = 0; jj < MAXJ; ++jj) {
      const int c = jj * 4 * vec + q * vec;
      nxt[jj] = *reinterpret_cast<const uint4*>(rp + (size_t)(c < w ? c : 0) * es);
    }
  };
  if (first < n) fetch(first, 0);
  for (int64_t base = first; base < n; base += stride) {
    v4d acc[NOUT][OT];
    v4f acc32[NOUT][OT];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
      for (int t = 0; t < OT; ++t) {
        acc[o][t] = v4d{0.0, 0.0, 0.0, 0.0};
        acc32[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    for (int j = 0; j < S.nb; ++j) {
      uint4 cur[MAXJ];
#pragma unroll
      for (int jj = 0; jj < MAXJ; ++jj) cur[jj] = nxt[jj];
      const int w = S.start[j + 1] - S.start[j];
      const int cb = S.cbase[j];
      const bool f32 = S.es[j] == 4;
      if (j + 1 < S.nb) fetch(base, j + 1);
      else if (base + stride < n) fetch(base + stride, 0);
      const bool two = TWO && S.start[j] + w > c2_first;
      if (f32) {
#pragma unroll
        for (int jj = 0; jj < MAXJ; ++jj) {
          if (jj * 16 >= w) break;
          __builtin_amdgcn_sched_barrier(0);
          const float keep = jj * 16 + q * 4 < w ? 1.f : 0.f;
          const float* fv = reinterpret_cast<const float*>(&cur[jj]);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float a = fv[e] * keep;
            const int kl = cb + jj * 16 + q * 4 + e;   // staged FP32 row (zero past the block)
#pragma unroll
            for (int t = 0; t < OT; ++t) acc32[0][t] = mfma_f32(a, sC32[kl * KP + t * 16 + rl], acc32[0][t]);
            if (two) {
#pragma unroll
              for (int t = 0; t < OT; ++t)
                acc32[NOUT - 1][t] = mfma_f32(a, sC32[(rows32 + kl) * KP + t * 16 + rl], acc32[NOUT - 1][t]);
            }
          }
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < MAXJ; ++jj) {
          if (jj * 8 >= w) break;
          __builtin_amdgcn_sched_barrier(0);
          const double keep = jj * 8 + q * 2 < w ? 1.0 : 0.0;
          const double* dv = reinterpret_cast<const double*>(&cur[jj]);
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const double a = dv[e] * keep;
            const int kl = cb + jj * 8 + q * 2 + e;    // staged FP64 row (zero past the block)
#pragma unroll
            for (int t = 0; t < OT; ++t) acc[0][t] = mfma_f64(a, sC64[kl * KP + t * 16 + rl], acc[0][t]);
            if (two) {
#pragma unroll
              for (int t = 0; t < OT; ++t)
                acc[NOUT - 1][t] = mfma_f64(a, sC64[(rows64 + kl) * KP + t * 16 + rl], acc[NOUT - 1][t]);
            }
          }
        }
      }
    }
    const int64_t r0 = base + wave * 16;
#pragma unroll
    for (int t = 0; t < OT; ++t) {
      const int col = t * 16 + rl;
      if (col >= k) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t orow = r0 + q + 4 * e;   // FP64 result row (the FP32 rows were permuted to match)
        if (orow < n) {
          Y1[orow * ldy1 + col] = acc[0][t][e] + (double)acc32[0][t][e];
          if (TWO) Y2[orow * ldy2 + col] = (TY2)(acc[NOUT - 1][t][e] + (double)acc32[NOUT - 1][t][e]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- residual / dots
template <typename T>
__global__ void residual_kernel(int64_t n, int b, const T* __restrict__ HX, int64_t ldh,
                                const T* __restrict__ MX, int64_t ldm, const double* __restrict__ lam,
                                T* __restrict__ R, int64_t ldr) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * b;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / b;
    const int c = (int)(t - r * b);
    R[r * ldr + c] = HX[r * ldh + c] - (T)lam[c] * MX[r * ldm + c];
  }
}

// complex residual with complex lambda (interleaved); bilinear (no conjugation)
template <typename T>
__global__ void residual_cplx_kernel(int64_t n, int b, const T* __restrict__ HX, int64_t ldh,
                                     const T* __restrict__ MX, int64_t ldm,
                                     const double* __restrict__ lam, T* __restrict__ R, int64_t ldr) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * b;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / b;
    const int c = (int)(t - r * b);
    const T lr = (T)lam[2 * c], li = (T)lam[2 * c + 1];
    const T mr = MX[2 * (r * ldm + c)], mi = MX[2 * (r * ldm + c) + 1];
    R[2 * (r * ldr + c)] = HX[2 * (r * ldh + c)] - (lr * mr - li * mi);
    R[2 * (r * ldr + c) + 1] = HX[2 * (r * ldh + c) + 1] - (lr * mi + li * mr);
  }
}

// partial column dots: mode 0 real x.y ; mode 1 complex conj(x).y (re part + im part);
// mode 2 complex bilinear x.y
template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock)
coldot_partial(int64_t n, int b, const T* __restrict__ X, int64_t ldx, const T* __restrict__ Y,
               int64_t ldy, int64_t rows_per_blk, double* __restrict__ partial) {
  // thread layout: threads cover columns (b <= 256 / groups), rows strided
  const int cols = MODE == 0 ? b : b;
  const int groups = kBlock / cols;
  const int tc = threadIdx.x % cols;
  const int tg = threadIdx.x / cols;
  __shared__ double red[2 * kBlock];
  double s0 = 0.0, s1 = 0.0;
  const int64_t r_begin = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t r_end = min(n, r_begin + rows_per_blk);
  if (tg < groups) {
    for (int64_t r = r_begin + tg; r < r_end; r += groups) {
      if (MODE == 0) {
        s0 += (double)X[r * ldx + tc] * (double)Y[r * ldy + tc];
      } else {
        const double xr = X[2 * (r * ldx + tc)], xi = X[2 * (r * ldx + tc) + 1];
        const double yr = Y[2 * (r * ldy + tc)], yi = Y[2 * (r * ldy + tc) + 1];
        if (MODE == 1) { s0 += xr * yr + xi * yi; s1 += xr * yi - xi * yr; }
        else { s0 += xr * yr - xi * yi; s1 += xr * yi + xi * yr; }
      }
    }
  }
  red[threadIdx.x] = s0;
  red[kBlock + threadIdx.x] = s1;
  __syncthreads();
  if (threadIdx.x < cols) {
    double a0 = 0.0, a1 = 0.0;
    for (int g = 0; g < groups; ++g) {
      a0 += red[g * cols + threadIdx.x];
      a1 += red[kBlock + g * cols + threadIdx.x];
    }
    partial[(int64_t)blockIdx.x * 2 * b + threadIdx.x] = a0;
    partial[(int64_t)blockIdx.x * 2 * b + b + threadIdx.x] = a1;
  }
}

// One block of 1024 threads; partial row k holds (re[b], im[b]) sums of block k.
__global__ void __launch_bounds__(1024)
coldot_reduce(int nblk, int b, const double* __restrict__ partial, double* __restrict__ out, int cplx) {
  __shared__ double red[1024];
  const int w = 2 * b;
  const int groups = 1024 / w;
  const int c = threadIdx.x % w;
  const int g = threadIdx.x / w;
  double a = 0.0;
  if (g < groups)
    for (int k = g; k < nblk; k += groups) a += partial[(int64_t)k * w + c];
  red[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x < w) {
    double s = 0.0;
    for (int q = 0; q < groups; ++q) s += red[q * w + threadIdx.x];
    const int j = threadIdx.x % b;
    const int part = threadIdx.x / b;   // 0: real sums, 1: imaginary sums
    if (cplx) out[2 * j + part] = s;
    else if (part == 0) out[j] = s;
  }
}

int gram_blocks(int64_t n) {
  int64_t chunks = (n + kRB - 1) / kRB;
  int64_t g = chunks < 512 ? chunks : 512;
  return (int)(g < 1 ? 1 : g);
}

GramShape make_shape(int m, int k1, int k2, int sym) {
  GramShape g;
  g.m = m; g.k1 = k1; g.k2 = k2;
  g.tm = (m + 15) / 16; g.tk1 = (k1 + 15) / 16; g.tk2 = (k2 + 15) / 16;
  g.mpad = 16 * g.tm; g.kpad = 16 * (g.tk1 + g.tk2);
  g.sym = sym;
  return g;
}

constexpr int kColdotBlocks = 512;
// fewest row slabs the no-alloc Gram accepts (the workspace form uses up to 512)
constexpr int kMinStaticSlabs = 64;

}  // namespace

QDS_API int64_t qds_gram_workspace(int64_t n, int32_t m, int32_t k1, int32_t k2) {
  GramShape g = make_shape(m, k1, k2, 0);
  return (int64_t)gram_blocks(n) * g.mpad * g.kpad * (int64_t)sizeof(double);
}

extern "C" {
// Host description of a logical matrix made of up to four column blocks.
typedef struct {
  const void* p[4];
  int64_t ld[4];
  int32_t cols[4];
  int32_t nb;
} qds_colblocks;
}

namespace {
// Split the caller's blocks into pieces of at most wcap columns.
bool to_device_blocks(const qds_colblocks* h, ColBlocks& d, int& total, int es, int wcap) {
  if (!h || h->nb < 1 || h->nb > 4) return false;
  int np = 0, acc = 0;
  for (int j = 0; j < h->nb; ++j) {
    if (!h->p[j] || h->cols[j] < 1 || h->ld[j] < h->cols[j]) return false;
    for (int c = 0; c < h->cols[j]; c += wcap) {
      if (np == kMaxPieces) return false;
      d.p[np] = static_cast<const char*>(h->p[j]) + (size_t)c * es;
      d.ld[np] = h->ld[j];
      d.start[np] = acc;
      acc += h->cols[j] - c < wcap ? h->cols[j] - c : wcap;
      ++np;
    }
  }
  for (int j = np; j <= kMaxPieces; ++j) {
    if (j < kMaxPieces) { d.p[j] = nullptr; d.ld[j] = 0; }
    d.start[j] = acc;
  }
  d.nb = np;
  total = acc;
  return true;
}

// 16-byte vector access is possible when every block starts 16-byte aligned and
// its leading dimension and width are multiples of the vector width
bool blocks_vec(const qds_colblocks* h, int es) {
  const int vw = 16 / es;
  for (int j = 0; j < h->nb; ++j)
    if ((reinterpret_cast<uintptr_t>(h->p[j]) % 16) || (h->ld[j] * es) % 16 || h->cols[j] % vw) return false;
  return true;
}

qds_colblocks single(const void* p, int64_t ld, int cols) {
  qds_colblocks b;
  for (int j = 0; j < 4; ++j) { b.p[j] = nullptr; b.ld[j] = 0; b.cols[j] = 0; }
  b.p[0] = p; b.ld[0] = ld; b.cols[0] = cols; b.nb = 1;
  return b;
}
}  // namespace

namespace {
// C (m x (k1+k2), row-major FP64) = A^T [B1 | B2] for column-block matrices, with
// the per-slab partial sums in `partial`.  The row slabs are gram_blocks(n), at
// most max_slabs of them; partial_bytes == 0 takes the caller's qds_gram_workspace
// size check instead of the slab cap.
// Pieces of one operand whose blocks each have their own storage type (dts[j]):
// 16-byte vector access, pieces of at most 16 vectors (64 FP32 / 32 FP64 columns).
bool typed_pieces(const qds_colblocks* h, const int32_t* dts, int op, GramStage& stg, int& total) {
  if (!h || !dts || h->nb < 1 || h->nb > 4) return false;
  int acc = 0;
  for (int j = 0; j < h->nb; ++j) {
    if (dts[j] != qds::F32 && dts[j] != qds::F64) return false;
    const int es = qds::dtype_size(dts[j]);
    const int vw = 16 / es, wcap = 16 * vw;
    if (!h->p[j] || h->cols[j] < 1 || h->ld[j] < h->cols[j] || (reinterpret_cast<uintptr_t>(h->p[j]) % 16) ||
        (h->ld[j] * es) % 16 || h->cols[j] % vw)
      return false;
    for (int c = 0; c < h->cols[j]; c += wcap) {
      if (stg.n == kGramPieces) return false;
      const int w = h->cols[j] - c < wcap ? h->cols[j] - c : wcap;
      stg.src[stg.n] = static_cast<const char*>(h->p[j]) + (size_t)c * es;
      stg.ld[stg.n] = h->ld[j];
      stg.w[stg.n] = w;
      stg.op[stg.n] = op;
      stg.col0[stg.n] = acc;
      stg.es[stg.n] = es;
      acc += w;
      ++stg.n;
    }
  }
  total = acc;
  return true;
}

// dts (optional): per-operand arrays of per-block storage types (F32 / F64) for
// operands whose blocks differ in type; then dtype is ignored, batch must be 1
// and every block needs 16-byte vector access.
int gram_blocks_impl(int32_t dtype, int64_t n, const qds_colblocks* A, const qds_colblocks* B1,
                     const qds_colblocks* B2, int32_t sym, double* C, double* partial, int64_t ws_bytes,
                     int64_t partial_bytes, void* stream, int batch = 1, int64_t s_stride = 0,
                     const int32_t* const* dts = nullptr) {
  QDS_CHECK_ARG(batch >= 1 && batch <= 65535, "qds_gram: batch %d out of range", batch);
  const bool typed = dts != nullptr;
  QDS_CHECK_ARG(typed || dtype == qds::F32 || dtype == qds::F64, "qds_gram: dtype must be F32 or F64");
  QDS_CHECK_ARG(!typed || batch == 1, "qds_gram: typed blocks need batch 1");
  QDS_CHECK_ARG(C && partial && A && B1, "qds_gram: null pointer");
  const int es = dtype == qds::F32 ? 4 : 8;
  bool vec = true;
  int m = 0, k1 = 0, k2 = 0;
  GramStage stg;
  stg.n = 0;
  if (typed) {
    QDS_CHECK_ARG(typed_pieces(A, dts[0], 0, stg, m) && typed_pieces(B1, dts[1], 1, stg, k1) &&
                      (!B2 || typed_pieces(B2, dts[2], 2, stg, k2)),
                  "qds_gram: typed blocks need F32/F64 types, 16-byte alignment, widths and leading dimensions "
                  "a multiple of the vector width, at most %d pieces", kGramPieces);
  } else {
    vec = blocks_vec(A, es) && blocks_vec(B1, es) && (!B2 || blocks_vec(B2, es));
    const int wcap = vec ? 16 * (16 / es) : 16;   // piece: at most one vector per thread for kRB rows
    ColBlocks dA, dB1, dB2;
    QDS_CHECK_ARG(to_device_blocks(A, dA, m, es, wcap) && to_device_blocks(B1, dB1, k1, es, wcap),
                  "qds_gram: bad or too wide A / B1 blocks (at most %d pieces of %d columns)", kMaxPieces, wcap);
    dB2.nb = 0;
    if (B2) QDS_CHECK_ARG(to_device_blocks(B2, dB2, k2, es, wcap), "qds_gram: bad or too wide B2 blocks");
    const ColBlocks* ops[3] = {&dA, &dB1, &dB2};
    for (int o = 0; o < 3; ++o)
      for (int j = 0; j < ops[o]->nb; ++j) {
        stg.src[stg.n] = ops[o]->p[j];
        stg.ld[stg.n] = ops[o]->ld[j];
        stg.w[stg.n] = ops[o]->start[j + 1] - ops[o]->start[j];
        stg.op[stg.n] = o;
        stg.col0[stg.n] = ops[o]->start[j];
        stg.es[stg.n] = es;
        ++stg.n;
      }
  }
  QDS_CHECK_ARG(!sym || (m == k1 && (k2 == 0 || k2 == m)), "qds_gram: sym needs square halves");
  GramShape g = make_shape(m, k1, k2, sym);
  QDS_CHECK_ARG(g.tm <= 255 && g.tk1 + g.tk2 <= 255, "qds_gram: block too wide");
  const int tk = g.tk1 + g.tk2;
  const int k1pad = 16 * g.tk1;
  // column groups: at most 16 accumulator tiles per wave, the staged rows in 64 KiB
  auto slots = [&](int tp) {
    int mx = 0;
    for (int j0 = 0; j0 < tk; j0 += tp) {
      const int j1 = j0 + tp < tk ? j0 + tp : tk;
      for (int w = 0; w < kWaves; ++w) {
        int c = 0;
        const int nt = g.tm * (j1 - j0);
        for (int t = w; t < nt; t += kWaves) {
          const int ti = t / (j1 - j0), tj = j0 + (t - ti * (j1 - j0));
          const int tl = tj < g.tk1 ? tj : tj - g.tk1;
          if (!sym || ti <= tl) ++c;
        }
        mx = c > mx ? c : mx;
      }
    }
    return mx;
  };
  // at most 8 accumulator tiles per wave: 16 spill registers to scratch
  int tpg = tk, max_slots = 0;
  for (; tpg >= 1; --tpg) {
    max_slots = slots(tpg);
    if (max_slots <= 8 && (size_t)kRB * (g.mpad + 16 * tpg) * sizeof(double) <= 64 * 1024) break;
  }
  QDS_CHECK_ARG(tpg >= 1, "qds_gram: operands of %d columns too wide", m);
  const int ngroups = (tk + tpg - 1) / tpg;
  int nblk = gram_blocks(n);
  if (partial_bytes == 0) {
    QDS_CHECK_ARG(ws_bytes >= batch * qds_gram_workspace(n, m, k1, k2), "qds_gram: workspace too small");
  } else {
    const int64_t per_slab = (int64_t)g.mpad * g.kpad * (int64_t)sizeof(double);
    const int64_t cap = partial_bytes / per_slab;
    QDS_CHECK_ARG(cap >= kMinStaticSlabs,
                  "qds_gram_blocks_noalloc: a %d x %d Gram (padded %d x %d) is too large for the %lld-byte "
                  "static scratch: it needs %d row slabs of %lld bytes; padded m * (k1 + k2) must be at "
                  "most %lld",
                  m, k1 + k2, g.mpad, g.kpad, (long long)partial_bytes, kMinStaticSlabs, (long long)per_slab,
                  (long long)(partial_bytes / ((int64_t)kMinStaticSlabs * (int64_t)sizeof(double))));
    if (nblk > cap) nblk = (int)cap;
  }
  const int64_t rows_per_blk = (((n + nblk - 1) / nblk) + kRB - 1) / kRB * kRB;
  const size_t shm = (size_t)kRB * (g.mpad + 16 * tpg) * sizeof(double);
  for (int q = 0; q < stg.n; ++q)
    QDS_CHECK_ARG(stg.ld[q] * kRB < (int64_t(1) << 31), "qds_gram: leading dimension too large");
  hipStream_t st = qds_stream(stream);
  // a thread keeps offsets and one staged vector per piece: more than 9 pieces
  // spill, so wide operands run one launch per column group with only the pieces
  // that group reads (all of A, the B pieces inside the group's columns)
  auto launch = [&](auto tag, auto vecc, const GramStage& sg, int g0, int ny) {
    using T = decltype(tag);
    constexpr bool VV = decltype(vecc)::value;
    if (typed) {
      if (sg.n <= 9)
        hipLaunchKernelGGL((gram_partial<double, true, 9, 8, true>), dim3(nblk, ny, batch), dim3(kBlock), shm, st,
                           n, rows_per_blk, g, tpg, g0, sg, partial, s_stride);
      else
        hipLaunchKernelGGL((gram_partial<double, true, kGramPieces, 8, true>), dim3(nblk, ny, batch), dim3(kBlock),
                           shm, st, n, rows_per_blk, g, tpg, g0, sg, partial, s_stride);
    } else if (sg.n <= 9) {
      hipLaunchKernelGGL((gram_partial<T, VV, 9, 8>), dim3(nblk, ny, batch), dim3(kBlock), shm, st, n,
                         rows_per_blk, g, tpg, g0, sg, partial, s_stride);
    } else {
      hipLaunchKernelGGL((gram_partial<T, VV, kGramPieces, 8>), dim3(nblk, ny, batch), dim3(kBlock), shm, st, n,
                         rows_per_blk, g, tpg, g0, sg, partial, s_stride);
    }
  };
  auto launch_all = [&](auto tag, auto vecc) {
    if (stg.n <= 9 || ngroups == 1) {
      launch(tag, vecc, stg, 0, ngroups);
      return;
    }
    for (int gi = 0; gi < ngroups; ++gi) {
      const int c_lo = gi * tpg * 16, c_hi = (gi + 1 < ngroups ? (gi + 1) * tpg : tk) * 16;
      GramStage sg;
      sg.n = 0;
      for (int q = 0; q < stg.n; ++q) {
        if (stg.op[q] != 0) {
          const int t0 = (stg.op[q] == 1 ? 0 : k1pad) + stg.col0[q];
          if (t0 + stg.w[q] <= c_lo || t0 >= c_hi) continue;
        }
        sg.src[sg.n] = stg.src[q];
        sg.ld[sg.n] = stg.ld[q];
        sg.w[sg.n] = stg.w[q];
        sg.op[sg.n] = stg.op[q];
        sg.col0[sg.n] = stg.col0[q];
        sg.es[sg.n] = stg.es[q];
        ++sg.n;
      }
      launch(tag, vecc, sg, gi, 1);
    }
  };
  if (dtype == qds::F32 && vec) launch_all(float(0), std::true_type());
  else if (dtype == qds::F32) launch_all(float(0), std::false_type());
  else if (vec) launch_all(double(0), std::true_type());
  else launch_all(double(0), std::false_type());
  QDS_LAUNCH_CHECK();
  const int tot = m * (k1 + k2);
  hipLaunchKernelGGL(gram_reduce, dim3((tot + kRedOut - 1) / kRedOut, batch), dim3(kBlock), 0, st, nblk, g,
                     partial, C);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
}  // namespace

// The Gram of operands whose column blocks differ in storage type: a_dt, b1_dt,
// b2_dt give the type (F32 / F64) of each block of A, B1, B2.  Rows are converted
// to FP64 when staged; the products and sums are FP64 as in qds_gram_blocks (an
// all-F64 or all-F32 call gives the same result as qds_gram_blocks).  Workspace:
// qds_gram_workspace.
QDS_API int qds_gram_blocks_typed(int64_t n, const qds_colblocks* A, const int32_t* a_dt, const qds_colblocks* B1,
                                  const int32_t* b1_dt, const qds_colblocks* B2, const int32_t* b2_dt, int32_t sym,
                                  double* C, void* workspace, int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(workspace && a_dt && b1_dt && (!B2 || b2_dt), "qds_gram_blocks_typed: null pointer");
  const int32_t* dts[3] = {a_dt, b1_dt, b2_dt};
  return gram_blocks_impl(qds::F64, n, A, B1, B2, sym, C, static_cast<double*>(workspace), ws_bytes, 0, stream, 1,
                          0, dts);
}

// C (m x (k1+k2), row-major FP64) = A^T [B1 | B2] for column-block matrices.
// B2 may be NULL.  sym = 1 asserts each half is symmetric (m == k1 == k2) and
// computes only its upper tiles.  dtype F32 or F64 gives the storage type.
QDS_API int qds_gram_blocks(int32_t dtype, int64_t n, const qds_colblocks* A, const qds_colblocks* B1,
                            const qds_colblocks* B2, int32_t sym, double* C, void* workspace,
                            int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(workspace, "qds_gram: null workspace");
  return gram_blocks_impl(dtype, n, A, B1, B2, sym, C, static_cast<double*>(workspace), ws_bytes, 0, stream);
}

// ---- the same Gram without a caller workspace ---------------------------------
// For callers that cannot allocate (HIP graph capture, serving loops): the
// per-slab partial sums live in a file-scope device array, one copy per GPU,
// held until the process ends.  Its device address is looked up once per device
// with hipGetSymbolAddress and only that address is handed to kernels.  Calls
// that overlap on one GPU share the array and race.  Every partial that the
// reduction reads is written by the same call first, so nothing needs clearing.
namespace {
constexpr int64_t kStaticGramBytes = 64ll << 20;
__device__ __align__(256) double scratch[kStaticGramBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(double** out) {
  int dev = -1;
  QDS_HIP(hipGetDevice(&dev));
  QDS_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "qds_gram_blocks_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    QDS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QDS_CHECK_ARG(p != nullptr, "qds_gram_blocks_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return QDS_OK;
}
}  // namespace

QDS_API int64_t qds_gram_noalloc_scratch_bytes(void) { return kStaticGramBytes; }

// The Gram of `batch` problems in one launch: problem i's blocks of A, B1, B2 at
// p[j] + i s_stride (elements, one stride for every block), its output at
// C + i m (k1 + k2); workspace: batch * qds_gram_workspace(n, m, k1, k2).
QDS_API int qds_gram_blocks_batched(int32_t dtype, int64_t n, const qds_colblocks* A, const qds_colblocks* B1,
                                    const qds_colblocks* B2, int32_t sym, double* C, int32_t batch,
                                    int64_t s_stride, void* workspace, int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(workspace, "qds_gram: null workspace");
  const int es = dtype == qds::F32 ? 4 : 8;
  QDS_CHECK_ARG(batch == 1 || (s_stride * es) % 16 == 0, "qds_gram: batch stride must keep 16-byte alignment");
  return gram_blocks_impl(dtype, n, A, B1, B2, sym, C, static_cast<double*>(workspace), ws_bytes, 0, stream, batch,
                          s_stride);
}

QDS_API int qds_gram_blocks_noalloc(int32_t dtype, int64_t n, const qds_colblocks* A, const qds_colblocks* B1,
                                    const qds_colblocks* B2, int32_t sym, double* C, void* stream) {
  double* partial = nullptr;
  const int rc = static_scratch(&partial);
  if (rc) return rc;
  return gram_blocks_impl(dtype, n, A, B1, B2, sym, C, partial, 0, kStaticGramBytes, stream);
}

// Single-block form: A n x m (lda), B1 n x k1, B2 n x k2 (or NULL).
QDS_API int qds_gram(int32_t dtype, int64_t n, int32_t m, const void* A, int64_t lda, int32_t k1,
                     const void* B1, int64_t ldb1, int32_t k2, const void* B2, int64_t ldb2,
                     int32_t sym, double* C, void* workspace, int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(A && B1 && m > 0 && k1 > 0 && k2 >= 0 && (k2 == 0 || B2), "qds_gram: bad shape");
  qds_colblocks a = single(A, lda, m), b1 = single(B1, ldb1, k1), b2 = single(B2, ldb2, k2);
  return qds_gram_blocks(dtype, n, &a, &b1, k2 ? &b2 : nullptr, sym, C, workspace, ws_bytes, stream);
}

// Y1 = S C1 (n x k), and Y2 = S C2 when C2 != NULL; S given as column blocks.
// C1, C2: m x k row-major FP64 (rounded to the storage type on the device).  k <= 64.  Every block's width and
// leading dimension must be multiples of 4 (F32) or 2 (F64), blocks 16-byte aligned.
// c2_first: the caller guarantees rows [0, c2_first) of C2 are zero (0: none known).
QDS_API int qds_block_update_blocks_batched(int32_t dtype, int64_t n, const qds_colblocks* S, int32_t k,
                                            const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                                            int64_t ldy2, int32_t c2_first, int32_t batch, int64_t s_stride,
                                            int64_t c_stride, int64_t y_stride, void* stream);

QDS_API int qds_block_update_blocks(int32_t dtype, int64_t n, const qds_colblocks* S, int32_t k,
                                    const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                                    int64_t ldy2, int32_t c2_first, void* stream) {
  return qds_block_update_blocks_batched(dtype, n, S, k, C1, Y1, ldy1, C2, Y2, ldy2, c2_first, 1, 0, 0, 0, stream);
}

// The update for `batch` problems in one launch: problem i reads its S blocks at
// S.p[j] + i s_stride, C1 / C2 at + i c_stride and writes Y1 / Y2 at + i y_stride
// (strides in elements; the S blocks of all problems share one stride).
QDS_API int qds_block_update_blocks_batched(int32_t dtype, int64_t n, const qds_colblocks* S, int32_t k,
                                            const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                                            int64_t ldy2, int32_t c2_first, int32_t batch, int64_t s_stride,
                                            int64_t c_stride, int64_t y_stride, void* stream) {
  QDS_CHECK_ARG(batch >= 1 && batch <= 65535, "qds_block_update: batch %d out of range", batch);
  QDS_CHECK_ARG(dtype == qds::F32 || dtype == qds::F64, "qds_block_update: dtype must be F32 or F64");
  const int es = dtype == qds::F32 ? 4 : 8;
  QDS_CHECK_ARG(S && blocks_vec(S, es), "qds_block_update: S blocks need 16-byte alignment and widths, "
                "leading dimensions multiple of %d", 16 / es);
  ColBlocks dS;
  int m = 0;
  QDS_CHECK_ARG(to_device_blocks(S, dS, m, es, 16 * (16 / es)), "qds_block_update: bad or too wide S blocks");
  QDS_CHECK_ARG(C1 && Y1 && (!C2 || Y2), "qds_block_update: null pointer");
  QDS_CHECK_ARG(k >= 1 && k <= 16 * kMaxOutTiles, "qds_block_update: bad shape");
  QDS_CHECK_ARG(c2_first >= 0 && c2_first <= m, "qds_block_update: bad c2_first");
  if (n == 0) return QDS_OK;
  hipStream_t st = qds_stream(stream);
  if (!C2) c2_first = 0;
  const double *c1 = C1, *c2 = C2;
  if (batch > 1)
    QDS_CHECK_ARG((s_stride * es) % 16 == 0 && (y_stride * es) % 16 == 0,
                  "qds_block_update: batch strides must keep 16-byte alignment");
  if (dtype == qds::F32) {
    float *y1 = (float*)Y1, *y2 = (float*)Y2;
    return C2 ? launch_update<float, 2>(n, m, k, dS, c2_first, c1, c2, y1, ldy1, y2, ldy2, st, batch, s_stride,
                                        c_stride, y_stride)
              : launch_update<float, 1>(n, m, k, dS, 0, c1, nullptr, y1, ldy1, nullptr, ldy2, st, batch, s_stride,
                                        c_stride, y_stride);
  }
  double *y1 = (double*)Y1, *y2 = (double*)Y2;
  return C2 ? launch_update<double, 2>(n, m, k, dS, c2_first, c1, c2, y1, ldy1, y2, ldy2, st, batch, s_stride,
                                       c_stride, y_stride)
            : launch_update<double, 1>(n, m, k, dS, 0, c1, nullptr, y1, ldy1, nullptr, ldy2, st, batch, s_stride,
                                       c_stride, y_stride);
}

// Single-block form: S n x m with leading dimension lds.
QDS_API int qds_block_update(int32_t dtype, int64_t n, int32_t m, int32_t k, const void* S, int64_t lds,
                             const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                             int64_t ldy2, void* stream) {
  QDS_CHECK_ARG(S && m >= 1, "qds_block_update: bad S");
  qds_colblocks s = single(S, lds, m);
  return qds_block_update_blocks(dtype, n, &s, k, C1, Y1, ldy1, C2, Y2, ldy2, 0, stream);
}

// Y1 = S C1 (FP64) and, when C2 != NULL, Y2 = S C2 (stored in y2_dtype, F32 or
// F64) for column blocks whose storage types differ (s_dt[j] for block j, F32 or
// F64): the update of the mixed-precision iteration (X_new in FP64, P_new in FP32
// from [X | W | P]).  The FP64 blocks' products and sums are FP64; the FP32 blocks'
// run on the FP32 MFMA with C rounded to FP32 and are added in FP64.  C resident in
// LDS (k <= 64); blocks 16-byte aligned with widths and leading dimensions a
// multiple of the vector width (4 F32, 2 F64), at most 32 (F64) / 64 (F32) columns
// each.  c2_first: rows of C2 before it are zero.
QDS_API int qds_block_update_typed(int64_t n, const qds_colblocks* S, const int32_t* s_dt, int32_t k,
                                   const double* C1, void* Y1, int64_t ldy1, const double* C2, int32_t y2_dtype,
                                   void* Y2, int64_t ldy2, int32_t c2_first, void* stream) {
  QDS_CHECK_ARG(S && s_dt && S->nb >= 1 && S->nb <= 4, "qds_block_update_typed: bad S");
  QDS_CHECK_ARG(C1 && Y1 && (!C2 || Y2), "qds_block_update_typed: null pointer");
  QDS_CHECK_ARG(!C2 || y2_dtype == qds::F32 || y2_dtype == qds::F64, "qds_block_update_typed: bad y2_dtype");
  QDS_CHECK_ARG(k >= 1 && k <= 16 * kMaxOutTiles, "qds_block_update_typed: bad shape");
  TypedBlocks d;
  int m = 0, maxj = 1, w64 = 0, w32 = 0;
  for (int j = 0; j < 4; ++j) { d.p[j] = nullptr; d.ld[j] = 0; d.es[j] = 8; d.cbase[j] = 0; }
  for (int j = 0; j < S->nb; ++j) {
    QDS_CHECK_ARG(s_dt[j] == qds::F32 || s_dt[j] == qds::F64, "qds_block_update_typed: block %d type", j);
    const int es = qds::dtype_size(s_dt[j]);
    const int vw = 16 / es;
    QDS_CHECK_ARG(S->p[j] && S->cols[j] >= 1 && S->cols[j] <= 16 * vw && S->ld[j] >= S->cols[j] &&
                      (reinterpret_cast<uintptr_t>(S->p[j]) % 16) == 0 && (S->ld[j] * es) % 16 == 0 &&
                      S->cols[j] % vw == 0,
                  "qds_block_update_typed: block %d needs 16-byte alignment, width and leading dimension a "
                  "multiple of %d, at most %d columns", j, vw, 16 * vw);
    d.p[j] = S->p[j];
    d.ld[j] = S->ld[j];
    d.es[j] = es;
    d.start[j] = m;
    d.cbase[j] = es == 8 ? w64 : w32;
    (es == 8 ? w64 : w32) += S->cols[j];
    m += S->cols[j];
    const int loads = (S->cols[j] + 4 * vw - 1) / (4 * vw);
    maxj = loads > maxj ? loads : maxj;
  }
  for (int j = S->nb; j <= 4; ++j) d.start[j] = m;
  d.nb = S->nb;
  QDS_CHECK_ARG(c2_first >= 0 && c2_first <= m, "qds_block_update_typed: bad c2_first");
  const int ot = (k + 15) / 16;
  const int kp = ot * 16;
  // each region padded by one load of rows (the last load of a block may run past it)
  const int rows64 = w64 ? w64 + 8 : 0, rows32 = w32 ? w32 + 16 : 0;
  const int nout = C2 ? 2 : 1;
  const size_t shm = (size_t)nout * kp * ((size_t)rows64 * sizeof(double) + (size_t)rows32 * sizeof(float));
  QDS_CHECK_ARG(shm <= (size_t)kUpdLds, "qds_block_update_typed: C of %d x %d too large for LDS", m, k);
  if (n == 0) return QDS_OK;
  hipStream_t st = qds_stream(stream);
  const int64_t passes = (n + kWaves * 16 - 1) / (kWaves * 16);
  const unsigned blocks = qds_blocks(passes, 1, 4096);
  QDS_CHECK_ARG(maxj <= 4, "qds_block_update_typed: blocks too wide");
#define QDS_UPDT(OTV, J, TY, TW)                                                                            \
  hipLaunchKernelGGL((block_update_typed_kernel<OTV, J, TY, TW>), dim3(blocks), dim3(kBlock), shm, st, n, m, k, d, \
                     C1, C2, (double*)Y1, ldy1, (TY*)Y2, ldy2, rows64, rows32, C2 ? c2_first : 0)
#define QDS_UPDT_Y(OTV, J)                                                                                  \
  do {                                                                                                       \
    if (!C2) QDS_UPDT(OTV, J, double, false);                                                                \
    else if (y2_dtype == qds::F32) QDS_UPDT(OTV, J, float, true);                                            \
    else QDS_UPDT(OTV, J, double, true);                                                                     \
  } while (0)
#define QDS_UPDT_J(OTV) \
  do { if (maxj <= 2) QDS_UPDT_Y(OTV, 2); else if (maxj == 3) QDS_UPDT_Y(OTV, 3); else QDS_UPDT_Y(OTV, 4); } while (0)
  switch (ot) {
    case 1: QDS_UPDT_J(1); break;
    case 2: QDS_UPDT_J(2); break;
    case 3: QDS_UPDT_J(3); break;
    default: QDS_UPDT_J(4); break;
  }
#undef QDS_UPDT_J
#undef QDS_UPDT_Y
#undef QDS_UPDT
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// R = HX - MX diag(lambda).  Real dtypes: lambda has b doubles.  Complex dtypes:
// lambda has b interleaved complex doubles, products without conjugation.
QDS_API int qds_block_residual(int32_t dtype, int64_t n, int32_t b, const void* HX, int64_t ldh,
                               const void* MX, int64_t ldm, const double* lam, void* R, int64_t ldr,
                               void* stream) {
  QDS_CHECK_ARG(HX && MX && lam && R && b > 0, "qds_block_residual: bad arguments");
  if (n == 0) return QDS_OK;
  hipStream_t st = qds_stream(stream);
  const dim3 g(qds_blocks(n * b, 256, 131072)), blk(256);
  switch (dtype) {
    case qds::F32: hipLaunchKernelGGL(residual_kernel<float>, g, blk, 0, st, n, b, (const float*)HX, ldh, (const float*)MX, ldm, lam, (float*)R, ldr); break;
    case qds::F64: hipLaunchKernelGGL(residual_kernel<double>, g, blk, 0, st, n, b, (const double*)HX, ldh, (const double*)MX, ldm, lam, (double*)R, ldr); break;
    case qds::C64: hipLaunchKernelGGL(residual_cplx_kernel<float>, g, blk, 0, st, n, b, (const float*)HX, ldh, (const float*)MX, ldm, lam, (float*)R, ldr); break;
    case qds::C128: hipLaunchKernelGGL(residual_cplx_kernel<double>, g, blk, 0, st, n, b, (const double*)HX, ldh, (const double*)MX, ldm, lam, (double*)R, ldr); break;
    default: QDS_CHECK_ARG(false, "qds_block_residual: bad dtype");
  }
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

QDS_API int64_t qds_coldot_workspace(int32_t b) { return (int64_t)kColdotBlocks * 2 * b * sizeof(double); }

// out[j] = sum_r X[r, j] * Y[r, j] in FP64.  Real dtypes: out has b doubles.
// Complex dtypes: out has b complex doubles; conj = 1 conjugates X.
QDS_API int qds_coldot(int32_t dtype, int64_t n, int32_t b, const void* X, int64_t ldx, const void* Y,
                       int64_t ldy, int32_t conj, double* out, void* workspace, int64_t ws_bytes,
                       void* stream) {
  QDS_CHECK_ARG(X && Y && out && workspace && b > 0 && b <= 256, "qds_coldot: bad arguments");
  QDS_CHECK_ARG(ws_bytes >= qds_coldot_workspace(b), "qds_coldot: workspace too small");
  hipStream_t st = qds_stream(stream);
  int64_t nblk = (n + 255) / 256;
  if (nblk > kColdotBlocks) nblk = kColdotBlocks;
  if (nblk < 1) nblk = 1;
  const int64_t rpb = (n + nblk - 1) / nblk;
  double* part = static_cast<double*>(workspace);
  const bool cplx = dtype == qds::C64 || dtype == qds::C128;
  switch (dtype) {
    case qds::F32: hipLaunchKernelGGL((coldot_partial<float, 0>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const float*)X, ldx, (const float*)Y, ldy, rpb, part); break;
    case qds::F64: hipLaunchKernelGGL((coldot_partial<double, 0>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const double*)X, ldx, (const double*)Y, ldy, rpb, part); break;
    case qds::C64:
      if (conj) hipLaunchKernelGGL((coldot_partial<float, 1>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const float*)X, ldx, (const float*)Y, ldy, rpb, part);
      else hipLaunchKernelGGL((coldot_partial<float, 2>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const float*)X, ldx, (const float*)Y, ldy, rpb, part);
      break;
    case qds::C128:
      if (conj) hipLaunchKernelGGL((coldot_partial<double, 1>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const double*)X, ldx, (const double*)Y, ldy, rpb, part);
      else hipLaunchKernelGGL((coldot_partial<double, 2>), dim3(nblk), dim3(kBlock), 0, st, n, b, (const double*)X, ldx, (const double*)Y, ldy, rpb, part);
      break;
    default: QDS_CHECK_ARG(false, "qds_coldot: bad dtype");
  }
  QDS_LAUNCH_CHECK();
  hipLaunchKernelGGL(coldot_reduce, dim3(1), dim3(1024), 0, st, (int)nblk, b, part, out, cplx ? 1 : 0);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// ---------------------------------------------------------------------------
// Electron density of the self-consistent loop.
// Reference: fusedElectronDensity (backend/src/fused_gpu_kernels.cu:127-157, unused)
// and the per-node loop of self_consistent_solver.pyx:262-274:
//   n[i] = sum_s occ[s] * |psi[i, s]|^2
namespace {
__global__ void density_kernel(int64_t n, int k, const double* __restrict__ psi, int64_t ld,
                               const double* __restrict__ occ, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) {
      const double v = psi[i * ld + j];
      s += occ[j] * v * v;
    }
    out[i] = s;
  }
}
}  // namespace

QDS_API int qds_density(int64_t n, int32_t k, const double* psi, int64_t ld, const double* occ,
                        double* out, void* stream) {
  QDS_CHECK_ARG(psi && occ && out && k >= 0 && ld >= k, "qds_density: bad arguments");
  if (n == 0) return QDS_OK;
  hipLaunchKernelGGL(density_kernel, dim3(qds_blocks(n, 256, 65536)), dim3(256), 0, qds_stream(stream),
                     n, k, psi, ld, occ, out);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}

// ---------------------------------------------------------------------------
// Fused residual and column norms for the eigensolver loop:
//   R = HX - MX diag(lambda),  out[j] = ||R_j||^2,  out[b + j] = ||MX_j||^2  (FP64)
// Real dtypes only.  Threads of a block form groups of b lanes; lane c of every
// group keeps the running sums of column c over the rows of its group.
namespace {
constexpr int kResBlocks = 2048;

// R is stored in TR: the storage type of the images, or FP32 from FP64 images (the
// mixed-precision iteration keeps its corrections in FP32).
template <typename T, typename TR>
__global__ void __launch_bounds__(kBlock)
residual_norms_partial(int64_t n, int b, const T* __restrict__ HX, int64_t ldh, const T* __restrict__ MX,
                       int64_t ldm, const double* __restrict__ lam, const double* __restrict__ colscale,
                       TR* __restrict__ R, int64_t ldr, int64_t rows_per_blk, double* __restrict__ partial,
                       int64_t sx) {
  // problem of a batch (blockIdx.y): blocks at + set * sx, its own lam, colscale, partials
  {
    const int64_t set = blockIdx.y;
    HX += set * sx;
    MX += set * sx;
    R += set * sx;
    lam += set * b;
    if (colscale) colscale += set * b;
    partial += set * (int64_t)gridDim.x * 2 * b;
  }
  const int groups = kBlock / b;
  const int c = threadIdx.x % b;
  const int gidx = threadIdx.x / b;
  __shared__ double red[2 * kBlock];
  double sr = 0.0, sm = 0.0;
  const int64_t r_begin = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t r_end = min(n, r_begin + rows_per_blk);
  if (gidx < groups) {
    const T l = (T)lam[c];
    const T cs = colscale ? (T)colscale[c] : T(1);
    for (int64_t r = r_begin + gidx; r < r_end; r += groups) {
      const T m = MX[r * ldm + c];
      const T v = HX[r * ldh + c] - l * m;
      R[r * ldr + c] = (TR)(v * cs);
      sr += (double)v * (double)v;
      sm += (double)m * (double)m;
    }
  }
  red[threadIdx.x] = sr;
  red[kBlock + threadIdx.x] = sm;
  __syncthreads();
  if (threadIdx.x < b) {
    double a0 = 0.0, a1 = 0.0;
    for (int g = 0; g < groups; ++g) {
      a0 += red[g * b + threadIdx.x];
      a1 += red[kBlock + g * b + threadIdx.x];
    }
    partial[(int64_t)blockIdx.x * 2 * b + threadIdx.x] = a0;
    partial[(int64_t)blockIdx.x * 2 * b + b + threadIdx.x] = a1;
  }
}

// One block of 1024 threads: thread t sums column t % 2b over partial rows
// t / 2b, t / 2b + groups, ...; the group sums are then combined in a fixed order.
__global__ void __launch_bounds__(1024)
residual_norms_reduce(int nblk, int b, const double* __restrict__ partial, double* __restrict__ out) {
  partial += (int64_t)blockIdx.x * nblk * 2 * b;
  out += (int64_t)blockIdx.x * 2 * b;
  __shared__ double red[1024];
  const int w = 2 * b;
  const int groups = 1024 / w;
  const int c = threadIdx.x % w;
  const int g = threadIdx.x / w;
  double a = 0.0;
  if (g < groups)
    for (int k = g; k < nblk; k += groups) a += partial[(int64_t)k * w + c];
  red[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x < w) {
    double s = 0.0;
    for (int q = 0; q < groups; ++q) s += red[q * w + threadIdx.x];
    out[threadIdx.x] = s;
  }
}
}  // namespace

QDS_API int64_t qds_residual_norms_workspace(int32_t b) {
  return (int64_t)kResBlocks * 2 * b * sizeof(double);
}

QDS_API int64_t qds_residual_norms_batched_workspace(int64_t n, int32_t b, int32_t batch) {
  int64_t nblk = (n + 511) / 512;
  if (nblk > kResBlocks) nblk = kResBlocks;
  if (nblk < 1) nblk = 1;
  return (int64_t)batch * nblk * 2 * b * sizeof(double);
}

// The residual step for `batch` problems in one launch: problem i's HX, MX, R at
// + i x_stride (elements), lam / colscale at + i b, out at + i 2b.
namespace {
int residual_norms_impl(int32_t dtype, int32_t r_dtype, int64_t n, int32_t b, int32_t batch, const void* HX,
                        int64_t ldh, const void* MX, int64_t ldm, int64_t x_stride, const double* lam,
                        const double* colscale, void* R, int64_t ldr, double* out, void* workspace,
                        int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(HX && MX && lam && R && out && workspace && b >= 1 && b <= 128 && batch >= 1 && batch <= 65535,
                "qds_residual_norms: bad arguments");
  QDS_CHECK_ARG(r_dtype == dtype || (dtype == qds::F64 && r_dtype == qds::F32 && batch == 1),
                "qds_residual_norms: R must be stored in the images' type, or in F32 from F64 images");
  int64_t nblk = (n + 511) / 512;
  if (nblk > kResBlocks) nblk = kResBlocks;
  if (nblk < 1) nblk = 1;
  QDS_CHECK_ARG(ws_bytes >= (int64_t)batch * nblk * 2 * b * (int64_t)sizeof(double),
                "qds_residual_norms: workspace too small");
  hipStream_t st = qds_stream(stream);
  const int64_t rpb = (n + nblk - 1) / nblk;
  double* part = static_cast<double*>(workspace);
  if (dtype == qds::F32)
    hipLaunchKernelGGL((residual_norms_partial<float, float>), dim3(nblk, batch), dim3(kBlock), 0, st, n, b,
                       (const float*)HX, ldh, (const float*)MX, ldm, lam, colscale, (float*)R, ldr, rpb, part,
                       x_stride);
  else if (dtype == qds::F64 && r_dtype == qds::F64)
    hipLaunchKernelGGL((residual_norms_partial<double, double>), dim3(nblk, batch), dim3(kBlock), 0, st, n, b,
                       (const double*)HX, ldh, (const double*)MX, ldm, lam, colscale, (double*)R, ldr, rpb, part,
                       x_stride);
  else if (dtype == qds::F64)
    hipLaunchKernelGGL((residual_norms_partial<double, float>), dim3(nblk, batch), dim3(kBlock), 0, st, n, b,
                       (const double*)HX, ldh, (const double*)MX, ldm, lam, colscale, (float*)R, ldr, rpb, part,
                       x_stride);
  else
    QDS_CHECK_ARG(false, "qds_residual_norms: dtype must be F32 or F64");
  QDS_LAUNCH_CHECK();
  hipLaunchKernelGGL(residual_norms_reduce, dim3(batch), dim3(1024), 0, st, (int)nblk, b, part, out);
  QDS_LAUNCH_CHECK();
  return QDS_OK;
}
}  // namespace

QDS_API int qds_residual_norms_batched(int32_t dtype, int64_t n, int32_t b, int32_t batch, const void* HX,
                                       int64_t ldh, const void* MX, int64_t ldm, int64_t x_stride,
                                       const double* lam, const double* colscale, void* R, int64_t ldr, double* out,
                                       void* workspace, int64_t ws_bytes, void* stream) {
  return residual_norms_impl(dtype, dtype, n, b, batch, HX, ldh, MX, ldm, x_stride, lam, colscale, R, ldr, out,
                             workspace, ws_bytes, stream);
}

// The residual step of the mixed-precision iteration: FP64 images, R stored in
// r_dtype (F32 or F64); the norms are FP64 sums of the FP64 residual.
QDS_API int qds_residual_norms_to(int32_t dtype, int32_t r_dtype, int64_t n, int32_t b, const void* HX,
                                  int64_t ldh, const void* MX, int64_t ldm, const double* lam,
                                  const double* colscale, void* R, int64_t ldr, double* out, void* workspace,
                                  int64_t ws_bytes, void* stream) {
  QDS_CHECK_ARG(ws_bytes >= qds_residual_norms_workspace(b), "qds_residual_norms: workspace too small");
  return residual_norms_impl(dtype, r_dtype, n, b, 1, HX, ldh, MX, ldm, 0, lam, colscale, R, ldr, out, workspace,
                             ws_bytes, stream);
}

// colscale (optional, b doubles) multiplies the stored residual columns after the
// norms are taken: the eigensolver zeroes the columns of converged pairs (soft locking).
QDS_API int qds_residual_norms(int32_t dtype, int64_t n, int32_t b, const void* HX, int64_t ldh,
                               const void* MX, int64_t ldm, const double* lam, const double* colscale,
                               void* R, int64_t ldr, double* out, void* workspace, int64_t ws_bytes,
                               void* stream) {
  QDS_CHECK_ARG(ws_bytes >= qds_residual_norms_workspace(b), "qds_residual_norms: workspace too small");
  return qds_residual_norms_batched(dtype, n, b, 1, HX, ldh, MX, ldm, 0, lam, colscale, R, ldr, out, workspace,
                                    ws_bytes, stream);
}
