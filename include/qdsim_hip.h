/*
 * qdsim_hip -- C-ABI of the MI355X (gfx950) FEM-Hamiltonian path.
 *
 * Conventions
 *   - Every call returns 0 on success or a negative code (QDS_E*); the message of
 *     the last failure on the calling thread is qds_last_error().
 *   - Device pointers are plain HBM addresses (e.g. torch tensor data_ptr()); host
 *     pointers are marked "host".  `stream` is a hipStream_t passed as void*.
 *   - Calls never allocate device memory.  Entry points that need scratch take a
 *     `workspace` pointer and size; the matching *_workspace() call gives the size.
 *   - dtype codes: 1 = FP32, 2 = FP64, 3 = complex64, 4 = complex128
 *     (complex values interleaved re, im).
 *   - Blocks are row-major n x b arrays with a leading dimension (elements).
 *
 * The reference binds its FEM path through pybind11 (backend/src/bindings.cpp:948-985)
 * and scipy; the Cython solvers call scipy.sparse / scipy.sparse.linalg directly.
 * Each entry point below names the reference operation it replaces.
 */
#ifndef QDSIM_HIP_H
#define QDSIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QDS_OK 0
#define QDS_EINVAL (-1)
#define QDS_EHIP (-2)
#define QDS_ENOMEM (-3)
#define QDS_EWORKSPACE (-4)

/* ---- library --------------------------------------------------------------- */
const char* qds_last_error(void);
int qds_version(void);
int qds_device_info(char* arch /* host */, int32_t arch_len, int32_t* cu_count /* host */);
/* No counterpart in the reference.  For callers that load the library themselves
 * (dlopen, ctypes) and could end up with two HIP runtimes in one process: launches a
 * probe kernel through the library's runtime and reads it back; QDS_EHIP, with the
 * mapped libamdhip64 images named in qds_last_error(), when that fails.  Allocates a
 * 4-byte device buffer.  qds_runtime_images: number of libamdhip64 images mapped. */
int qds_runtime_check(void);
int qds_runtime_images(void);

/* ---- CSR pattern and scatter maps (host) -------------------------------------
 * Replaces the COO->CSR conversion of the element triplets:
 *   qdsim_cython/solvers/fem_solver.pyx:142-145 (sp.csr_matrix((data,(row,col))))
 *   qdsim_cython/solvers/fixed_open_system_solver.pyx:259-267
 *   backend/src/fem.cpp:148-150 (setFromTriplets)
 * conn: host int32 [n_elem][npe].  seg_src entries index packed symmetric element
 * matrices: e*nsym + s(min(i,j), max(i,j)), s(i,j) = i*npe - i*(i-1)/2 + (j-i).  */
int qds_pattern_count(const int32_t* conn, int64_t n_elem, int32_t npe, int64_t n_dof,
                      int64_t* nnz_out);
int qds_pattern_fill(const int32_t* conn, int64_t n_elem, int32_t npe, int64_t n_dof, int64_t nnz,
                     int32_t* indptr, int32_t* indices, int32_t* seg_ptr, int32_t* seg_src,
                     int32_t* elem_map /* optional [n_elem*npe*npe] */);
/* Submatrix on renumbered DOFs (new_of_old[old] = new id or -1).  Replaces the
 * Dirichlet row replacement via LIL of schrodinger_solver.pyx:253-283 by a
 * restriction to the free DOFs.  All arrays host.                               */
int qds_pattern_restrict_count(const int32_t* indptr, const int32_t* indices, int64_t n_dof,
                               const int32_t* new_of_old, int64_t n_new, int64_t* nnz_out);
int qds_pattern_restrict_fill(const int32_t* indptr, const int32_t* indices, int64_t n_dof,
                              const int32_t* new_of_old, int64_t n_new, int64_t nnz_r,
                              int32_t* r_indptr, int32_t* r_indices, int32_t* r_src);

/* ---- assembly (device) ---------------------------------------------------------
 * Element matrices: replaces FEMSolver::assemble_element_matrix
 * (backend/src/fem.cpp:736-989) and the P1 element formulas of
 * fem_solver.pyx:277-320 / schrodinger_solver.pyx:201-251.
 * tables = [w(nq) | phi(nq*npe) | dphi_dxi(nq*npe) | dphi_deta(nq*npe)].
 * Coefficient c/v/r/w: pointer, element stride, quadrature stride (0: per element).
 * out0 = K(c) + M(v), out1 = M(r) (r NULL: 1), out2 = M(w); each [n_elem*nsym].  */
int qds_assemble_elements(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                          const double* x, const double* y, const double* tables,
                          const double* c_q, int64_t c_se, int32_t c_sq,
                          const double* v_q, int64_t v_se, int32_t v_sq,
                          const double* r_q, int64_t r_se, int32_t r_sq,
                          const double* w_q, int64_t w_se, int32_t w_sq,
                          double* out0, double* out1, double* out2, void* stream);
/* CSR values = segmented sums of packed element entries (duplicate summation of
 * the COO->CSR conversion, fem_solver.pyx:142-145), deterministic order.          */
int qds_assemble_gather(int64_t nnz, const int32_t* seg_ptr, const int32_t* seg_src,
                        const double* e0, const double* e1, const double* e2,
                        double* v0, double* v1, double* v2, void* stream);
/* Physical quadrature points (coefficient sampling; replaces the per-element
 * centroid callbacks of fem_solver.pyx:126-128 and per-point calls of
 * fem.cpp:926-943).  bary: [nq][3].  xq, yq: [n_elem][nq].                        */
int qds_quad_points(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn, const double* x,
                    const double* y, const double* bary, double* xq, double* yq, void* stream);
/* out[k] = re[src[k]] * scale (real dtypes) or re + i*scale*im (complex dtypes);
 * src[k] < 0 gives 0.                                                              */
int qds_csr_take(int32_t dtype, int64_t n, const int32_t* src, const double* re, const double* im,
                 double scale, void* out, void* stream);

/* FE interpolation of a nodal field at the quadrature points (replaces the
 * nearest-node lookup of self_consistent_solver.pyx:119-128,225-236):
 * out[e][q] = sum_a phi[q][a] u[conn[e][a]].                                      */
int qds_interpolate_q(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                      const double* phi, const double* u, double* out, void* stream);
/* FE field at arbitrary points of a structured order-p mesh (cx x cy cells of
 * dx x dy; element 2 (j cx + i) + t).  Replaces FEInterpolator::interpolate /
 * interpolateWithGradient / findElement (backend/src/fe_interpolator.cpp:54-410)
 * and interpolate_field_kernel (backend/src/gpu_kernels.cu:352-423): same element
 * choice (lowest index whose barycentrics are >= -1e-10), nearest-node value and
 * inverse-distance gradient outside the mesh.  lattice_node[(p cy + 1)(p cx + 1)]:
 * lowest node id at each refined-lattice position.  grad_x/grad_y (both or
 * neither) and elem (-1 outside) are optional.                                    */
int qds_interpolate_points(int32_t order, int64_t cx, int64_t cy, double dx, double dy,
                           const int32_t* conn, const double* x, const double* y,
                           const int32_t* lattice_node,
                           const double* field /* [node][nfield] */, int32_t nfield, int64_t n_points,
                           const double* px, const double* py, double* val /* [point][nfield] */,
                           double* grad_x, double* grad_y, int32_t* elem, void* stream);
/* Hellmann-Feynman weights out[e][q][s] = w_q |J_e|/2 psi_s(x_q)^2 (no reference
 * counterpart: adjoint of the eigenvalues w.r.t. the potential at x_q).           */
int qds_state_density_q(int32_t npe, int32_t nq, int64_t n_elem, const int32_t* conn,
                        const double* x, const double* y, const double* tables, const double* psi,
                        int64_t ldpsi, int32_t k, double* out, void* stream);
/* Electron density n[i] = sum_s occ[s] psi[i][s]^2 (fusedElectronDensity,
 * backend/src/fused_gpu_kernels.cu:127-157; self_consistent_solver.pyx:262-274).  */
int qds_density(int64_t n, int32_t k, const double* psi, int64_t ld, const double* occ, double* out,
                void* stream);

/* ---- sparse products (device) --------------------------------------------------
 * Y_a = A_a X (and Y_b = A_b X on the same pattern).  Replaces scipy csr @ x inside
 * ARPACK (eigen_solver.pyx:225-235) and the unused fusedSpMV
 * (backend/src/fused_gpu_kernels.cu:97-125).  Complex products do not conjugate.  */
int qds_spmm(int32_t dtype, int64_t n_rows, int64_t nnz, int32_t b, const int32_t* indptr,
             const int32_t* indices, const void* vals_a, const void* vals_b, const void* X,
             int64_t ldx, void* Y_a, int64_t ldy_a, void* Y_b, int64_t ldy_b, void* stream);

/* CAP Hamiltonian on complex blocks: Y_h = (H - iW) X, Y_m = M X with real value
 * arrays h, w, m on one pattern (dtype C64: FP32 values; C128: FP64).  Replaces
 * the complex V + i eta of backend/src/fem.cpp:942,974 and the constant lifetime
 * shift of fixed_open_system_solver.pyx:321-328.                                  */
int qds_spmm_cap(int32_t dtype, int64_t n_rows, int32_t b, const int32_t* indptr,
                 const int32_t* indices, const void* vals_h, const void* vals_w, const void* vals_m,
                 const void* X, int64_t ldx, void* Y_h, int64_t ldy_h, void* Y_m, int64_t ldy_m,
                 void* stream);

/* ---- lattice (stencil) form of the free-DOF operator -----------------------------
 * The eigensolver's operator lives on the interior points of the p-times refined
 * lattice of a structured order-p mesh (nx x ny points, row-major).  Every row of
 * class (I mod p, J mod p) (I = i + 1, J = j + 1: full-lattice coordinates) couples
 * to the same set of lattice offsets (dI, dJ), |dI|, |dJ| <= p, so the operator is
 * stored as value arrays without column indices.  Cells (I / p, J / p) are grouped
 * in tiles of `tile` cells along x; the values of one tile are contiguous
 * (tile_slots entries), class after class, and within class c in groups of four
 * entries, cell after cell:
 *   value of entry k of row (I, J), class c, cell (X, Y) = (I / p, J / p):
 *     vals[(Y * tiles_x + X / tile) * tile_slots + base[c]
 *          + ((k / 4) * tile + X % tile) * 4 + k % 4]
 * entries ordered by ascending column (dJ, then dI); missing entries at the
 * boundary, the padding of len[c] to a multiple of 4 and the cells past the last
 * column hold 0.  Same matrix as the CSR form of qds_pattern_restrict_fill (scipy
 * csr of the reference, eigen_solver.pyx:225-235), products bitwise equal to
 * qds_spmm's.                                                                   */
#define QDS_LATTICE_MAX_CLASS 9
#define QDS_LATTICE_MAX_ENTRY 52
typedef struct {
  int32_t order;                 /* p (1..3) */
  int32_t n_class;               /* p * p */
  int64_t nx, ny;                /* interior lattice points per row / column */
  int64_t cells_x, cells_y;      /* nx / p + 1, ny / p + 1 */
  int32_t tile;                  /* cells per tile along x */
  int32_t pad0;
  int64_t tiles_x;               /* ceil(cells_x / tile) */
  int64_t tile_slots;            /* tile * sum(len) */
  int64_t n_slots;               /* length of each value array: cells_y * tiles_x * tile_slots */
  int32_t len[QDS_LATTICE_MAX_CLASS];   /* entries per row of class c, multiple of 4 */
  int64_t base[QDS_LATTICE_MAX_CLASS];  /* offset of class c within a tile: tile * sum(len[<c]) */
  int8_t di[QDS_LATTICE_MAX_CLASS][QDS_LATTICE_MAX_ENTRY];
  int8_t dj[QDS_LATTICE_MAX_CLASS][QDS_LATTICE_MAX_ENTRY];
} qds_lattice_stencil;

/* Host: stencil of a lattice-ordered CSR pattern (rows = nx * ny interior points)
 * and, when slot_src != NULL, slot_src[slot] = CSR position of each value slot or
 * -1 (a qds_csr_take map after composing with the CSR value map).  Fails when a
 * column is not within p lattice steps of its row.                              */
int qds_lattice_stencil_build(int32_t order, int64_t nx, int64_t ny, const int32_t* indptr,
                              const int32_t* indices, qds_lattice_stencil* st /* host */,
                              int64_t* slot_src /* host, optional [st->n_slots] */);

/* Y_a = A_a X (and Y_b = A_b X) for operators in the lattice form: the product of
 * qds_spmm without index loads; X rows of a lattice tile and its p-point halo are
 * staged in LDS once per workgroup.  dtype F32 or F64.  Replaces scipy csr @ X in
 * the block eigensolvers (eigen_solver.pyx:225-235, 254-266).                     */
int qds_lattice_spmm(int32_t dtype, const qds_lattice_stencil* st /* host */, int32_t b,
                     const void* vals_a, const void* vals_b, const void* X, int64_t ldx,
                     void* Y_a, int64_t ldy_a, void* Y_b, int64_t ldy_b, void* stream);

/* The product for `batch` problems on one stencil in one launch (no reference
 * counterpart: the reference solves one parameter set per call): problem i reads
 * vals_a + i sva, vals_b + i svb (svb = 0: one shared matrix such as M),
 * X + i sx and writes Y_a + i sya, Y_b + i syb (strides in elements, each keeping
 * 16-byte alignment).                                                            */
int qds_lattice_spmm_batched(int32_t dtype, const qds_lattice_stencil* st /* host */, int32_t b, int32_t batch,
                             const void* vals_a, int64_t sva, const void* vals_b, int64_t svb,
                             const void* X, int64_t ldx, int64_t sx, void* Y_a, int64_t ldy_a, int64_t sya,
                             void* Y_b, int64_t ldy_b, int64_t syb, void* stream);

/* Block shape of the lattice product (no reference counterpart), chosen per
 * operator by the package's autotuner: values_in_lds = 1 stages each block's value
 * runs in LDS with its X rows (the default), 0 reads them from HBM in the compute
 * phase; lds_budget (4096 .. 65536 bytes) sizes the block (cells per workgroup),
 * and with it how many workgroups share a CU.                                    */
typedef struct {
  int32_t values_in_lds;
  int32_t lds_budget;
} qds_lattice_tuning;
int qds_lattice_spmm_tuned(int32_t dtype, const qds_lattice_stencil* st /* host */, int32_t b, int32_t batch,
                           const void* vals_a, int64_t sva, const void* vals_b, int64_t svb,
                           const void* X, int64_t ldx, int64_t sx, void* Y_a, int64_t ldy_a, int64_t sya,
                           void* Y_b, int64_t ldy_b, int64_t syb, const qds_lattice_tuning* tune /* host */,
                           void* stream);

/* The fused product of qds_lattice_spmm_tuned (batch 1, both matrices) with the
 * block iteration's residual step in its epilogue: R = (Y_a - Y_b diag(lam))
 * diag(colscale) (colscale may be NULL), stored as r_dtype (F32, or the product's
 * type), and norms (device, 2b FP64) = the column sums of (Y_a - Y_b diag(lam))^2,
 * then of Y_b^2 -- what qds_residual_norms_to computes from the stored products
 * (same R values; the sums in another fixed order), without reading Y_a, Y_b back.
 * Replaces the residual / convergence step of scipy's lobpcg
 * (eigen_solver.pyx:254-266) fused into the operator apply.  b <= 128; workspace:
 * qds_lattice_spmm_residual_workspace bytes.                                      */
int64_t qds_lattice_spmm_residual_workspace(const qds_lattice_stencil* st /* host */, int32_t b);
int qds_lattice_spmm_residual(int32_t dtype, const qds_lattice_stencil* st /* host */, int32_t b,
                              const void* vals_a, const void* vals_b, const void* X, int64_t ldx, void* Y_a,
                              int64_t ldy_a, void* Y_b, int64_t ldy_b, const double* lam, const double* colscale,
                              int32_t r_dtype, void* R, int64_t ldr, double* norms, void* workspace,
                              int64_t ws_bytes, const qds_lattice_tuning* tune /* host */, void* stream);

/* ---- dense block kernels for the block eigensolver (device) --------------------
 * Replace the dense Gram / projection steps inside scipy's lobpcg
 * (eigen_solver.pyx:254-266) and ARPACK's orthogonalisation.                      */
int64_t qds_gram_workspace(int64_t n, int32_t m, int32_t k1, int32_t k2);
int qds_gram(int32_t dtype, int64_t n, int32_t m, const void* A, int64_t lda, int32_t k1,
             const void* B1, int64_t ldb1, int32_t k2, const void* B2, int64_t ldb2, int32_t sym,
             double* C, void* workspace, int64_t ws_bytes, void* stream);
/* Y1 = S C1 (and Y2 = S C2 when C2 != NULL) for one row-major S (n x m, lds);
 * C1, C2: m x k row-major FP64 (staged in the storage type).                     */
int qds_block_update(int32_t dtype, int64_t n, int32_t m, int32_t k, const void* S, int64_t lds,
                     const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2, int64_t ldy2,
                     void* stream);
/* Column-block forms: the logical n x m operand is up to four separately stored
 * row-major blocks side by side (block j: p[j], leading dimension ld[j], cols[j]
 * columns), so iterate blocks X, W, P can live in their own contiguous buffers.
 * Gram: C = A^T [B1 | B2] (B2 may be NULL).  Update: Y1 = S C1 (Y2 = S C2), C row-major
 * m x k FP64 (staged in the storage type, MFMA in the storage type), k <= 64, each block
 * 16-byte aligned with widths and leading dimensions a multiple of 4 (FP32) / 2 (FP64). */
typedef struct {
  const void* p[4];
  int64_t ld[4];
  int32_t cols[4];
  int32_t nb;
} qds_colblocks;
int qds_gram_blocks(int32_t dtype, int64_t n, const qds_colblocks* A /* host */,
                    const qds_colblocks* B1 /* host */, const qds_colblocks* B2 /* host */,
                    int32_t sym, double* C, void* workspace, int64_t ws_bytes, void* stream);
/* qds_gram_blocks without a workspace argument.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The per-slab partial sums live
 * in a static device array of qds_gram_noalloc_scratch_bytes() (64 MiB) on each
 * GPU, kept until the process ends; padded m x (k1 + k2) (each padded to a
 * multiple of 16) may be at most 131072 entries (e.g. m = 96 with k1 + k2 = 192
 * fits), larger shapes return QDS_EINVAL.  The rows are split into the same slabs
 * as qds_gram_blocks (so the results are bitwise equal) whenever the array holds
 * them all: padded m x (k1 + k2) <= 16384 entries (m = 72, k1 + k2 = 144 fits),
 * or fewer than 16 * (64 MiB / (8 * padded entries)) rows; otherwise it uses fewer,
 * longer slabs.  Calls that overlap on one GPU share the array and race: serialise
 * them (one stream, or events).                                                  */
int64_t qds_gram_noalloc_scratch_bytes(void);
int qds_gram_blocks_noalloc(int32_t dtype, int64_t n, const qds_colblocks* A /* host */,
                            const qds_colblocks* B1 /* host */, const qds_colblocks* B2 /* host */,
                            int32_t sym, double* C, void* stream);
int qds_block_update_blocks(int32_t dtype, int64_t n, const qds_colblocks* S /* host */, int32_t k,
                            const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                            int64_t ldy2, int32_t c2_first /* rows of C2 before it are zero */,
                            void* stream);
/* Batched forms (one launch for `batch` problems of equal shape, no reference
 * counterpart): every block of problem i is at its pointer + i * stride elements
 * (one stride for all blocks of a call); outputs C / Y at + i m (k1 + k2) / + i
 * y_stride; coefficients at + i c_stride.  Workspace of the Gram: batch times
 * qds_gram_workspace.                                                             */
int qds_gram_blocks_batched(int32_t dtype, int64_t n, const qds_colblocks* A /* host */,
                            const qds_colblocks* B1 /* host */, const qds_colblocks* B2 /* host */, int32_t sym,
                            double* C, int32_t batch, int64_t s_stride, void* workspace, int64_t ws_bytes,
                            void* stream);
int qds_block_update_blocks_batched(int32_t dtype, int64_t n, const qds_colblocks* S /* host */, int32_t k,
                                    const double* C1, void* Y1, int64_t ldy1, const double* C2, void* Y2,
                                    int64_t ldy2, int32_t c2_first, int32_t batch, int64_t s_stride,
                                    int64_t c_stride, int64_t y_stride, void* stream);
int qds_block_residual(int32_t dtype, int64_t n, int32_t b, const void* HX, int64_t ldh,
                       const void* MX, int64_t ldm, const double* lam, void* R, int64_t ldr,
                       void* stream);
/* Fused R = HX - MX diag(lambda) with out[j] = ||R_j||^2, out[b+j] = ||MX_j||^2
 * (FP64 sums, real dtypes): the residual step of each block-eigensolver iteration. */
int64_t qds_residual_norms_workspace(int32_t b);
int qds_residual_norms(int32_t dtype, int64_t n, int32_t b, const void* HX, int64_t ldh,
                       const void* MX, int64_t ldm, const double* lam,
                       const double* colscale /* optional: R columns scaled after the norms */,
                       void* R, int64_t ldr, double* out, void* workspace, int64_t ws_bytes,
                       void* stream);
/* Batched residual step: problem i's HX, MX, R at + i x_stride, lam and colscale
 * at + i b, out at + i 2b.                                                        */
int64_t qds_residual_norms_batched_workspace(int64_t n, int32_t b, int32_t batch);
int qds_residual_norms_batched(int32_t dtype, int64_t n, int32_t b, int32_t batch, const void* HX, int64_t ldh,
                               const void* MX, int64_t ldm, int64_t x_stride, const double* lam,
                               const double* colscale, void* R, int64_t ldr, double* out, void* workspace,
                               int64_t ws_bytes, void* stream);
/* Mixed-precision forms (no reference counterpart; they replace the same dense
 * steps of scipy's lobpcg, eigen_solver.pyx:254-266, for a basis whose blocks
 * differ in storage type: X and its images in FP64, the corrections W, P and
 * their images in FP32).
 * qds_residual_norms_to: qds_residual_norms with R stored in r_dtype (F32 from
 *   F64 images, or the images' type); the norms are FP64 sums of the FP64 residual.
 * qds_gram_blocks_typed: qds_gram_blocks for operands whose block j has storage
 *   type *_dt[j] (F32 / F64); rows converted to FP64 when staged, FP64 products
 *   and sums; workspace qds_gram_workspace.
 * qds_block_update_typed: Y1 = S C1 in FP64 and Y2 = S C2 in y2_dtype (F32 / F64)
 *   for blocks of storage type s_dt[j]: FP64 blocks with FP64 products and sums,
 *   FP32 blocks on the FP32 MFMA (C rounded to FP32), the two parts added in FP64;
 *   k <= 64, blocks of at most 32 (F64) / 64 (F32) columns, 16-byte aligned.     */
int qds_residual_norms_to(int32_t dtype, int32_t r_dtype, int64_t n, int32_t b, const void* HX, int64_t ldh,
                          const void* MX, int64_t ldm, const double* lam, const double* colscale, void* R,
                          int64_t ldr, double* out, void* workspace, int64_t ws_bytes, void* stream);
int qds_gram_blocks_typed(int64_t n, const qds_colblocks* A /* host */, const int32_t* a_dt /* host */,
                          const qds_colblocks* B1 /* host */, const int32_t* b1_dt /* host */,
                          const qds_colblocks* B2 /* host */, const int32_t* b2_dt /* host */, int32_t sym,
                          double* C, void* workspace, int64_t ws_bytes, void* stream);
int qds_block_update_typed(int64_t n, const qds_colblocks* S /* host */, const int32_t* s_dt /* host */, int32_t k,
                           const double* C1, void* Y1, int64_t ldy1, const double* C2, int32_t y2_dtype, void* Y2,
                           int64_t ldy2, int32_t c2_first, void* stream);
int64_t qds_coldot_workspace(int32_t b);
int qds_coldot(int32_t dtype, int64_t n, int32_t b, const void* X, int64_t ldx, const void* Y,
               int64_t ldy, int32_t conj, double* out, void* workspace, int64_t ws_bytes,
               void* stream);

/* ---- Rayleigh-Ritz on the device -------------------------------------------------
 * Replaces the dense projected eigensolve inside scipy's lobpcg / ARPACK
 * (qdsim_cython/solvers/eigen_solver.pyx:225-266): for each of `batch` problems,
 * G = [S^T H S | S^T M S] (m x 2m FP64, row-major, problem stride g_stride;
 * its M half is overwritten) gives the lowest b Ritz pairs: lam (in: the current
 * Ritz values; out: the new ones), C (m x b: X_new = S C) and CP (C with rows
 * < b zeroed, all zero after dropped directions: P_new = S CP; may be NULL), one
 * workgroup per problem (m even, m <= 96).  norms ([batch][2b], NULL for the
 * first projection) holds ||R_j||^2 and ||M x_j||^2 of the current block (from
 * qds_residual_norms); with them the call also writes the locking mask of the next
 * residual (colscale, may be NULL) and the status [batch][qds_rr_status_len(b)]:
 * [0] worst relative residual of the first k columns, [1] converged (< tol),
 * [2] dropped directions, [3] 0 Cholesky / 1 SVQB / 2 frozen, [4] Jacobi sweeps
 * (0 when the tridiagonal eigensolver produced the pairs), [5] non-vanished basis
 * columns, [6] 1 when the tridiagonal eigensolver produced the pairs (Cholesky
 * path, m >= 10, 2 b <= m + 1, every check passed), [8 + j] residual of column j.
 * With freeze, a converged problem keeps its block (C = [I; 0], lam unchanged). */
/* The same projected eigensolve on the host CPU, for one problem between two GPU
 * launches: G (m x 2m row-major FP64, host memory, m <= 192) gives theta (b
 * ascending Ritz values) and C (m x b row-major, host).  Cholesky of the
 * diagonally scaled M-Gram (vanished columns left out), Hager condition estimate,
 * T = L^-1 H L^-T, Householder tridiagonalisation, bisection, inverse iteration.
 * Returns 0, 1 when the basis is too close to dependent for the Cholesky path
 * (reciprocal condition below drop_tol; the caller selects a basis otherwise), or
 * a negative error code.                                                          */
int qds_host_rayleigh_ritz(int32_t m, int32_t b, const double* G /* host */, double drop_tol,
                           double* theta /* host */, double* C /* host */);
int32_t qds_rr_status_len(int32_t b);
int qds_rayleigh_ritz(int64_t batch, int32_t m, int32_t b, int32_t k, double* G, int64_t g_stride,
                      const double* norms, double* lam, double* C, double* CP, int64_t c_stride,
                      double* colscale, double* status, double tol, double drop_tol, int32_t freeze,
                      void* stream);

/* ---- lattice multigrid preconditioner (device) ---------------------------------
 * No reference counterpart: replaces the sparse LU of ARPACK shift-invert
 * (qdsim_cython/advanced_eigenvalue_solvers.py:106-122) as the spectral
 * accelerator of the eigensolver.                                                  */
typedef struct {
  int64_t nx, ny;
  const void* d;
  const void* we;
  const void* wn;
  void* u;
  void* f;
  void* t;
  int64_t ldu, ldf; /* per-point strides of u and f (0 means b) */
  int32_t neumann_x, neumann_y; /* 0: Dirichlet outside, coarse point j at fine 2j+1 (nx = 2 nx_c + 1);
                                   1: the boundary rows are unknowns, coarse point j at fine 2j
                                   (nx = 2 nx_c - 1); describes the coarsening of this level     */
} qds_mg_level;
int qds_mg_vcycle(int32_t dtype, int32_t n_levels, const qds_mg_level* levels, int32_t b,
                  int32_t nu_pre, int32_t nu_post, int32_t coarse_sweeps, double omega,
                  void* stream);
/* The same V-cycle with an FP32 hierarchy (dtype 1) on FP64 level-0 blocks
 * (io_dtype 2): levels[0].f / .u are FP64 (n, b) blocks, every other level FP32.
 * Lets an FP64 eigensolver phase use the FP32 preconditioner without converting
 * its residual and correction blocks.  Needs nu_pre = nu_post = 1 and a level 0
 * that is not inside the single-workgroup coarse cycle (else QDS_EINVAL).      */
int qds_mg_vcycle_io(int32_t dtype, int32_t io_dtype, int32_t n_levels, const qds_mg_level* levels,
                     int32_t b, int32_t nu_pre, int32_t nu_post, int32_t coarse_sweeps, double omega,
                     void* stream);
/* Batched V-cycle (no reference counterpart): `batch` problems on one hierarchy
 * shape, problem i's arrays at the level pointers plus i times the level's strides
 * (elements; 0 = shared, e.g. the edge weights when every problem has the same
 * mass).  dtype / io_dtype as qds_mg_vcycle_io.                                    */
typedef struct {
  int64_t sd, sw, su, sf, st;   /* d; we and wn; u; f; t */
} qds_mg_level_strides;
int qds_mg_vcycle_batched(int32_t dtype, int32_t io_dtype, int32_t n_levels, const qds_mg_level* levels,
                          const qds_mg_level_strides* strides /* host [n_levels] */, int32_t batch, int32_t b,
                          int32_t nu_pre, int32_t nu_post, int32_t coarse_sweeps, double omega, void* stream);
/* Construction of the eigensolver's lattice hierarchy (no reference counterpart: the
 * reference has no preconditioner).  Full-lattice FP64 arrays, row-major:
 * WE [NY][NX-1], WN [NY-1][NX], Mterm and Mass [NY][NX].
 * qds_mg_lattice_coefficients: level-0 edge weights (mean kinetic coefficient of
 * the two elements touching an edge x h_perp/h_par), mass term ((mean V over four
 * quarter points) - sigma) h_x h_y and lumped mass, from per-element c and V
 * (element 2(j cx + i) + upper of the cell holding a point; 1/dxc, 1/dyc given).
 * qds_mg_lattice_coarsen: NX x NY (odd) -> (NX+1)/2 x (NY+1)/2, harmonic-mean edge
 * weights, bilinear restriction of Mterm and Mass.
 * qds_mg_lattice_interior: the level's unknowns (interior points): d, we, wn in
 * dtype (qds_mg_level arrays), FP64 d and lumped mass (n = (NX-2)(NY-2)).          */
int qds_mg_lattice_coefficients(int64_t NX, int64_t NY, int64_t cx, int64_t cy, double hx, double hy,
                                double inv_dxc, double inv_dyc, const double* c_elem, const double* v_elem,
                                double sigma, double* WE, double* WN, double* Mterm, double* Mass, void* stream);
int qds_mg_lattice_coarsen(int64_t NX, int64_t NY, const double* WE, const double* WN, const double* Mterm,
                           const double* Mass, double* WEc, double* WNc, double* Mterm_c, double* Mass_c,
                           void* stream);
int qds_mg_lattice_interior(int32_t dtype, int64_t NX, int64_t NY, const double* WE, const double* WN,
                            const double* Mterm, const double* Mass, void* d, void* we, void* wn, double* d64,
                            double* mass, void* stream);
int qds_mg_apply(int32_t dtype, const qds_mg_level* level, int32_t b, const void* u, void* out,
                 void* stream);

/* ---- Poisson CG (device) -------------------------------------------------------
 * Replaces spla.spsolve of the Dirichlet-modified Poisson system
 * (qdsim_cython/solvers/poisson_solver.pyx:241).                                   */
int64_t qds_cg_workspace(int64_t n);
int qds_cg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
           const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
           void* workspace, int64_t ws_bytes, int32_t* iters_out /* host */,
           double* relres_out /* host */, void* stream);
/* The same CG preconditioned by one multigrid V(nu, nu)-cycle per iteration on
 * an FP64 lattice hierarchy (levels[0] holds the n unknowns in lattice order;
 * the Poisson problem's insulating top and bottom rows use neumann_y).  Same
 * workspace (qds_cg_workspace), stopping rule and outputs as qds_cg; replaces
 * the same spsolve (qdsim_cython/solvers/poisson_solver.pyx:241).               */
int qds_cg_mg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
              const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
              int32_t n_levels, const qds_mg_level* levels, int32_t nu, int32_t coarse_sweeps, double omega,
              void* workspace, int64_t ws_bytes, int32_t* iters_out /* host */,
              double* relres_out /* host */, void* stream);

/* ---- nonsymmetric Krylov (device) -------------------------------------------------
 * Jacobi left-preconditioned BiCGStab in FP64.  Replaces the SimplicialLDLT
 * factorisation and solve of the drift-diffusion systems
 * (backend/src/full_poisson_dd_solver.cpp:473-480, :1018-1040; on the S matrices
 * of the ldlt_lower mode it solves what that factorisation solves).  x holds the
 * initial guess; stops at ||D^-1 (b - A x)|| <= rtol ||D^-1 b|| (D = diag A;
 * recursive residual, read every check_every iterations) or max_iter; a
 * breakdown restarts from the last iterate (restarts_out, host, optional).  A
 * non-finite b gives a non-finite x and relres_out.                               */
int64_t qds_bicgstab_workspace(int64_t n);
int qds_bicgstab(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
                 const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
                 void* workspace, int64_t ws_bytes, int32_t* iters_out /* host */,
                 double* relres_out /* host */, int32_t* restarts_out /* host */, void* stream);
/* The same BiCGStab right-preconditioned by one multigrid V(nu, nu)-cycle per
 * product on an FP64 lattice hierarchy (levels[0] holds the n unknowns in lattice
 * order, as for qds_cg_mg); stops at ||b - A x|| <= rtol ||b||.  Replaces the same
 * factorisations for the continuity systems restricted to the free nodes of a
 * structured mesh (full_poisson_dd_solver.cpp:1018-1040).  Workspace as
 * qds_bicgstab_workspace.                                                         */
int qds_bicgstab_mg(int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* b, double* x, double rtol, int32_t max_iter, int32_t check_every,
                    int32_t n_levels, const qds_mg_level* levels, int32_t nu, int32_t coarse_sweeps, double omega,
                    void* workspace, int64_t ws_bytes, int32_t* iters_out /* host */,
                    double* relres_out /* host */, int32_t* restarts_out /* host */, void* stream);

/* ---- Poisson-drift-diffusion (device) ----------------------------------------------
 * FullPoissonDriftDiffusionSolver (backend/src/full_poisson_dd_solver.cpp) on a P1
 * triangle mesh.  Device arrays, FP64 node-major; fields and currents [n][2].
 * ne_ptr/ne_elem: the elements of each node in ascending id (the order of the
 * reference's per-node element loops).  mat: [n_mat][8] = N_c, N_v, E_g, mu_n,
 * mu_p, eps_r, n_i = sqrt(N_c N_v) exp(-E_g / 2kT), 0; mat_idx per node and
 * emat_idx per element (centroid) select the material of get_material_at
 * (:185-211).  contact: 0 interior, 1 x < tol (V_p contact), 2 x > Lx - tol
 * (V_n contact), as the reference's boundary tests (:404-416, :790-838).         */
typedef struct {
  int64_t n_nodes, n_elem;
  const int32_t* conn;                      /* [n_elem][3] */
  const double* x;
  const double* y;
  const int32_t* ne_ptr;                    /* [n_nodes + 1] */
  const int32_t* ne_elem;                   /* [3 n_elem] */
  const int32_t* indptr;                    /* CSR pattern of the mesh (row kernels) */
  const int32_t* indices;
  const int32_t* mat_idx;                   /* [n_nodes] */
  const int32_t* emat_idx;                  /* [n_elem] (Poisson rows) */
  const double* mat;                        /* [n_mat][8] */
  const int8_t* contact;                    /* [n_nodes] */
} qds_dd_mesh;
/* n, p from phi, phi_n, phi_p: Boltzmann N_c exp((phi_n - phi)/kT), N_v exp((phi -
 * phi_p - E_g)/kT), or with fermi_dirac the F_1/2 approximation; floor n_min with
 * std::max.  Replaces compute_carrier_concentrations_boltzmann / _fermi_dirac
 * (:656-718).                                                                     */
int qds_dd_carriers(const qds_dd_mesh* m, int32_t fermi_dirac, double kT, double n_min, const double* phi,
                    const double* phi_n, const double* phi_p, double* n, double* p, void* stream);
/* Ohmic contact values of apply_boundary_conditions (:778-839), N_a = N_d = n_contact;
 * fn / fp (optional): the contact rows of the continuity loads get the new phi_n /
 * phi_p (the Dirichlet values of :1026-1027).                                      */
int qds_dd_contacts(const qds_dd_mesh* m, double V_p, double V_n, double kT, double n_contact, double* phi,
                    double* phi_n, double* phi_p, double* n, double* p, double* fn, double* fp, void* stream);
/* Initial carriers and quasi-Fermi levels from the doping
 * (initialize_carrier_concentrations, :216-260).                                    */
int qds_dd_initial(const qds_dd_mesh* m, double kT, double n_min, const double* doping, double* n, double* p,
                   double* phi_n, double* phi_p, void* stream);
/* rho = scale (p - n + N); compute_charge_density (:267-288) with scale = q.        */
int qds_dd_charge(int64_t n_nodes, double scale, const double* n, const double* p, const double* doping,
                  double* rho, void* stream);
/* mode 0: phi = phi_old + damping (phi - phi_old) (apply_damping, :1168-1173).
 * mode 1: non-finite phi -> phi_old and sums[3] = {|phi|^2, |phi - phi_old|^2,
 * #non-finite} on the device (the Gummel test of solve, :1383-1408).
 * workspace: qds_dd_workspace() bytes (mode 1).                                    */
int64_t qds_dd_workspace(void);
int qds_dd_potential_update(int64_t n_nodes, int32_t mode, double damping, const double* phi_old, double* phi,
                            double* sums, void* workspace, int64_t ws_bytes, void* stream);
/* E = -grad phi averaged over each node's elements (compute_electric_field, :524-583). */
int qds_dd_field(const qds_dd_mesh* m, const double* phi, double* E, void* stream);
/* Jn, Jp = mean over the node's elements of q mu_c n_c grad phi_n, q mu_c p_c grad
 * phi_p (compute_current_densities, :1068-1159).                                     */
int qds_dd_currents(const qds_dd_mesh* m, double q, const double* phi_n, const double* phi_p, const double* n,
                    const double* p, const double* mu_n, const double* mu_p, double* Jn, double* Jp,
                    void* stream);
/* Continuity systems of solve_drift_diffusion_equations (:847-1040): values of
 * K_n, K_p on the mesh pattern and loads f_n, f_p, Dirichlet rows replaced by the
 * contact phi_n / phi_p.  G: per-node generation-recombination rate, or NULL for
 * the default SRH model (:72-82) evaluated from n, p.  ldlt_lower = 1 writes the
 * matrix the reference's SimplicialLDLT factors (lower triangle mirrored).         */
int qds_dd_continuity(const qds_dd_mesh* m, int32_t ldlt_lower, double kT, const double* E, const double* mu_n,
                      const double* mu_p, const double* G, const double* n, const double* p, const double* phi_n,
                      const double* phi_p, double* Kn, double* Kp, double* fn, double* fp, void* stream);
/* Poisson system of solve_poisson_equation (:296-471): K = eps_r eps0 grad.grad A,
 * f_i -= sum_q rho_q lambda_i (1/6) A, rho = q (p - n + N), Dirichlet rows V_p / V_n;
 * ldlt_lower as above.                                                               */
int qds_dd_poisson_system(const qds_dd_mesh* m, int32_t ldlt_lower, double eps0, double q, const double* n,
                          const double* p, const double* doping, double V_p, double V_n, double* K, double* f,
                          void* stream);

/* ---- adaptive refinement indicator (device) --------------------------------
 * Replaces AdaptiveMesh::computeRefinementFlags
 * (backend/src/adaptive_mesh.cpp:431-475): out[e] = |gradient of the least-squares
 * fit psi ~ a + b x + c y over element e's npe nodes|, coordinates scaled by
 * coord_scale first.  The caller compares out with the threshold.               */
int qds_gradient_indicator(int32_t npe, int64_t n_elem, const int32_t* conn, const double* x, const double* y,
                           const double* psi, double coord_scale, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* QDSIM_HIP_H */
