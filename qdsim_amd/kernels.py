"""Thin Python wrappers over the qdsim_hip C-ABI (one function per entry point).

Blocks are torch tensors of shape (n, b) with unit column stride; a column slice
of a wider buffer is passed with its row stride as the leading dimension.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, host_ptr, i32, lib
from .device import code, device, ptr, stream_ptr, workspace


# ------------------------------------------------------------------ host pattern
def pattern_build(conn, n_dof, with_elem_map=False):
    """CSR pattern + scatter maps of the element cliques (host, native C++)."""
    conn = i32(conn)
    ne, npe = conn.shape
    nnz = ctypes.c_int64()
    check(lib.qds_pattern_count(host_ptr(conn), ne, npe, n_dof, ctypes.byref(nnz)), "qds_pattern_count")
    nnz = nnz.value
    indptr = np.empty(n_dof + 1, np.int32)
    indices = np.empty(nnz, np.int32)
    seg_ptr = np.empty(nnz + 1, np.int32)
    seg_src = np.empty(ne * npe * npe, np.int32)
    emap = np.empty(ne * npe * npe, np.int32) if with_elem_map else None
    check(lib.qds_pattern_fill(host_ptr(conn), ne, npe, n_dof, nnz, host_ptr(indptr), host_ptr(indices),
                               host_ptr(seg_ptr), host_ptr(seg_src),
                               host_ptr(emap) if emap is not None else None), "qds_pattern_fill")
    out = {"indptr": indptr, "indices": indices, "seg_ptr": seg_ptr, "seg_src": seg_src, "nnz": nnz}
    if emap is not None:
        out["elem_map"] = emap.reshape(ne, npe, npe)
    return out


def pattern_restrict(indptr, indices, new_of_old, n_new):
    indptr = i32(indptr); indices = i32(indices); new_of_old = i32(new_of_old)
    n = len(indptr) - 1
    nnz = ctypes.c_int64()
    check(lib.qds_pattern_restrict_count(host_ptr(indptr), host_ptr(indices), n, host_ptr(new_of_old),
                                         n_new, ctypes.byref(nnz)), "qds_pattern_restrict_count")
    nnz = nnz.value
    r_ip = np.empty(n_new + 1, np.int32)
    r_ix = np.empty(nnz, np.int32)
    r_src = np.empty(nnz, np.int32)
    check(lib.qds_pattern_restrict_fill(host_ptr(indptr), host_ptr(indices), n, host_ptr(new_of_old), n_new,
                                        nnz, host_ptr(r_ip), host_ptr(r_ix), host_ptr(r_src)),
          "qds_pattern_restrict_fill")
    return r_ip, r_ix, r_src


# ------------------------------------------------------------------ assembly
def _coef(t):
    """(pointer, element stride, quadrature stride) of a coefficient tensor.

    Shapes: (E, Q) per quadrature point, (E,) per element, or None."""
    if t is None:
        return None, 0, 0
    assert t.dtype == torch.float64 and t.is_contiguous()
    if t.dim() == 1:
        return ptr(t), 1, 0
    return ptr(t), t.shape[1], 1


def assemble_elements(npe, nq, conn, x, y, tables, c=None, v=None, r=None, w=None,
                      out0=True, out1=False, out2=False):
    ne = conn.shape[0]
    nsym = npe * (npe + 1) // 2
    dev = conn.device
    o0 = torch.empty(ne * nsym, dtype=torch.float64, device=dev) if out0 else None
    o1 = torch.empty(ne * nsym, dtype=torch.float64, device=dev) if out1 else None
    o2 = torch.empty(ne * nsym, dtype=torch.float64, device=dev) if out2 else None
    cp, cse, csq = _coef(c)
    vp, vse, vsq = _coef(v)
    rp, rse, rsq = _coef(r)
    wp, wse, wsq = _coef(w)
    check(lib.qds_assemble_elements(npe, nq, ne, ptr(conn), ptr(x), ptr(y), ptr(tables),
                                    cp, cse, csq, vp, vse, vsq, rp, rse, rsq, wp, wse, wsq,
                                    ptr(o0), ptr(o1), ptr(o2), stream_ptr()), "qds_assemble_elements")
    return o0, o1, o2


def assemble_gather(nnz, seg_ptr, seg_src, e0=None, e1=None, e2=None):
    dev = seg_ptr.device
    v = [torch.empty(nnz, dtype=torch.float64, device=dev) if e is not None else None for e in (e0, e1, e2)]
    check(lib.qds_assemble_gather(nnz, ptr(seg_ptr), ptr(seg_src), ptr(e0), ptr(e1), ptr(e2),
                                  ptr(v[0]), ptr(v[1]), ptr(v[2]), stream_ptr()), "qds_assemble_gather")
    return v


def quad_points(npe, conn, x, y, bary):
    ne = conn.shape[0]
    nq = bary.shape[0]
    xq = torch.empty((ne, nq), dtype=torch.float64, device=conn.device)
    yq = torch.empty_like(xq)
    check(lib.qds_quad_points(npe, nq, ne, ptr(conn), ptr(x), ptr(y), ptr(bary), ptr(xq), ptr(yq),
                              stream_ptr()), "qds_quad_points")
    return xq, yq


def csr_take(src, re, im=None, scale=1.0, dtype=torch.float64):
    n = src.numel()
    out = torch.empty(n, dtype=dtype, device=src.device)
    check(lib.qds_csr_take(code(out), n, ptr(src), ptr(re), ptr(im), float(scale), ptr(out), stream_ptr()),
          "qds_csr_take")
    return out


# ------------------------------------------------------------------ blocks
def _ld(t):
    if t.dim() == 1:
        return 1
    assert t.stride(1) == 1, "blocks need unit column stride"
    return t.stride(0)


def _width(t):
    return 1 if t.dim() == 1 else t.shape[1]


def spmm(indptr, indices, n_rows, nnz, vals_a, X, Ya, vals_b=None, Yb=None):
    assert vals_a.dtype == X.dtype == Ya.dtype
    b = _width(X)
    check(lib.qds_spmm(code(X), n_rows, nnz, b, ptr(indptr), ptr(indices), ptr(vals_a), ptr(vals_b),
                       ptr(X), _ld(X), ptr(Ya), _ld(Ya), ptr(Yb), _ld(Yb) if Yb is not None else 1,
                       stream_ptr()), "qds_spmm")


def lattice_stencil(order, nx, ny, indptr, indices, with_map=True):
    """Stencil form of a lattice-ordered CSR pattern (host arrays).

    Returns (qds_lattice_stencil, slot_src int64 numpy array or None): slot_src[s]
    is the CSR position feeding value slot s, -1 for structural zeros."""
    st = _lib.LatticeStencil()
    ip = i32(indptr)
    ix = i32(indices)
    check(lib.qds_lattice_stencil_build(int(order), int(nx), int(ny), host_ptr(ip), host_ptr(ix),
                                        ctypes.byref(st), None), "qds_lattice_stencil_build")
    if not with_map:
        return st, None
    slot_src = np.empty(st.n_slots, dtype=np.int64)
    check(lib.qds_lattice_stencil_build(int(order), int(nx), int(ny), host_ptr(ip), host_ptr(ix),
                                        ctypes.byref(st), host_ptr(slot_src)), "qds_lattice_stencil_build")
    return st, slot_src


def lattice_spmm(st, vals_a, X, Ya, vals_b=None, Yb=None):
    """Ya = A_a X (and Yb = A_b X) with A in the lattice form described by st."""
    assert vals_a.dtype == X.dtype == Ya.dtype
    b = _width(X)
    check(lib.qds_lattice_spmm(code(X), ctypes.byref(st), b, ptr(vals_a), ptr(vals_b), ptr(X), _ld(X),
                               ptr(Ya), _ld(Ya), ptr(Yb), _ld(Yb) if Yb is not None else 1, stream_ptr()),
          "qds_lattice_spmm")


def lattice_spmm_tuned(st, vals_a, X, Ya, vals_b=None, Yb=None, choice=None):
    """The lattice product with an explicit block shape (autotune choice): X, Ya, Yb
    (n, b) or (batch, n, b); vals (slots,) shared or (batch, slots)."""
    assert vals_a.dtype == X.dtype == Ya.dtype
    from . import autotune
    t = autotune.to_struct(choice or autotune.DEFAULT)
    if X.dim() == 2:
        batch, sx, sya, syb = 1, 0, 0, 0
        ldx, ldya, ldyb = _ld(X), _ld(Ya), _ld(Yb) if Yb is not None else 1
        b = _width(X)
    else:
        batch, b = X.shape[0], X.shape[2]
        sx, sya, syb = X.stride(0), Ya.stride(0), Yb.stride(0) if Yb is not None else 0
        ldx, ldya, ldyb = X.stride(1), Ya.stride(1), Yb.stride(1) if Yb is not None else 1

    def vstride(V):
        return 0 if V is None or V.dim() == 1 else V.stride(0)
    check(lib.qds_lattice_spmm_tuned(code(X), ctypes.byref(st), b, batch, ptr(vals_a), vstride(vals_a), ptr(vals_b),
                                     vstride(vals_b), ptr(X), ldx, sx, ptr(Ya), ldya, sya, ptr(Yb), ldyb, syb,
                                     ctypes.byref(t), stream_ptr()), "qds_lattice_spmm_tuned")


def lattice_spmm_auto(st, vals_a, X, Ya, vals_b=None, Yb=None):
    """The lattice product with the block shape the autotuner chose for this shape
    (timed on first use, see qdsim_amd.autotune)."""
    from . import autotune
    b = X.shape[-1]
    batch = 1 if X.dim() == 2 else X.shape[0]
    choice = autotune.choice_for(st, X.dtype, b, vals_b is not None, batch,
                                 lambda ch: lattice_spmm_tuned(st, vals_a, X, Ya, vals_b, Yb, ch))
    lattice_spmm_tuned(st, vals_a, X, Ya, vals_b, Yb, choice)


def lattice_spmm_residual(st, vals_a, vals_b, X, Ya, Yb, lam, R, colscale=None, out=None, choice=None):
    """Ya = H X, Yb = M X (lattice form, the block shape ``choice``) with the residual
    step fused in: R = (Ya - Yb diag(lam)) diag(colscale) in R's dtype, and
    (||Ya - Yb lam||_j^2, ||Yb_j||^2) returned as views of ``out`` (2b FP64 device
    values, allocated when None) -- the results of ``residual_norms``."""
    from . import autotune
    assert vals_a.dtype == vals_b.dtype == X.dtype == Ya.dtype == Yb.dtype
    n, b = X.shape
    if out is None:
        out = torch.empty(2 * b, dtype=torch.float64, device=X.device)
    assert out.dtype == torch.float64 and out.is_contiguous() and out.numel() >= 2 * b
    t = autotune.to_struct(choice or autotune.DEFAULT)
    wsb = lib.qds_lattice_spmm_residual_workspace(ctypes.byref(st), b)
    wp, wn = workspace("lattice_res").ptr(wsb)
    cs = colscale.contiguous() if colscale is not None else None
    check(lib.qds_lattice_spmm_residual(code(X), ctypes.byref(st), b, ptr(vals_a), ptr(vals_b), ptr(X), _ld(X),
                                        ptr(Ya), _ld(Ya), ptr(Yb), _ld(Yb), ptr(lam.contiguous()), ptr(cs), code(R),
                                        ptr(R), _ld(R), ptr(out), wp, wn, ctypes.byref(t), stream_ptr()),
          "qds_lattice_spmm_residual")
    return out[:b], out[b:2 * b]


def spmm_cap(indptr, indices, n_rows, h, w, m, X, YH, YM):
    """YH = (H - iW) X, YM = M X for complex blocks, real value arrays."""
    b = _width(X)
    check(lib.qds_spmm_cap(code(X), n_rows, b, ptr(indptr), ptr(indices), ptr(h), ptr(w), ptr(m),
                           ptr(X), _ld(X), ptr(YH), _ld(YH), ptr(YM), _ld(YM), stream_ptr()), "qds_spmm_cap")


def gram(A, B1, B2=None, sym=False, out=None):
    """A^T [B1 | B2] in FP64, shape (m, k1 + k2)."""
    n, m = A.shape
    k1 = B1.shape[1]
    k2 = B2.shape[1] if B2 is not None else 0
    if out is None:
        out = torch.empty((m, k1 + k2), dtype=torch.float64, device=A.device)
    wsb = lib.qds_gram_workspace(n, m, k1, k2)
    wp, wn = workspace("gram").ptr(wsb)
    check(lib.qds_gram(code(A), n, m, ptr(A), _ld(A), k1, ptr(B1), _ld(B1), k2, ptr(B2),
                       _ld(B2) if B2 is not None else 1, 1 if sym else 0, ptr(out), wp, wn, stream_ptr()),
          "qds_gram")
    return out


def block_update(S, C1, Y1, C2=None, Y2=None):
    """Y1 = S C1 (and Y2 = S C2); C passed in FP64."""
    n, m = S.shape
    k = C1.shape[1]
    C1 = C1.to(torch.float64).contiguous()
    C2 = C2.to(torch.float64).contiguous() if C2 is not None else None
    check(lib.qds_block_update(code(S), n, m, k, ptr(S), _ld(S), ptr(C1), ptr(Y1), _ld(Y1),
                               ptr(C2), ptr(Y2), _ld(Y2) if Y2 is not None else 1, stream_ptr()),
          "qds_block_update")


def _colblocks(blocks):
    """qds_colblocks for a list of (n, c_j) row-major blocks (views with unit column stride)."""
    assert 1 <= len(blocks) <= 4
    cb = _lib.ColBlocks()
    for j, B in enumerate(blocks):
        assert B.stride(1) == 1, "column blocks need unit column stride"
        cb.p[j] = B.data_ptr()
        cb.ld[j] = _ld(B)
        cb.cols[j] = B.shape[1]
    cb.nb = len(blocks)
    return cb


def gram_blocks(A, B1, B2=None, sym=False, out=None):
    """[A_0 | A_1 | ...]^T [B1 | B2] in FP64 for lists of column blocks."""
    n = A[0].shape[0]
    m = sum(a.shape[1] for a in A)
    k1 = sum(x.shape[1] for x in B1)
    k2 = sum(x.shape[1] for x in B2) if B2 else 0
    if out is None:
        out = torch.empty((m, k1 + k2), dtype=torch.float64, device=A[0].device)
    ca, c1 = _colblocks(A), _colblocks(B1)
    c2 = _colblocks(B2) if B2 else None
    wsb = lib.qds_gram_workspace(n, m, k1, k2)
    wp, wn = workspace("gram").ptr(wsb)
    check(lib.qds_gram_blocks(code(A[0]), n, ctypes.byref(ca), ctypes.byref(c1),
                              ctypes.byref(c2) if c2 is not None else None, 1 if sym else 0, ptr(out),
                              wp, wn, stream_ptr()), "qds_gram_blocks")
    return out


def host_rayleigh_ritz(G, m, b, drop_tol):
    """Lowest b Ritz pairs of the projected pencil G = [S^T H S | S^T M S] (host
    FP64 array, m x 2m) by qds_host_rayleigh_ritz: (theta (b,), C (m, b)), or None
    when the Cholesky path does not apply (near-dependent basis)."""
    G = np.ascontiguousarray(G, dtype=np.float64)
    assert G.shape == (m, 2 * m)
    theta = np.empty(b)
    C = np.empty((m, b))
    rc = lib.qds_host_rayleigh_ritz(int(m), int(b), G.ctypes.data, float(drop_tol), theta.ctypes.data,
                                    C.ctypes.data)
    if rc == 1:
        return None
    check(rc, "qds_host_rayleigh_ritz")
    return theta, C


def _dtypes(blocks):
    arr = (ctypes.c_int32 * 4)()
    for j, B in enumerate(blocks):
        arr[j] = code(B)
    return arr


def gram_blocks_typed(A, B1, B2=None, sym=False, out=None):
    """gram_blocks for column blocks of different storage types (FP32 / FP64):
    FP64 products and sums (qds_gram_blocks_typed)."""
    n = A[0].shape[0]
    m = sum(a.shape[1] for a in A)
    k1 = sum(x.shape[1] for x in B1)
    k2 = sum(x.shape[1] for x in B2) if B2 else 0
    if out is None:
        out = torch.empty((m, k1 + k2), dtype=torch.float64, device=A[0].device)
    ca, c1 = _colblocks(A), _colblocks(B1)
    c2 = _colblocks(B2) if B2 else None
    ta, t1 = _dtypes(A), _dtypes(B1)
    t2 = _dtypes(B2) if B2 else None
    wsb = lib.qds_gram_workspace(n, m, k1, k2)
    wp, wn = workspace("gram").ptr(wsb)
    check(lib.qds_gram_blocks_typed(n, ctypes.byref(ca), ta, ctypes.byref(c1), t1,
                                    ctypes.byref(c2) if c2 is not None else None, t2, 1 if sym else 0, ptr(out),
                                    wp, wn, stream_ptr()), "qds_gram_blocks_typed")
    return out


def block_update_typed(S, C1, Y1, C2=None, Y2=None, c2_first=0):
    """Y1 = [S_0 | S_1 | ...] C1 in FP64 (and Y2 = ... C2 in Y2's type) for blocks of
    different storage types; FP64 products (qds_block_update_typed).  C1, C2:
    contiguous FP64 (m, k)."""
    n = S[0].shape[0]
    k = C1.shape[1]
    assert Y1.dtype == torch.float64 and C1.dtype == torch.float64 and C1.is_contiguous()
    assert C2 is None or (C2.dtype == torch.float64 and C2.is_contiguous())
    cs = _colblocks(S)
    check(lib.qds_block_update_typed(n, ctypes.byref(cs), _dtypes(S), k, ptr(C1), ptr(Y1), _ld(Y1), ptr(C2),
                                     code(Y2) if Y2 is not None else 0, ptr(Y2),
                                     _ld(Y2) if Y2 is not None else 1, int(c2_first) if C2 is not None else 0,
                                     stream_ptr()), "qds_block_update_typed")


# ---- batched forms: every operand (batch, n, b) contiguous, one launch per stage
def _bcolblocks(blocks):
    """qds_colblocks of set 0 for (batch, n, c_j) tensors and their common set stride."""
    stride = blocks[0].stride(0)
    assert all(B.dim() == 3 and B.stride(0) == stride and B.stride(2) == 1 for B in blocks)
    cb = _lib.ColBlocks()
    for j, B in enumerate(blocks):
        cb.p[j] = B.data_ptr()
        cb.ld[j] = B.stride(1)
        cb.cols[j] = B.shape[2]
    cb.nb = len(blocks)
    return cb, stride


def lattice_spmm_batched(stencil, A, X, YA, B=None, YB=None):
    """YA[i] = A_i X[i] (and YB[i] = B_i X[i]) for a batch on one stencil; A, B are
    (batch, slots) or (slots,) (shared by every problem)."""
    batch, n, b = X.shape

    def vstride(V):
        return 0 if V.dim() == 1 else V.stride(0)
    check(lib.qds_lattice_spmm_batched(code(X), ctypes.byref(stencil), b, batch, ptr(A), vstride(A), ptr(B),
                                       vstride(B) if B is not None else 0, ptr(X), X.stride(1), X.stride(0),
                                       ptr(YA), YA.stride(1), YA.stride(0), ptr(YB),
                                       YB.stride(1) if YB is not None else 1, YB.stride(0) if YB is not None else 0,
                                       stream_ptr()), "qds_lattice_spmm_batched")


def gram_blocks_batched(A, B1, B2=None, sym=False, out=None):
    """out[i] = [A_0[i] | ...]^T [B1[i] | B2[i]] in FP64 for lists of (batch, n, c) blocks."""
    batch, n = A[0].shape[:2]
    m = sum(a.shape[2] for a in A)
    k1 = sum(x.shape[2] for x in B1)
    k2 = sum(x.shape[2] for x in B2) if B2 else 0
    if out is None:
        out = torch.empty((batch, m, k1 + k2), dtype=torch.float64, device=A[0].device)
    ca, st = _bcolblocks(A)
    c1, st1 = _bcolblocks(B1)
    c2, st2 = _bcolblocks(B2) if B2 else (None, st)
    assert st == st1 == st2 and out.is_contiguous()
    wsb = batch * lib.qds_gram_workspace(n, m, k1, k2)
    wp, wn = workspace("gram").ptr(wsb)
    check(lib.qds_gram_blocks_batched(code(A[0]), n, ctypes.byref(ca), ctypes.byref(c1),
                                      ctypes.byref(c2) if c2 is not None else None, 1 if sym else 0, ptr(out), batch,
                                      st, wp, wn, stream_ptr()), "qds_gram_blocks_batched")
    return out


def block_update_blocks_batched(S, C1, Y1, C2=None, Y2=None, c2_first=0):
    """Y1[i] = [S_0[i] | ...] C1[i] (and Y2[i] = ... C2[i]); C (batch, m, k) FP64."""
    batch, n = S[0].shape[:2]
    k = C1.shape[2]
    cs, st = _bcolblocks(S)
    # coefficients: (batch, m, k) with contiguous rows; their set stride may be a
    # larger buffer's (a row slice of it)
    for Cx in (C1, C2):
        assert Cx is None or (Cx.dtype == torch.float64 and Cx.stride(2) == 1 and Cx.stride(1) == k)
    assert C2 is None or C2.stride(0) == C1.stride(0)
    check(lib.qds_block_update_blocks_batched(code(S[0]), n, ctypes.byref(cs), k, ptr(C1), ptr(Y1), Y1.stride(1),
                                              ptr(C2), ptr(Y2), Y2.stride(1) if Y2 is not None else 1,
                                              int(c2_first) if C2 is not None else 0, batch, st, C1.stride(0),
                                              Y1.stride(0), stream_ptr()), "qds_block_update_blocks_batched")


def residual_norms_batched(HX, MX, lam, R, colscale, out):
    """R[i] = (HX[i] - MX[i] diag(lam[i])) diag(colscale[i]); out[i] = (||r_j||^2, ||MX_j||^2)."""
    batch, n, b = HX.shape
    assert HX.stride(0) == MX.stride(0) == R.stride(0)
    wsb = lib.qds_residual_norms_batched_workspace(n, b, batch)
    wp, wn = workspace("resnorm").ptr(wsb)
    check(lib.qds_residual_norms_batched(code(HX), n, b, batch, ptr(HX), HX.stride(1), ptr(MX), MX.stride(1),
                                         HX.stride(0), ptr(lam), ptr(colscale), ptr(R), R.stride(1), ptr(out), wp, wn,
                                         stream_ptr()), "qds_residual_norms_batched")


def rr_status_len(b):
    return int(lib.qds_rr_status_len(int(b)))


def rayleigh_ritz(G, lam, C, status, k, tol, drop_tol, norms=None, CP=None, colscale=None, freeze=True):
    """Device Rayleigh-Ritz (qds_rayleigh_ritz) of a batch of projected pencils.
    G (batch, m, 2m) FP64 (its M half is overwritten), lam (batch, b) in/out,
    C / CP (batch, m, b), norms (batch, 2b), colscale (batch, b), status
    (batch, rr_status_len(b)); all contiguous float64 device tensors."""
    batch, m, m2 = G.shape
    b = lam.shape[-1]
    assert m2 == 2 * m and C.shape[-2:] == (m, b) and status.shape[-1] == rr_status_len(b)
    for t in (G, lam, C, status, norms, CP, colscale):
        assert t is None or (t.dtype == torch.float64 and t.is_contiguous())
    check(lib.qds_rayleigh_ritz(batch, m, b, int(k), ptr(G), m * 2 * m, ptr(norms), ptr(lam), ptr(C), ptr(CP), m * b,
                                ptr(colscale), ptr(status), float(tol), float(drop_tol), 1 if freeze else 0,
                                stream_ptr()), "qds_rayleigh_ritz")


def block_update_blocks(S, C1, Y1, C2=None, Y2=None, c2_first=0):
    """Y1 = [S_0 | S_1 | ...] C1 (and Y2 = ... C2); C in the storage dtype of S.
    c2_first: rows of C2 before it are zero (the kernel skips them)."""
    n = S[0].shape[0]
    k = C1.shape[1]
    # coefficients in FP64 (the kernel stages them in the storage type)
    C1 = C1.to(torch.float64).contiguous()
    C2 = C2.to(torch.float64).contiguous() if C2 is not None else None
    cs = _colblocks(S)
    check(lib.qds_block_update_blocks(code(S[0]), n, ctypes.byref(cs), k, ptr(C1), ptr(Y1), _ld(Y1),
                                      ptr(C2), ptr(Y2), _ld(Y2) if Y2 is not None else 1,
                                      int(c2_first) if C2 is not None else 0, stream_ptr()),
          "qds_block_update_blocks")


def block_residual(HX, MX, lam, R):
    """R = HX - MX diag(lam); lam float64 (b,) or complex128 (b,)."""
    n, b = HX.shape
    lam = lam.contiguous()
    check(lib.qds_block_residual(code(HX), n, b, ptr(HX), _ld(HX), ptr(MX), _ld(MX), ptr(lam),
                                 ptr(R), _ld(R), stream_ptr()), "qds_block_residual")


def residual_norms(HX, MX, lam, R, colscale=None, out=None):
    """R = (HX - MX diag(lam)) diag(colscale); returns (||HX - MX lam||_j^2, ||MX_j||^2)
    (views of ``out``, 2b FP64 device values, allocated when None)."""
    n, b = HX.shape
    if out is None:
        out = torch.empty(2 * b, dtype=torch.float64, device=HX.device)
    assert out.dtype == torch.float64 and out.is_contiguous() and out.numel() >= 2 * b
    wsb = lib.qds_residual_norms_workspace(b)
    wp, wn = workspace("resnorm").ptr(wsb)
    cs = colscale.contiguous() if colscale is not None else None
    if R.dtype != HX.dtype:
        # FP32 residual from FP64 images (mixed-precision iteration)
        check(lib.qds_residual_norms_to(code(HX), code(R), n, b, ptr(HX), _ld(HX), ptr(MX), _ld(MX),
                                        ptr(lam.contiguous()), ptr(cs), ptr(R), _ld(R), ptr(out), wp, wn,
                                        stream_ptr()), "qds_residual_norms_to")
        return out[:b], out[b:]
    check(lib.qds_residual_norms(code(HX), n, b, ptr(HX), _ld(HX), ptr(MX), _ld(MX), ptr(lam.contiguous()),
                                 ptr(cs), ptr(R), _ld(R), ptr(out), wp, wn, stream_ptr()), "qds_residual_norms")
    return out[:b], out[b:]


def coldot(X, Y, conj=True):
    n, b = X.shape
    cplx = X.is_complex()
    out = torch.empty(b, dtype=torch.complex128 if cplx else torch.float64, device=X.device)
    wsb = lib.qds_coldot_workspace(b)
    wp, wn = workspace("coldot").ptr(wsb)
    check(lib.qds_coldot(code(X), n, b, ptr(X), _ld(X), ptr(Y), _ld(Y), 1 if conj else 0, ptr(out),
                         wp, wn, stream_ptr()), "qds_coldot")
    return out


def mg_vcycle(dtype_code, levels_struct, n_levels, b, nu_pre, nu_post, coarse, omega):
    check(lib.qds_mg_vcycle(dtype_code, n_levels, ctypes.cast(levels_struct, ctypes.c_void_p), b,
                            nu_pre, nu_post, coarse, float(omega), stream_ptr()), "qds_mg_vcycle")


def mg_vcycle_io(dtype_code, io_code, levels_struct, n_levels, b, nu_pre, nu_post, coarse, omega):
    check(lib.qds_mg_vcycle_io(dtype_code, io_code, n_levels, ctypes.cast(levels_struct, ctypes.c_void_p), b,
                               nu_pre, nu_post, coarse, float(omega), stream_ptr()), "qds_mg_vcycle_io")


def mg_vcycle_batched(dtype_code, io_code, levels_struct, strides_struct, n_levels, batch, b, nu_pre, nu_post, coarse,
                      omega):
    check(lib.qds_mg_vcycle_batched(dtype_code, io_code, n_levels, ctypes.cast(levels_struct, ctypes.c_void_p),
                                    ctypes.cast(strides_struct, ctypes.c_void_p), batch, b, nu_pre, nu_post, coarse,
                                    float(omega), stream_ptr()), "qds_mg_vcycle_batched")


def mg_apply(dtype_code, level_struct, b, u, out):
    check(lib.qds_mg_apply(dtype_code, ctypes.cast(ctypes.byref(level_struct), ctypes.c_void_p), b,
                           ptr(u), ptr(out), stream_ptr()), "qds_mg_apply")


def cg_mg(indptr, indices, vals, n, nnz, b, x, levels_struct, n_levels, nu=1, coarse=10, omega=0.8, rtol=1e-10,
          max_iter=10000, check_every=5):
    """CG preconditioned by a multigrid V-cycle (qds_cg_mg); x holds the initial guess."""
    wsb = lib.qds_cg_workspace(n)
    wp, wn = workspace("cg").ptr(wsb)
    it = ctypes.c_int32()
    rr = ctypes.c_double()
    check(lib.qds_cg_mg(n, nnz, ptr(indptr), ptr(indices), ptr(vals), ptr(b), ptr(x), float(rtol), int(max_iter),
                        int(check_every), int(n_levels), levels_struct, int(nu), int(coarse), float(omega), wp, wn,
                        ctypes.byref(it), ctypes.byref(rr), stream_ptr()), "qds_cg_mg")
    return it.value, rr.value


def cg(indptr, indices, vals, n, nnz, b, x, rtol=1e-10, max_iter=10000, check_every=10):
    wsb = lib.qds_cg_workspace(n)
    wp, wn = workspace("cg").ptr(wsb)
    it = ctypes.c_int32()
    rr = ctypes.c_double()
    check(lib.qds_cg(n, nnz, ptr(indptr), ptr(indices), ptr(vals), ptr(b), ptr(x), float(rtol),
                     int(max_iter), int(check_every), wp, wn, ctypes.byref(it), ctypes.byref(rr),
                     stream_ptr()), "qds_cg")
    return it.value, rr.value


def bicgstab(indptr, indices, vals, n, nnz, b, x, rtol=1e-12, max_iter=20000, check_every=10):
    """Jacobi-preconditioned BiCGStab (qds_bicgstab); x holds the initial guess.
    Returns (iterations, relative residual, breakdown restarts)."""
    wsb = lib.qds_bicgstab_workspace(n)
    wp, wn = workspace("bicgstab").ptr(wsb)
    it, rr, rs = ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
    check(lib.qds_bicgstab(n, nnz, ptr(indptr), ptr(indices), ptr(vals), ptr(b), ptr(x), float(rtol), int(max_iter),
                           int(check_every), wp, wn, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(rs),
                           stream_ptr()), "qds_bicgstab")
    return it.value, rr.value, rs.value


def bicgstab_mg(indptr, indices, vals, n, nnz, b, x, levels_struct, n_levels, nu=1, coarse=10, omega=0.8,
                rtol=1e-12, max_iter=2000, check_every=5):
    """BiCGStab right-preconditioned by a multigrid V-cycle (qds_bicgstab_mg)."""
    wsb = lib.qds_bicgstab_workspace(n)
    wp, wn = workspace("bicgstab").ptr(wsb)
    it, rr, rs = ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
    check(lib.qds_bicgstab_mg(n, nnz, ptr(indptr), ptr(indices), ptr(vals), ptr(b), ptr(x), float(rtol),
                              int(max_iter), int(check_every), int(n_levels), levels_struct, int(nu), int(coarse),
                              float(omega), wp, wn, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(rs),
                              stream_ptr()), "qds_bicgstab_mg")
    return it.value, rr.value, rs.value


def mg_lattice_coefficients(NX, NY, cx, cy, hx, hy, inv_dxc, inv_dyc, c_elem, v_elem, sigma, WE, WN, Mterm, Mass):
    check(lib.qds_mg_lattice_coefficients(int(NX), int(NY), int(cx), int(cy), float(hx), float(hy), float(inv_dxc),
                                          float(inv_dyc), ptr(c_elem), ptr(v_elem), float(sigma), ptr(WE), ptr(WN),
                                          ptr(Mterm), ptr(Mass), stream_ptr()), "qds_mg_lattice_coefficients")


def mg_lattice_coarsen(NX, NY, WE, WN, Mterm, Mass, WEc, WNc, Mc, Massc):
    check(lib.qds_mg_lattice_coarsen(int(NX), int(NY), ptr(WE), ptr(WN), ptr(Mterm), ptr(Mass), ptr(WEc), ptr(WNc),
                                     ptr(Mc), ptr(Massc), stream_ptr()), "qds_mg_lattice_coarsen")


def mg_lattice_interior(dtype, NX, NY, WE, WN, Mterm, Mass, d, we, wn, d64, mass):
    code_ = _lib.F32 if dtype == torch.float32 else _lib.F64
    check(lib.qds_mg_lattice_interior(code_, int(NX), int(NY), ptr(WE), ptr(WN), ptr(Mterm), ptr(Mass), ptr(d),
                                      ptr(we), ptr(wn), ptr(d64), ptr(mass), stream_ptr()), "qds_mg_lattice_interior")
