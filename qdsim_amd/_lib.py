"""ctypes binding of libqdsim_hip.so (the C-ABI declared in include/qdsim_hip.h).

The library is required: importing a solver without it raises immediately.  There
is no CPU fallback anywhere in the package.
"""
import ctypes
import os

import numpy as np
# torch first: it loads the HIP runtime it ships with, and libqdsim_hip.so then
# binds to that same runtime.  Loaded the other way round, the dynamic linker
# brings in the system libamdhip64 as a second runtime, and kernels launched
# through it reported "no ROCm-capable device" on the GPU host although torch
# saw the device.  Device buffers come from torch, so one runtime per process.
import torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqdsim_hip.so")

F32, F64, C64, C128 = 1, 2, 3, 4

_c = ctypes
_p = _c.c_void_p
_i32 = _c.c_int32
_i64 = _c.c_int64
_d = _c.c_double

_SIGNATURES = {
    "qds_last_error": ([], _c.c_char_p),
    "qds_version": ([], _c.c_int),
    "qds_device_info": ([_c.c_char_p, _i32, _c.POINTER(_i32)], _c.c_int),
    "qds_runtime_check": ([], _c.c_int),
    "qds_runtime_images": ([], _c.c_int),
    "qds_pattern_count": ([_p, _i64, _i32, _i64, _c.POINTER(_i64)], _c.c_int),
    "qds_pattern_fill": ([_p, _i64, _i32, _i64, _i64, _p, _p, _p, _p, _p], _c.c_int),
    "qds_pattern_restrict_count": ([_p, _p, _i64, _p, _i64, _c.POINTER(_i64)], _c.c_int),
    "qds_pattern_restrict_fill": ([_p, _p, _i64, _p, _i64, _i64, _p, _p, _p], _c.c_int),
    "qds_assemble_elements": ([_i32, _i32, _i64, _p, _p, _p, _p,
                               _p, _i64, _i32, _p, _i64, _i32, _p, _i64, _i32, _p, _i64, _i32,
                               _p, _p, _p, _p], _c.c_int),
    "qds_assemble_gather": ([_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_quad_points": ([_i32, _i32, _i64, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_csr_take": ([_i32, _i64, _p, _p, _p, _d, _p, _p], _c.c_int),
    "qds_spmm": ([_i32, _i64, _i64, _i32, _p, _p, _p, _p, _p, _i64, _p, _i64, _p, _i64, _p], _c.c_int),
    "qds_spmm_cap": ([_i32, _i64, _i32, _p, _p, _p, _p, _p, _p, _i64, _p, _i64, _p, _i64, _p], _c.c_int),
    "qds_gram_workspace": ([_i64, _i32, _i32, _i32], _i64),
    "qds_gram": ([_i32, _i64, _i32, _p, _i64, _i32, _p, _i64, _i32, _p, _i64, _i32, _p, _p, _i64, _p],
                 _c.c_int),
    "qds_block_update": ([_i32, _i64, _i32, _i32, _p, _i64, _p, _p, _i64, _p, _p, _i64, _p], _c.c_int),
    "qds_gram_blocks": ([_i32, _i64, _p, _p, _p, _i32, _p, _p, _i64, _p], _c.c_int),
    "qds_gram_noalloc_scratch_bytes": ([], _i64),
    "qds_gram_blocks_noalloc": ([_i32, _i64, _p, _p, _p, _i32, _p, _p], _c.c_int),
    "qds_rr_status_len": ([_i32], _i32),
    "qds_rayleigh_ritz": ([_i64, _i32, _i32, _i32, _p, _i64, _p, _p, _p, _p, _i64, _p, _p, _d, _d, _i32, _p],
                          _c.c_int),
    "qds_lattice_spmm_batched": ([_i32, _p, _i32, _i32, _p, _i64, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, _p,
                                  _i64, _i64, _p], _c.c_int),
    "qds_lattice_spmm_tuned": ([_i32, _p, _i32, _i32, _p, _i64, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, _p,
                                _i64, _i64, _p, _p], _c.c_int),
    "qds_lattice_spmm_residual_workspace": ([_p, _i32], _i64),
    "qds_lattice_spmm_residual": ([_i32, _p, _i32, _p, _p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _i32, _p, _i64,
                                   _p, _p, _i64, _p, _p], _c.c_int),
    "qds_gram_blocks_batched": ([_i32, _i64, _p, _p, _p, _i32, _p, _i32, _i64, _p, _i64, _p], _c.c_int),
    "qds_block_update_blocks_batched": ([_i32, _i64, _p, _i32, _p, _p, _i64, _p, _p, _i64, _i32, _i32, _i64, _i64,
                                         _i64, _p], _c.c_int),
    "qds_residual_norms_batched_workspace": ([_i64, _i32, _i32], _i64),
    "qds_residual_norms_batched": ([_i32, _i64, _i32, _i32, _p, _i64, _p, _i64, _i64, _p, _p, _p, _i64, _p, _p,
                                    _i64, _p], _c.c_int),
    "qds_block_update_blocks": ([_i32, _i64, _p, _i32, _p, _p, _i64, _p, _p, _i64, _i32, _p], _c.c_int),
    "qds_block_residual": ([_i32, _i64, _i32, _p, _i64, _p, _i64, _p, _p, _i64, _p], _c.c_int),
    "qds_coldot_workspace": ([_i32], _i64),
    "qds_residual_norms_workspace": ([_i32], _i64),
    "qds_residual_norms": ([_i32, _i64, _i32, _p, _i64, _p, _i64, _p, _p, _p, _i64, _p, _p, _i64, _p],
                           _c.c_int),
    "qds_coldot": ([_i32, _i64, _i32, _p, _i64, _p, _i64, _i32, _p, _p, _i64, _p], _c.c_int),
    "qds_residual_norms_to": ([_i32, _i32, _i64, _i32, _p, _i64, _p, _i64, _p, _p, _p, _i64, _p, _p, _i64, _p],
                              _c.c_int),
    "qds_host_rayleigh_ritz": ([_i32, _i32, _p, _d, _p, _p], _c.c_int),
    "qds_gram_blocks_typed": ([_i64, _p, _p, _p, _p, _p, _p, _i32, _p, _p, _i64, _p], _c.c_int),
    "qds_block_update_typed": ([_i64, _p, _p, _i32, _p, _p, _i64, _p, _i32, _p, _i64, _i32, _p], _c.c_int),
    "qds_mg_vcycle": ([_i32, _i32, _p, _i32, _i32, _i32, _i32, _d, _p], _c.c_int),
    "qds_mg_vcycle_io": ([_i32, _i32, _i32, _p, _i32, _i32, _i32, _i32, _d, _p], _c.c_int),
    "qds_mg_vcycle_batched": ([_i32, _i32, _i32, _p, _p, _i32, _i32, _i32, _i32, _i32, _d, _p], _c.c_int),
    "qds_mg_apply": ([_i32, _p, _i32, _p, _p, _p], _c.c_int),
    "qds_interpolate_q": ([_i32, _i32, _i64, _p, _p, _p, _p, _p], _c.c_int),
    "qds_state_density_q": ([_i32, _i32, _i64, _p, _p, _p, _p, _p, _i64, _i32, _p, _p], _c.c_int),
    "qds_density": ([_i64, _i32, _p, _i64, _p, _p, _p], _c.c_int),
    "qds_interpolate_points": ([_i32, _i64, _i64, _d, _d, _p, _p, _p, _p, _p, _i32, _i64, _p, _p, _p, _p, _p,
                                _p, _p], _c.c_int),
    "qds_lattice_stencil_build": ([_i32, _i64, _i64, _p, _p, _p, _p], _c.c_int),
    "qds_lattice_spmm": ([_i32, _p, _i32, _p, _p, _p, _i64, _p, _i64, _p, _i64, _p], _c.c_int),
    "qds_cg_workspace": ([_i64], _i64),
    "qds_cg": ([_i64, _i64, _p, _p, _p, _p, _p, _d, _i32, _i32, _p, _i64, _c.POINTER(_i32),
                _c.POINTER(_d), _p], _c.c_int),
    "qds_gradient_indicator": ([_i32, _i64, _p, _p, _p, _p, _d, _p, _p], _c.c_int),
    "qds_cg_mg": ([_i64, _i64, _p, _p, _p, _p, _p, _d, _i32, _i32, _i32, _p, _i32, _i32, _d, _p, _i64,
                   _c.POINTER(_i32), _c.POINTER(_d), _p], _c.c_int),
    "qds_mg_lattice_coefficients": ([_i64, _i64, _i64, _i64, _d, _d, _d, _d, _p, _p, _d, _p, _p, _p, _p, _p],
                                    _c.c_int),
    "qds_mg_lattice_coarsen": ([_i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_mg_lattice_interior": ([_i32, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_bicgstab_workspace": ([_i64], _i64),
    "qds_bicgstab": ([_i64, _i64, _p, _p, _p, _p, _p, _d, _i32, _i32, _p, _i64, _c.POINTER(_i32),
                      _c.POINTER(_d), _c.POINTER(_i32), _p], _c.c_int),
    "qds_bicgstab_mg": ([_i64, _i64, _p, _p, _p, _p, _p, _d, _i32, _i32, _i32, _p, _i32, _i32, _d, _p, _i64,
                         _c.POINTER(_i32), _c.POINTER(_d), _c.POINTER(_i32), _p], _c.c_int),
    "qds_dd_carriers": ([_p, _i32, _d, _d, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_contacts": ([_p, _d, _d, _d, _d, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_initial": ([_p, _d, _d, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_charge": ([_i64, _d, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_workspace": ([], _i64),
    "qds_dd_potential_update": ([_i64, _i32, _d, _p, _p, _p, _p, _i64, _p], _c.c_int),
    "qds_dd_field": ([_p, _p, _p, _p], _c.c_int),
    "qds_dd_currents": ([_p, _d, _p, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_continuity": ([_p, _i32, _d, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p], _c.c_int),
    "qds_dd_poisson_system": ([_p, _i32, _d, _d, _p, _p, _p, _d, _d, _p, _p, _p], _c.c_int),
}


class ColBlocks(ctypes.Structure):
    """Mirror of qds_colblocks."""
    _fields_ = [("p", _p * 4), ("ld", _i64 * 4), ("cols", _i32 * 4), ("nb", _i32)]


MAX_CLASS, MAX_ENTRY = 9, 52


class LatticeStencil(ctypes.Structure):
    """Mirror of qds_lattice_stencil."""
    _fields_ = [("order", _i32), ("n_class", _i32), ("nx", _i64), ("ny", _i64), ("cells_x", _i64),
                ("cells_y", _i64), ("tile", _i32), ("pad0", _i32), ("tiles_x", _i64), ("tile_slots", _i64),
                ("n_slots", _i64), ("len", _i32 * MAX_CLASS), ("base", _i64 * MAX_CLASS),
                ("di", (_c.c_int8 * MAX_ENTRY) * MAX_CLASS), ("dj", (_c.c_int8 * MAX_ENTRY) * MAX_CLASS)]


class MGLevel(ctypes.Structure):
    """Mirror of qds_mg_level."""
    _fields_ = [("nx", _i64), ("ny", _i64), ("d", _p), ("we", _p), ("wn", _p),
                ("u", _p), ("f", _p), ("t", _p), ("ldu", _i64), ("ldf", _i64),
                ("neumann_x", _i32), ("neumann_y", _i32)]


class LatticeTuning(ctypes.Structure):
    """Mirror of qds_lattice_tuning."""
    _fields_ = [("values_in_lds", _i32), ("lds_budget", _i32)]


class MGLevelStrides(ctypes.Structure):
    """Mirror of qds_mg_level_strides."""
    _fields_ = [("sd", _i64), ("sw", _i64), ("su", _i64), ("sf", _i64), ("st", _i64)]


class DDMesh(ctypes.Structure):
    """Mirror of qds_dd_mesh."""
    _fields_ = [("n_nodes", _i64), ("n_elem", _i64), ("conn", _p), ("x", _p), ("y", _p), ("ne_ptr", _p),
                ("ne_elem", _p), ("indptr", _p), ("indices", _p), ("mat_idx", _p), ("emat_idx", _p), ("mat", _p),
                ("contact", _p)]


class QDSError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"qdsim_amd: native library {LIB_PATH} is missing; build it with "
            "`python qdsim_amd/_build.py` (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc, what=""):
    if rc != 0:
        msg = lib.qds_last_error().decode(errors="replace")
        raise QDSError(f"{what or 'qdsim_hip'} failed (rc={rc}): {msg}")


def host_ptr(a):
    """Pointer to a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
