"""Device plumbing: torch tensors as device memory, streams, workspaces.

PyTorch-ROCm provides allocation and the current HIP stream; every computation
on those buffers is a qdsim_hip kernel.  A GPU is required.
"""
import ctypes
import threading

import torch

from . import _lib

DTYPE_CODE = {torch.float32: _lib.F32, torch.float64: _lib.F64,
              torch.complex64: _lib.C64, torch.complex128: _lib.C128}


_runtime_checked = False
_blas_ctl = None


def blas_single_thread():
    """Context manager keeping host BLAS / LAPACK on one thread (the block
    iteration's small dense solves lose to thread start-up otherwise).  The
    library scan of threadpoolctl runs once per process (~1-6 ms); entering the
    cached controller's limit costs microseconds, where a fresh
    ``threadpool_limits`` per solve cost ~1 ms."""
    global _blas_ctl
    if _blas_ctl is None:
        from threadpoolctl import ThreadpoolController
        _blas_ctl = ThreadpoolController()
    return _blas_ctl.limit(limits=1, user_api="blas")


def device():
    """The current GPU.  The first call also checks that the library's kernels run
    through the HIP runtime torch uses (qds_runtime_check: a probe kernel launched
    and read back), so that a process holding two HIP runtimes fails here with the
    runtime images named instead of at the first kernel."""
    global _runtime_checked
    if not torch.cuda.is_available():
        raise RuntimeError("qdsim_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    if not _runtime_checked:
        _lib.check(_lib.lib.qds_runtime_check(), "qds_runtime_check")
        _runtime_checked = True
    return torch.device("cuda", torch.cuda.current_device())


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_CUR_DEVICE = getattr(torch._C, "_cuda_getDevice", None)


def stream_ptr():
    """The calling thread's current HIP stream.  Every kernel wrapper asks for it,
    several times per block iteration: torch's raw accessor skips building a
    Stream object (the public call cost ~10 us each on the host)."""
    if _RAW_STREAM is not None and _CUR_DEVICE is not None:
        return ctypes.c_void_p(_RAW_STREAM(_CUR_DEVICE()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def sync():
    """Wait for the current stream (every kernel of the calling thread runs there;
    other threads' streams keep running)."""
    torch.cuda.current_stream().synchronize()


def ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def code(t):
    return DTYPE_CODE[t.dtype]


def real_dtype(dt):
    return {torch.complex64: torch.float32, torch.complex128: torch.float64}.get(dt, dt)


class Workspace:
    """Grow-only scratch buffer (bytes) reused across calls on one device."""

    def __init__(self):
        self._buf = None

    def get(self, nbytes):
        nbytes = max(int(nbytes), 256)
        if self._buf is None or self._buf.numel() < nbytes:
            self._buf = torch.empty(nbytes, dtype=torch.uint8, device=device())
        return self._buf

    def ptr(self, nbytes):
        b = self.get(nbytes)
        return ctypes.c_void_p(b.data_ptr()), b.numel()


_WS = threading.local()


def workspace(name="default"):
    """Named scratch buffer of the calling thread: threads that run independent
    problems on their own streams never share scratch memory."""
    d = getattr(_WS, "d", None)
    if d is None:
        d = _WS.d = {}
    if name not in d:
        d[name] = Workspace()
    return d[name]
