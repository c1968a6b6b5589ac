"""qds_gram_blocks_noalloc (static device scratch, no workspace argument) against
qds_gram_blocks (caller workspace), both called through the C-ABI.

Two calls in a row, the second on a smaller input, so that the second runs over
a scratch array that still holds the first call's partial sums.  The slabs are
the same for shapes whose partials fit the array at 512 slabs, so the results
must be bitwise equal there; a wider Gram uses fewer slabs and is compared to
1e-12 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _blocks(n, widths, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return [torch.rand((n, w), generator=g, device="cuda", dtype=torch.float64).to(dtype) - 0.5
            for w in widths]


def _gram_pair(n, widths, dtype, seed):
    from qdsim_amd import _lib, kernels
    from qdsim_amd.device import code, stream_ptr
    S = _blocks(n, widths, dtype, seed)
    HS = _blocks(n, widths, dtype, seed + 1)
    MS = _blocks(n, widths, dtype, seed + 2)
    m = sum(widths)
    ca, c1, c2 = kernels._colblocks(S), kernels._colblocks(HS), kernels._colblocks(MS)
    ref = torch.full((m, 2 * m), np.nan, dtype=torch.float64, device="cuda")
    out = torch.full((m, 2 * m), np.nan, dtype=torch.float64, device="cuda")
    wsb = _lib.lib.qds_gram_workspace(n, m, m, m)
    ws = torch.empty(wsb // 8 + 1, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib.qds_gram_blocks(code(S[0]), n, ctypes.byref(ca), ctypes.byref(c1), ctypes.byref(c2), 0,
                                        ref.data_ptr(), ws.data_ptr(), wsb, stream_ptr()), "qds_gram_blocks")
    _lib.check(_lib.lib.qds_gram_blocks_noalloc(code(S[0]), n, ctypes.byref(ca), ctypes.byref(c1),
                                                ctypes.byref(c2), 0, out.data_ptr(), stream_ptr()),
               "qds_gram_blocks_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_gram_matches_workspace_gram_twice(dtype):
    print(f"no-alloc Gram: {dtype}, n=300000 then 40000", flush=True)
    # first call: the LOBPCG shape [X | W | P] of 24 columns each (72 x 144 Gram)
    ref1, out1 = _gram_pair(300_000, [24, 24, 24], dtype, 11)
    assert np.isfinite(ref1).all()
    assert np.array_equal(ref1, out1)
    # second call, smaller: fewer rows and columns, over the first call's partials
    ref2, out2 = _gram_pair(40_000, [16, 16], dtype, 23)
    assert np.isfinite(ref2).all()
    assert np.array_equal(ref2, out2)


def test_noalloc_gram_wide_uses_fewer_slabs():
    print("no-alloc Gram: 96 x 192, n=200000", flush=True)
    # padded 96 x 192 partials: 455 slabs fit the array instead of 512
    ref, out = _gram_pair(200_000, [24, 24, 24, 24], torch.float32, 5)
    scale = np.abs(ref).max()
    assert np.abs(out - ref).max() <= 1e-12 * scale


def test_noalloc_gram_rejects_oversized_shape():
    from qdsim_amd import _lib, kernels
    from qdsim_amd.device import code, stream_ptr
    # four blocks of 72 columns: padded 288 x 576 partials exceed the array at 64 slabs;
    # rejected on the host before any launch
    n = 1024
    S = [torch.zeros((n, 72), dtype=torch.float32, device="cuda") for _ in range(4)]
    ca = kernels._colblocks(S)
    out = torch.zeros((288, 576), dtype=torch.float64, device="cuda")
    rc = _lib.lib.qds_gram_blocks_noalloc(code(S[0]), n, ctypes.byref(ca), ctypes.byref(ca), ctypes.byref(ca), 0,
                                          out.data_ptr(), stream_ptr())
    assert rc == -1
    assert b"static scratch" in _lib.lib.qds_last_error()
    assert _lib.lib.qds_gram_noalloc_scratch_bytes() == 64 << 20
