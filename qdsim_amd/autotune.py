"""Autotuner of the lattice product's block shape (qds_lattice_tuning).

No reference counterpart.  The lattice product (qds_lattice_spmm_tuned) has two
launch-shape choices that do not change its arithmetic: whether each block's
value runs are staged in LDS with its X rows or read from HBM by the compute
phase, and the LDS budget a block is sized for (the cells per workgroup, and with
it how many workgroups share a CU).  Every candidate computes the same sums in the
same order, so the choice only changes time.  On the first product of an operator
shape (order, lattice size, dtype, block width, fused or not, batch) the tuner
times every candidate with HIP events on the stream the product runs on, with the
operator's own values (after a warm-up, in two interleaved rounds, the faster
round of each), and keeps the fastest for the process; the choices are also
written to a JSON file (``QDSIM_AMD_TUNE_CACHE``, default
~/.cache/qdsim_amd/lattice_tuning.json) keyed by the device and the shape, and read
back by later processes.  ``QDSIM_AMD_TUNE=0`` skips the timing and uses the
default shape, as do products below MIN_TUNED_ELEMENTS (n b batch) elements.
"""
import json
import os
import threading

import torch

from . import _lib

# (values staged in LDS, LDS budget in bytes); the first is the untuned default
CANDIDATES = [(1, 52 * 1024), (0, 52 * 1024), (0, 40 * 1024), (0, 32 * 1024), (0, 24 * 1024),
              (0, 64 * 1024), (1, 40 * 1024)]
DEFAULT = CANDIDATES[0]
# n * b * batch below which the default shape is used untimed: a sweep of uneven
# meshes (256 self-consistent instances, 75 shapes of up to 591k DOF x 16
# columns, eight worker processes) spent ~9 of 78 s tuning them
MIN_TUNED_ELEMENTS = 16 << 20

_lock = threading.Lock()
_choices = {}
_file_loaded = False


def _cache_path():
    return os.environ.get("QDSIM_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "qdsim_amd", "lattice_tuning.json"))


_tags = {}


def _device_tag():
    """Architecture and CU count of the current device, queried once per device
    under the lock: the batch drivers tune from several host threads at once, and
    concurrent property queries failed there ("Invalid device id")."""
    idx = torch.cuda.current_device()
    tag = _tags.get(idx)
    if tag is None:
        with _lock:
            if idx not in _tags:
                p = torch.cuda.get_device_properties(idx)
                _tags[idx] = f"{p.gcnArchName}/{p.multi_processor_count}"
            tag = _tags[idx]
    return tag


def _key(stencil, dtype, b, fused, batch):
    return f"{_device_tag()}|p{stencil.order}|{stencil.nx}x{stencil.ny}|{str(dtype)[6:]}|b{b}|f{int(fused)}|n{batch}"


def _read_file():
    try:
        with open(_cache_path()) as f:
            data = json.load(f)
        return {k: (int(v[0]), int(v[1])) for k, v in data.items()}
    except (OSError, ValueError, TypeError, IndexError, AttributeError):
        return {}


def _load_file(force=False):
    """Merge the cache file's choices (force: read it again -- the worker processes
    of a sweep tune shapes concurrently and share them through the file)."""
    global _file_loaded
    if _file_loaded and not force:
        return
    _file_loaded = True
    for k, v in _read_file().items():
        _choices.setdefault(k, v)


_untuned = {}


def _save_file():
    """Write this process's tuned choices merged into the file's (atomic replace)."""
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _read_file()
        data.update({k: v for k, v in _choices.items() if k not in _untuned})
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump({k: list(v) for k, v in data.items()}, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def to_struct(choice):
    t = _lib.LatticeTuning()
    t.values_in_lds, t.lds_budget = int(choice[0]), int(choice[1])
    return t


def choice_for(stencil, dtype, b, fused, batch, run):
    """The tuned (values_in_lds, lds_budget) for this product shape.  ``run(choice)``
    launches the product once with that choice on the current stream (it must not
    synchronise); the first call for a shape times every candidate."""
    key = _key(stencil, dtype, b, fused, batch)
    with _lock:
        _load_file()
        if key in _choices:
            return _choices[key]
    if os.environ.get("QDSIM_AMD_TUNE", "1") == "0":
        return DEFAULT
    if int(stencil.nx) * int(stencil.ny) * int(b) * int(batch) < MIN_TUNED_ELEMENTS:
        # small products: the shapes differ by a few percent and ~100 tuning
        # launches per shape cost more than they save (remembered in this process
        # only, so that every later product skips this test)
        with _lock:
            _untuned[key] = DEFAULT
            _choices.setdefault(key, DEFAULT)
        return DEFAULT
    with _lock:
        # another worker process may have tuned this shape meanwhile
        _load_file(force=True)
        if key in _choices:
            return _choices[key]
    st = torch.cuda.current_stream()
    valid = []
    for cand in CANDIDATES:
        try:
            run(cand)                      # a shape the kernel rejects raises here
            valid.append(cand)
        except _lib.QDSError:
            continue
    # the first launches of a product run while the GPU's clocks ramp up (the first
    # candidate timed cold came out 15% slower than in steady state): warm up on the
    # default shape, then time the candidates in two interleaved rounds and keep
    # each one's faster round
    for _ in range(8 if valid else 0):
        run(valid[0])
    times = {}
    for _ in range(2):
        for cand in valid:
            run(cand)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(5):
                run(cand)
            e1.record(st)
            e1.synchronize()
            t = e0.elapsed_time(e1) / 5
            times[cand] = min(times.get(cand, t), t)
    best = min(times, key=times.get) if times else DEFAULT
    with _lock:
        _choices[key] = best
        _timings[key] = {f"{c[0]}:{c[1]}": t for c, t in times.items()}
        _save_file()
    return best


_timings = {}


def report():
    """Chosen shapes and the candidate times (ms) of the shapes tuned in this process."""
    return {k: {"choice": list(v), "times_ms": _timings.get(k)} for k, v in _choices.items()}
