"""Host logic of the process-pool sweep driver (no GPU): results keep the order
of the items, a worker's exception reaches the caller, an empty batch starts no
process."""
import operator

import pytest

from qdsim_amd.batch import run_processes


def test_run_processes_keeps_order():
    items = [-5, 3, -1, 0, 7, -2, 4]
    assert run_processes(operator.neg, items, 3) == [5, -3, 1, 0, -7, 2, -4]


def test_run_processes_more_workers_than_items():
    assert run_processes(abs, [-2], 8) == [2]


def test_run_processes_empty():
    assert run_processes(abs, [], 4) == []


def test_run_processes_reraises():
    with pytest.raises(ValueError):
        run_processes(int, ["1", "x", "3"], 2)


def test_run_processes_progress():
    seen = []
    out = run_processes(abs, [-1, -2, -3], 2, progress=lambda d, t: seen.append((d, t)))
    assert out == [1, 2, 3]
    assert seen == [(1, 3), (2, 3), (3, 3)]
