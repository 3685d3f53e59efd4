"""Host-side logic of the product and the C ABI surface (no GPU calls)."""
import os
import re
import subprocess

import numpy as np
import pytest

import race_amd as R
from race_amd import _lib

from .conftest import ROOT


def test_library_exports_every_declared_entry_point():
    hdr = open(os.path.join(ROOT, "include", "race_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*|double)\s+\**(race_\w+)\s*\(", hdr, re.M))
    assert len(declared) > 30
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH]).decode()
    exported = set(re.findall(r"\sT\s(race_\w+)$", out, re.M))
    assert declared <= exported, declared - exported
    L = _lib.load()
    for name in declared:
        assert hasattr(L, name)
        assert name in _lib.PROTOTYPES, f"{name} has no ctypes prototype"


def test_error_strings_match_the_reference_enum():
    L = _lib.load()
    for code, name in _lib.STATUS_NAMES.items():
        assert L.race_error_string(code).decode() == name


@pytest.mark.parametrize("seed", range(300))
def test_product_load_balancer_matches_oracle(oracle, seed):
    rng = np.random.default_rng(seed)
    L = int(rng.integers(1, 260))
    lr = rng.integers(0, 2500, L)
    lr[0] = max(lr[0], 1)
    ln = lr * rng.integers(1, 30, L)
    T = int(rng.integers(1, 80))
    eff = float(rng.choice([0, 10, 40, 60, 80, 95]))
    dist = int(rng.choice([1, 2]))
    a = oracle.split_zones(lr, ln, T, eff, dist)
    b = R.lb_split(lr, ln, T, eff, dist)
    assert a["levelPtr"] == b["levelPtr"]
    assert a["scale"] == b["scale"]
    assert a["zone_rows"] == b["zone_rows"]


def test_lb_rejects_bad_arguments():
    with pytest.raises(R.RaceError):
        R.lb_split([], [], 4, 40.0, 1)
    with pytest.raises(R.RaceError):
        R.lb_split([1, 2], [3, 4], 0, 40.0, 1)


def test_bench_rejects_multi_gpu():
    p = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr


def test_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks that run before any device call (host only)."""
    import ctypes

    from race_amd import _lib as L
    lib = L.load()
    rp = np.array([0, 1, 2], np.int32)
    col = np.array([0, 1], np.int32)
    h = ctypes.c_void_p()
    ip = lambda a: a.ctypes.data_as(L.i32p)  # noqa: E731
    INVALID = 1
    # race_create: empty matrix, no thread, unknown distance, half a permutation
    assert lib.race_create(ctypes.byref(h), 0, 1, 1, ip(rp), ip(col), 1, 1, 1, None, None, 0, 1) == INVALID
    assert lib.race_create(ctypes.byref(h), 2, 0, 1, ip(rp), ip(col), 1, 1, 1, None, None, 0, 1) == INVALID
    assert lib.race_create(ctypes.byref(h), 2, 1, 7, ip(rp), ip(col), 1, 1, 1, None, None, 0, 1) == INVALID
    assert lib.race_create(ctypes.byref(h), 2, 1, 1, ip(rp), ip(col), 1, 1, 1, ip(rp), None, 0, 1) == INVALID
    # batches and the tuner's model
    assert lib.race_batch_create(ctypes.byref(h), 99, 0, 0, 1) == INVALID
    assert lib.race_batch_create(ctypes.byref(h), 2, 0, 0, 0) == INVALID
    assert lib.race_mlp_create(None, 0) == INVALID
    assert lib.race_mlp_forward(None, None, 1, None, 1) == INVALID
    # SELL helpers: missing arrays, unknown value type
    v = np.zeros(4)
    assert lib.race_simdify_resolve(4, 4, 4, None, ip(rp), ip(rp), v.ctypes.data, 0, 0) == INVALID
    assert lib.race_simdify_d1(4, 4, 4, ip(col), ip(rp), ip(rp), v.ctypes.data, 7, 0) == INVALID
    assert lib.race_set_option(None, b"layout", 0) == INVALID
