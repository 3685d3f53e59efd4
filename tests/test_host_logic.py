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


@pytest.mark.parametrize("T,dist,m", [(2, 1, 4), (3, 1, 3), (4, 2, 2), (5, 2, 3), (8, 1, 2), (16, 1, 3)])
def test_product_load_balancer_uniform_levels_known_answer(T, dist, m):
    """Equal nnz on every level, T x blocks-per-thread x m levels: the zones
    and blocks stay the uniform split (zero variance from the start; the
    same known answer pins the oracle in test_oracle.py)."""
    bpt = 2  # getBlockPerThread for distance 1 and for distance 2 with two blocks
    L = T * bpt * m
    b = R.lb_split(np.full(L, 7), np.full(L, 21), T, 40.0, dist)
    assert list(b["levelPtr"]) == [k * m for k in range(T * bpt + 1)]
    assert list(b["scale"]) == [1] * T


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
    # RACE_MAX_DENSE has no serial part for POWER (checked before any device call)
    os.environ["RACE_MAX_DENSE"] = "4"
    try:
        assert lib.race_create(ctypes.byref(h), 2, 1, 3, ip(rp), ip(col), 1, 1, 1, None, None, 0, 1) == 10
    finally:
        del os.environ["RACE_MAX_DENSE"]
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


def _fnv(data, h=1469598103934665603):
    for c in data:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _schedule_bytes(n=6, perm=None, leaves=None, dist=1, checksum=True, mpk=None):
    """A schedule byte string written field by field: header, permutation,
    zone tree (root with two threads of two colours each), checksum."""
    import struct
    perm = list(range(n))[::-1] if perm is None else perm
    inv = [0] * n
    for i, p in enumerate(perm):
        if 0 <= p < n:
            inv[p] = i
    if leaves is None:
        # root [0, n) split into 4 blocks: thread 0 = leaves 1, 2; thread 1 = 3, 4
        q = [0, 2, 3, 5, n]
        leaves = [(2, 2, 1, -1, n, 0, [0, n], [1, 5])]
        leaves += [(1, 1, 1, 0, q[i + 1] - q[i], 1, [q[i], q[i + 1]], []) for i in range(4)]
    vec = lambda v: struct.pack("<i", len(v)) + struct.pack(f"<{len(v)}i", *v)  # noqa: E731
    b = b"RACESCHD" + struct.pack("<I", 1) + struct.pack("<8i", n, dist, 0, 1, 2, 2, 1, 1)
    b += struct.pack("<Q", 12345) + vec(perm) + vec(inv)
    if mpk is None:
        b += struct.pack("<ii", 1, len(leaves))
        for lf in leaves:
            b += struct.pack("<7i", *lf[:6], 0) + vec(lf[6]) + vec(lf[7])
        b += struct.pack("<i", 0)
    else:
        # level groups: levelPtr, groupPtr, groupLevel, levelNnz, hopeless
        b += struct.pack("<ii", 0, 1) + b"".join(vec(v) for v in mpk) + struct.pack("<ddi", 1e6, 20.0, 2)
    return b + struct.pack("<Q", _fnv(b) if checksum else 0)


def test_schedule_bytes_are_checked_without_a_device():
    """race_schedule_info parses and range-checks a saved schedule on the host."""
    ok = _schedule_bytes()
    assert R.schedule_info(ok) == {"nrow": 6, "dist": 1, "nthreads": 2}
    bad = [
        _schedule_bytes(checksum=False),                       # damaged
        ok[:-9] + ok[-8:],                                     # truncated
        b"RACESCHX" + ok[8:],                                  # magic (checksum fails too)
        _schedule_bytes(perm=[0, 0, 2, 3, 4, 5]),              # not a permutation
        _schedule_bytes(leaves=[(2, 2, 1, -1, 6, 0, [0, 6], [0, 4])]),  # child before parent
        _schedule_bytes(leaves=[(2, 2, 1, -1, 6, 0, [0, 6], [1, 9])]),  # child out of range
        _schedule_bytes(leaves=[(1, 1, 1, -1, 6, 0, [0, 7], [])]),      # rows beyond nrow
        _schedule_bytes(leaves=[(2, 2, 1, -1, 6, 0, [0, 6], [1, 5]),    # rows 5 not covered
                                (1, 1, 1, 0, 2, 1, [0, 2], []), (1, 1, 1, 0, 1, 1, [2, 3], []),
                                (1, 1, 1, 0, 2, 1, [3, 5], []), (1, 1, 1, 0, 0, 1, [5, 5], [])]),
        _schedule_bytes(leaves=[(2, 2, 1, -1, 6, 0, [0, 6], [1, 5]),    # row 1 twice
                                (1, 1, 1, 0, 2, 1, [0, 2], []), (1, 1, 1, 0, 1, 1, [1, 3], []),
                                (1, 1, 1, 0, 2, 1, [3, 5], []), (1, 1, 1, 0, 1, 1, [5, 6], [])]),
        _schedule_bytes(dist=3),                               # POWER without level groups
        _schedule_bytes(dist=3, mpk=[[0, 1, 3, 6], [0, 3, 6], [0, 1, 3], [4, 8, 12], []]),  # group 1 at row 3
        _schedule_bytes(dist=3, mpk=[[0, 1, 3, 6], [0, 1, 6], [0, 1, 3], [4, 8, 12], [2]]),  # hopeless level
        _schedule_bytes(dist=1, mpk=[[0, 1, 3, 6], [0, 1, 6], [0, 1, 3], [4, 8, 12], []]),   # groups, not POWER
    ]
    power = _schedule_bytes(dist=3, mpk=[[0, 1, 3, 6], [0, 1, 6], [0, 1, 3], [4, 8, 12], []])
    assert R.schedule_info(power) == {"nrow": 6, "dist": 3, "nthreads": 2}
    for i, b in enumerate(bad):
        with pytest.raises(R.RaceError):
            R.schedule_info(b)
            pytest.fail(f"case {i} accepted")


def test_schedule_decoder_survives_resealed_corruption():
    """Random field changes in a valid schedule, checksum recomputed so the
    content checks (not the checksum) must catch them: every byte string is
    either accepted or refused with RACE_ERR_INVALID_ARG, never crashes,
    and every accepted one still describes a valid permutation."""
    import struct
    base = bytearray(_schedule_bytes()[:-8])
    rng = np.random.default_rng(12)
    accepted = 0
    for _ in range(400):
        b = bytearray(base)
        for _ in range(int(rng.integers(1, 4))):
            pos = 12 + 4 * int(rng.integers(0, (len(b) - 12) // 4))
            val = int(rng.choice([-1, 0, 1, 2, 3, 5, 6, 7, 1 << 20, -(1 << 30), int(rng.integers(-8, 16))]))
            b[pos:pos + 4] = struct.pack("<i", val)
        blob = bytes(b) + struct.pack("<Q", _fnv(bytes(b)))
        try:
            info = R.schedule_info(blob)
        except R.RaceError as e:
            assert e.status == 1
            continue
        accepted += 1
        assert info["nrow"] > 0
    assert accepted < 400


def test_schedule_decoder_under_sanitizers(tmp_path):
    """tests/cpp/schedule_fuzz.cpp: 100000 corrupted, re-sealed schedules
    through the decoder built for the host with AddressSanitizer and UBSan
    (g++ only; no device code is involved)."""
    import shutil
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    csrc = os.path.join(ROOT, "race_amd", "csrc")
    exe = str(tmp_path / "schedule_fuzz")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", csrc, "-I", "/opt/rocm/include",
                           os.path.join(ROOT, "tests", "cpp", "schedule_fuzz.cpp"),
                           os.path.join(csrc, "schedule_io.cpp"), os.path.join(csrc, "engine.cpp"), "-o", exe],
                          timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "accepted" in out.stdout


def test_flow_dependency_tables_under_sanitizers(tmp_path):
    """tests/cpp/flow_deps.cpp: the one-launch sweeps' dependency tables and
    coherent-read bits (chainDeps, flowExt) against conflicts computed
    directly from a 5-point grid coloured by lines, GS and KACZ, forward and
    backward, every line a chain of several bundles; with AddressSanitizer
    and UBSan (g++ only, no device code)."""
    import shutil
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    csrc = os.path.join(ROOT, "race_amd", "csrc")
    exe = str(tmp_path / "flow_deps")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", csrc, "-I", os.path.join(ROOT, "include"),
                           "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "flow_deps.cpp"),
                           os.path.join(csrc, "bundle.cpp"), os.path.join(csrc, "engine.cpp"), "-o", exe],
                          timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "flow deps ok" in out.stdout
