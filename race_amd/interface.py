"""Python mirror of RACE::Interface over the C ABI (include/race_amd.h).

Method names follow the reference (include/interface.h:54-205) so tests read
like the reference's drivers:

    ce = Interface(nrow, nthreads, TWO, rowPtr, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    fid = ce.registerDevice(K_SYMMSPMV, rowPtr_u, col_u, val_u, x_dev, b_dev)
    ce.executeFunction(fid)
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import (BF16, FP32, FP64, K_GS, K_KACZ, K_MPK, K_MPK_HIGHEST, K_SPMTV, K_SPMV, K_SYMMSPMV, LB_NNZ,
                   LB_ROW, ONE, POWER, SCATTER, TWO, TWO_BLOCK, RaceError)

__all__ = ["Interface", "DeviceArray", "lb_split", "split_threads", "mpk_hopeless_regions", "mpk_consolidate", "RaceError", "ONE", "TWO", "POWER", "K_SYMMSPMV",
           "K_GS", "K_KACZ", "K_SPMV", "K_MPK", "K_SPMTV", "K_MPK_HIGHEST", "FP64", "FP32", "BF16", "LB_NNZ", "LB_ROW",
           "TWO_BLOCK", "device_count", "schedule_info", "rcm"]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ip(a):
    return a.ctypes.data_as(L.i32p) if a is not None else None


def device_count():
    c = ctypes.c_int(0)
    L.load().race_device_count(ctypes.byref(c))
    return c.value


class DeviceArray:
    """Device allocation made through the C ABI (no torch needed)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        L.check(L.load().race_device_malloc(ctypes.byref(p), self.nbytes), "race_device_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or np.asarray(a).dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        L.check(L.load().race_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes), "race_memcpy_h2d")

    def download(self):
        out = np.empty(self.n, self.dtype)
        L.check(L.load().race_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "race_memcpy_d2h")
        return out

    def zero(self):
        L.check(L.load().race_memset_device(self.ptr, 0, self.nbytes), "race_memset_device")

    def free(self):
        if self.ptr:
            L.load().race_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Interface:
    """RACE::Interface(nrow, nthreads, dist, rowPtr, col, symm_hint, SMT,
    pinMethod, initPerm, initInvPerm, d2Type, lbTarget)."""

    def __init__(self, nrow, nthreads, dist, rowPtr, col, symm_hint=False, SMT=1, pinMethod=SCATTER,
                 initPerm=None, initInvPerm=None, d2Type=TWO_BLOCK, lbTarget=LB_NNZ):
        self._L = L.load()
        self.nrow = int(nrow)
        self._keep = []
        self._vecs = {}
        rp, c = _i32(rowPtr), _i32(col)
        ip = _i32(initPerm) if initPerm is not None else None
        iip = _i32(initInvPerm) if initInvPerm is not None else None
        h = ctypes.c_void_p()
        L.check(self._L.race_create(ctypes.byref(h), self.nrow, int(nthreads), int(dist), _ip(rp), _ip(c),
                                    int(bool(symm_hint)), int(SMT), int(pinMethod), _ip(ip), _ip(iip),
                                    int(d2Type), int(lbTarget)), "race_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.race_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- preprocessing ----
    def RACEColor(self, highestPower=None, cacheSize=None, safetyFactor=2.0, mtxType="N",
                  highestSubPower=1):
        if highestPower is None:
            L.check(self._L.race_color(self.h), "RACEColor")
        else:
            L.check(self._L.race_color_power(self.h, int(highestPower), float(cacheSize),
                                             float(safetyFactor), mtxType.encode()[:1],
                                             int(highestSubPower)), "RACEColor(power)")

    def getPerm(self):
        out = np.empty(self.nrow, np.int32)
        L.check(self._L.race_get_perm(self.h, _ip(out), self.nrow), "getPerm")
        return out

    def getInvPerm(self):
        out = np.empty(self.nrow, np.int32)
        L.check(self._L.race_get_inv_perm(self.h, _ip(out), self.nrow), "getInvPerm")
        return out

    def getNumThreads(self):
        v = ctypes.c_int()
        L.check(self._L.race_get_num_threads(self.h, ctypes.byref(v)), "getNumThreads")
        return v.value

    def getEfficiency(self):
        v = ctypes.c_double()
        L.check(self._L.race_get_efficiency(self.h, ctypes.byref(v)), "getEfficiency")
        return v.value

    def getMaxStageDepth(self):
        v = ctypes.c_int()
        L.check(self._L.race_get_max_stage_depth(self.h, ctypes.byref(v)), "getMaxStageDepth")
        return v.value

    def getHighestPower(self):
        v = ctypes.c_int()
        L.check(self._L.race_get_highest_power(self.h, ctypes.byref(v)), "getHighestPower")
        return v.value

    # ---- registration / execution ----
    def registerFunction(self, f, args=None, power=None, subPower=1, numaSplit=0):
        """Host callback: f(start, end, args) for colouring, or
        f(start, end, pow, subPow, numa, args) for POWER."""
        fid = ctypes.c_int()
        if power is None:
            cb = L.COLOR_FN(lambda s, e, a: f(s, e, args))
            L.check(self._L.race_register_function(self.h, cb, None, ctypes.byref(fid)), "registerFunction")
        else:
            cb = L.POWER_FN(lambda s, e, p, sp, nu, a: f(s, e, p, sp, nu, args))
            L.check(self._L.race_register_power_function(self.h, cb, None, int(power), int(subPower),
                                                         int(numaSplit), ctypes.byref(fid)),
                    "registerFunction(power)")
        self._keep.append(cb)
        return fid.value

    def attachCommunicationToFunction(self, fid, f, args=None, simple=False):
        """Interface::attachCommunicationToFunction: f(pow, subPow, args), or
        f(pow, args) with simple=True.  Stored, never called on one device
        (no remote rows), as in the reference's non-MPI build."""
        if simple:
            cb = L.COMM_SIMPLE_FN(lambda p, a: f(p, args))
            st = self._L.race_attach_communication_simple(self.h, int(fid), cb, None)
        else:
            cb = L.COMM_FN(lambda p, sp, a: f(p, sp, args))
            st = self._L.race_attach_communication(self.h, int(fid), cb, None)
        L.check(st, "attachCommunicationToFunction")
        self._keep.append(cb)

    def registerDevice(self, kind, rowPtr, col, val, x, b=None, nrhs=1, power=1,
                       val_precision=FP64, vec_precision=FP64, subPower=1):
        rp, c = _i32(rowPtr), _i32(col)
        v = np.ascontiguousarray(val, dtype=np.float64)
        d = L.KernelDesc(kind=int(kind), nrow=self.nrow, rowPtr=_ip(rp), col=_ip(c),
                         val=v.ctypes.data_as(L.f64p), val_precision=int(val_precision),
                         vec_precision=int(vec_precision), nrhs=int(nrhs),
                         x=x.ptr if isinstance(x, DeviceArray) else x,
                         b=(b.ptr if isinstance(b, DeviceArray) else b), power=int(power),
                         sub_power=int(subPower))
        fid = ctypes.c_int()
        L.check(self._L.race_register_device(self.h, ctypes.byref(d), ctypes.byref(fid)), "registerDevice")
        # device arrays stay alive while a registered function may use them
        self._vecs[fid.value] = (x, b)
        return fid.value

    def registerRefinement(self, kind, rowPtr, col, val, x, b, inner_id, sweeps=1, nrhs=1):
        """FP64 iterative refinement around the registered function inner_id;
        (rowPtr, col, val) is the full FP64 matrix, x and b FP64 device vectors."""
        rp, c = _i32(rowPtr), _i32(col)
        v = np.ascontiguousarray(val, dtype=np.float64)
        d = L.KernelDesc(kind=int(kind), nrow=self.nrow, rowPtr=_ip(rp), col=_ip(c),
                         val=v.ctypes.data_as(L.f64p), val_precision=FP64, vec_precision=FP64,
                         nrhs=int(nrhs), x=x.ptr if isinstance(x, DeviceArray) else x,
                         b=b.ptr if isinstance(b, DeviceArray) else b, power=1)
        fid = ctypes.c_int()
        L.check(self._L.race_register_refinement(self.h, ctypes.byref(d), int(inner_id), int(sweeps),
                                                 ctypes.byref(fid)), "registerRefinement")
        self._vecs[fid.value] = (x, b)
        return fid.value

    def refine(self, fid, tol, max_outer=1000):
        """run refinement steps to ||b-Ax||/||b|| <= tol: (steps, inner sweeps, relres)"""
        o, sw, rr = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
        L.check(self._L.race_refine(self.h, fid, float(tol), int(max_outer), ctypes.byref(o),
                                    ctypes.byref(sw), ctypes.byref(rr)), "refine")
        return o.value, sw.value, rr.value

    def residual(self, fid):
        rr = ctypes.c_double()
        L.check(self._L.race_residual(self.h, fid, ctypes.byref(rr)), "residual")
        return rr.value

    def setOption(self, key, value):
        L.check(self._L.race_set_option(self.h, key.encode(), int(value)), f"setOption({key})")

    def setVectors(self, fid, x, b=None):
        L.check(self._L.race_set_vectors(self.h, fid, x.ptr if isinstance(x, DeviceArray) else x,
                                         b.ptr if isinstance(b, DeviceArray) else b), "setVectors")
        self._vecs[fid] = (x, b)

    def executeFunction(self, fid, rev=False):
        L.check(self._L.race_execute(self.h, int(fid), int(bool(rev))), "executeFunction")

    def executeAsync(self, fid, rev=False, stream=None):
        L.check(self._L.race_execute_async(self.h, int(fid), int(bool(rev)), stream), "executeAsync")

    def setPower(self, fid, p):
        L.check(self._L.race_set_power(self.h, fid, int(p)), "setPower")

    def getPower(self, fid):
        v = ctypes.c_int()
        L.check(self._L.race_get_power(self.h, fid, ctypes.byref(v)), "getPower")
        return v.value

    def setSubPower(self, fid, sp):
        L.check(self._L.race_set_sub_power(self.h, fid, int(sp)), "setSubPower")

    def getSubPower(self, fid):
        v = ctypes.c_int()
        L.check(self._L.race_get_sub_power(self.h, fid, ctypes.byref(v)), "getSubPower")
        return v.value

    def setSerial(self, fid):
        L.check(self._L.race_set_serial(self.h, fid, 1), "setSerial")

    def unsetSerial(self, fid):
        L.check(self._L.race_set_serial(self.h, fid, 0), "unsetSerial")

    def resetTime(self):
        L.check(self._L.race_reset_time(self.h), "resetTime")

    def sleep(self):
        L.check(self._L.race_sleep(self.h), "sleep")

    def pinThread(self, threadId):
        L.check(self._L.race_pin_thread(self.h, int(threadId)), "pinThread")

    def numaInitRowPtr(self, rowPtr):
        """first-touch zeroing of rowPtr (an int32 numpy array, modified in place)"""
        L.check(self._L.race_numa_init_row_ptr(self.h, _ip(rowPtr)), "numaInitRowPtr")

    def numaInitMtxVec(self, rowPtr, col, val, x=None, power=1):
        """first-touch zeroing of col, val (and x[0:2n]) in place"""
        xp = None if x is None else x.ctypes.data_as(L.f64p)
        L.check(self._L.race_numa_init_mtx_vec(self.h, _ip(rowPtr), _ip(col), val.ctypes.data_as(L.f64p), xp,
                                               int(power)), "numaInitMtxVec")

    def getNumaSplitting(self):
        split = np.zeros(2, np.int32)
        n = ctypes.c_int()
        L.check(self._L.race_get_numa_splitting(self.h, _ip(split), 2, ctypes.byref(n)), "getNumaSplitting")
        return split[:n.value]

    def tuneFunction(self, fid, rev=False):
        v = ctypes.c_int()
        L.check(self._L.race_tune_function(self.h, fid, int(bool(rev)), ctypes.byref(v)), "tuneFunction")
        return v.value

    def checkPowerCoverage(self):
        return self._L.race_check_power_coverage(self.h) == L.SUCCESS

    # ---- introspection ----
    def schedule(self, rev=False):
        nph, nl = ctypes.c_int(), ctypes.c_int()
        L.check(self._L.race_get_schedule(self.h, int(rev), ctypes.byref(nph), ctypes.byref(nl), None,
                                          None, None, 0), "schedule")
        cap = max(nph.value + 1, nl.value)
        pp = np.empty(nph.value + 1, np.int32)
        lo = np.empty(nl.value, np.int32)
        hi = np.empty(nl.value, np.int32)
        L.check(self._L.race_get_schedule(self.h, int(rev), ctypes.byref(nph), ctypes.byref(nl), _ip(pp),
                                          _ip(lo), _ip(hi), cap), "schedule")
        return pp, lo, hi

    def tree(self):
        n = ctypes.c_int()
        L.check(self._L.race_get_tree(self.h, ctypes.byref(n), None, 0), "tree")
        d = np.empty(6 * n.value, np.int32)
        L.check(self._L.race_get_tree(self.h, ctypes.byref(n), _ip(d), n.value), "tree")
        return d.reshape(n.value, 6)

    def mpkChildren(self):
        """recursion-stage leaves: list of (lo, hi, stage-0 group, own groups)"""
        n = ctypes.c_int()
        L.check(self._L.race_get_mpk_children(self.h, ctypes.byref(n), None, 0), "mpkChildren")
        d = np.zeros(4 * max(n.value, 1), np.int32)
        L.check(self._L.race_get_mpk_children(self.h, ctypes.byref(n), _ip(d), n.value), "mpkChildren")
        return [tuple(int(v) for v in d[4 * i:4 * i + 4]) for i in range(n.value)]

    def mpkGroups(self):
        n = ctypes.c_int()
        L.check(self._L.race_get_mpk_groups(self.h, ctypes.byref(n), None, 0), "mpkGroups")
        g = np.empty(n.value + 1, np.int32)
        L.check(self._L.race_get_mpk_groups(self.h, ctypes.byref(n), _ip(g), n.value + 1), "mpkGroups")
        return g

    def saveSchedule(self):
        """What RACEColor computed, as bytes (race_schedule_save)."""
        n = ctypes.c_size_t()
        L.check(self._L.race_schedule_save(self.h, None, 0, ctypes.byref(n)), "saveSchedule")
        buf = ctypes.create_string_buffer(n.value)
        L.check(self._L.race_schedule_save(self.h, buf, n.value, ctypes.byref(n)), "saveSchedule")
        return buf.raw[:n.value]

    def loadSchedule(self, data):
        """Instead of RACEColor: a schedule saved from a handle with the same
        matrix pattern and constructor arguments (race_schedule_load)."""
        data = bytes(data)
        L.check(self._L.race_schedule_load(self.h, data, len(data)), "loadSchedule")

    def launchCount(self, fid, rev=False):
        """kernel launches one execution of a device function issues"""
        v = ctypes.c_int()
        L.check(self._L.race_get_launch_count(self.h, fid, int(rev), ctypes.byref(v)), "launchCount")
        return v.value

    def flowAborts(self, fid):
        """one-launch sweeps whose dependency waits gave up (0 when the
        schedule is consistent); synchronises with the device"""
        v = ctypes.c_int()
        L.check(self._L.race_get_flow_aborts(self.h, fid, ctypes.byref(v)), "flowAborts")
        return v.value

    def waveCount(self, fid, rev=False):
        v = ctypes.c_int()
        L.check(self._L.race_get_wave_count(self.h, fid, int(rev), ctypes.byref(v)), "waveCount")
        return v.value

    def stream(self):
        """the handle's HIP stream (as an integer handle)"""
        v = ctypes.c_void_p()
        L.check(self._L.race_get_stream(self.h, ctypes.byref(v)), "race_get_stream")
        return v.value

    def timerBegin(self):
        L.check(self._L.race_timer_begin(self.h), "timerBegin")

    def timerEnd(self):
        v = ctypes.c_float()
        L.check(self._L.race_timer_end(self.h, ctypes.byref(v)), "timerEnd")
        return float(v.value)

    def sell_c_sigmize(self, simdWidth, C, col, chunkStart, rl, clp):
        """sell_c_sigmize(simdWidth, C, col, chunkStart, rl, clp, ce) in place."""
        for a in (col, chunkStart, rl, clp):
            if not (isinstance(a, np.ndarray) and a.dtype == np.int32 and a.flags.c_contiguous):
                raise TypeError("SELL arrays must be contiguous int32")
        L.check(self._L.race_sell_c_sigmize(self.h, int(simdWidth), int(C), _ip(col), _ip(chunkStart), _ip(rl),
                                            _ip(clp)), "sell_c_sigmize")

    def simdify(self, simdWidth, C, nrows, col, chunkStart, rl, clp, val, diagFirst=False):
        """Interface::simdify in place: padding by the colouring, then the
        distance-2 rearrangement.  True on success."""
        prec = _sell_args(col, chunkStart, clp, val)
        rl = _i32(rl)
        st = self._L.race_simdify(self.h, int(simdWidth), int(C), int(nrows), _ip(col), _ip(chunkStart), _ip(rl),
                                  _ip(clp), val.ctypes.data, prec, int(bool(diagFirst)))
        if st == 6:
            return False
        L.check(st, "simdify")
        return True

    def simdifyD1(self, simdWidth, C, nrows, col, chunkStart, rl, clp, val, dist2Compatible=False):
        return simdifyD1(simdWidth, C, nrows, col, chunkStart, rl, clp, val, dist2Compatible)

    def preprocessStats(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        L.check(self._L.race_get_preprocess_stats(self.h, ctypes.byref(ms), ctypes.byref(n)), "stats")
        return ms.value, n.value


class Batch:
    """One GS / KACZ / SpMV / SymmSpMV function on many matrices, run as
    fused launches (race_batch_* in the C ABI).  Each matrix is added with
    its coloured Interface, its (permuted) CRS and its device vectors."""

    def __init__(self, kind, val_precision=FP64, vec_precision=FP64, nrhs=1):
        self._L = L.load()
        self.kind = int(kind)
        h = ctypes.c_void_p()
        L.check(self._L.race_batch_create(ctypes.byref(h), self.kind, int(val_precision), int(vec_precision),
                                          int(nrhs)), "race_batch_create")
        self.h = h
        self.val_precision, self.vec_precision, self.nrhs = int(val_precision), int(vec_precision), int(nrhs)
        self._keep = []

    def add(self, iface, rowPtr, col, val, x, b):
        rp, c = _i32(rowPtr), _i32(col)
        v = np.ascontiguousarray(val, dtype=np.float64)
        d = L.KernelDesc(kind=self.kind, nrow=iface.nrow, rowPtr=_ip(rp), col=_ip(c),
                         val=v.ctypes.data_as(L.f64p), val_precision=self.val_precision,
                         vec_precision=self.vec_precision, nrhs=self.nrhs,
                         x=x.ptr if isinstance(x, DeviceArray) else x,
                         b=b.ptr if isinstance(b, DeviceArray) else b, power=1)
        L.check(self._L.race_batch_add(self.h, iface.h, ctypes.byref(d)), "race_batch_add")
        self._keep.append((iface, x, b))

    def setOption(self, key, value):
        L.check(self._L.race_batch_set_option(self.h, key.encode(), int(value)), "race_batch_set_option")

    def getOption(self, key):
        v = ctypes.c_int()
        L.check(self._L.race_batch_get_option(self.h, key.encode(), ctypes.byref(v)), "race_batch_get_option")
        return v.value

    def launchCount(self, rev=False):
        """kernel launches of one execution of the built batch"""
        v = ctypes.c_int()
        L.check(self._L.race_batch_get_launch_count(self.h, int(rev), ctypes.byref(v)), "race_batch_get_launch_count")
        return v.value

    def flowAborts(self):
        """one-launch sweeps whose dependency waits gave up (0 when consistent)"""
        v = ctypes.c_int()
        L.check(self._L.race_batch_get_flow_aborts(self.h, ctypes.byref(v)), "race_batch_get_flow_aborts")
        return v.value

    def build(self):
        L.check(self._L.race_batch_build(self.h), "race_batch_build")

    def execute(self, rev=False):
        L.check(self._L.race_batch_execute(self.h, int(bool(rev))), "race_batch_execute")

    def executeAsync(self, rev=False, stream=None):
        L.check(self._L.race_batch_execute_async(self.h, int(bool(rev)), stream), "race_batch_execute_async")

    def time(self, rev=False, reps=5):
        ms = ctypes.c_float()
        L.check(self._L.race_batch_time(self.h, int(bool(rev)), int(reps), ctypes.byref(ms)), "race_batch_time")
        return float(ms.value)

    def useCandidate(self, i):
        L.check(self._L.race_batch_use_candidate(self.h, int(i)), "race_batch_use_candidate")

    def autotune(self, rev=False, reps=3, progress=None):
        """Time the 12 candidate layouts, keep the fastest; returns (index,
        times, -1 for a layout these matrices cannot use).  With progress, the
        candidates are timed one call at a time (race_batch_autotune's loop)
        and progress(i, ms) is called after each."""
        if progress is None:
            best = ctypes.c_int()
            t = (ctypes.c_float * 12)()
            L.check(self._L.race_batch_autotune(self.h, int(bool(rev)), int(reps), ctypes.byref(best), t, 12),
                    "race_batch_autotune")
            return best.value, [float(v) for v in t]
        times = []
        for i in range(12):
            self.useCandidate(i)
            st = self._L.race_batch_build(self.h)
            ms = ctypes.c_float(-1.0)
            if st == L.SUCCESS:
                st = self._L.race_batch_time(self.h, int(bool(rev)), int(reps), ctypes.byref(ms))
            if st == 9:  # RACE_ERR_INTERNAL
                L.check(st, "race_batch_autotune")
            times.append(float(ms.value) if st == L.SUCCESS else -1.0)
            progress(i, times[-1])
        ok = [i for i in range(12) if times[i] >= 0]
        if not ok:
            L.check(6, "race_batch_autotune")
        best = min(ok, key=lambda i: times[i])
        self.useCandidate(best)
        self.build()
        return best, times

    @staticmethod
    def candidate(kind, i):
        a, b_, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.check(L.load().race_batch_candidate(int(kind), int(i), ctypes.byref(a), ctypes.byref(b_),
                                              ctypes.byref(c)), "race_batch_candidate")
        return a.value, b_.value, c.value

    def close(self):
        if getattr(self, "h", None):
            self._L.race_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rcm(rowPtr, col):
    """Reverse Cuthill-McKee order of a symmetric pattern: (perm, invPerm),
    perm[new] = old, for Interface(..., initPerm=perm, initInvPerm=invPerm)."""
    rp, c = _i32(rowPtr), _i32(col)
    n = len(rp) - 1
    perm, inv = np.empty(n, np.int32), np.empty(n, np.int32)
    L.check(L.load().race_rcm(n, _ip(rp), _ip(c), _ip(perm), _ip(inv)), "rcm")
    return perm, inv


def schedule_info(data):
    """Check a saved schedule without a handle: dict(nrow, dist, nthreads)."""
    data = bytes(data)
    n, d, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.check(L.load().race_schedule_info(data, len(data), ctypes.byref(n), ctypes.byref(d), ctypes.byref(t)),
            "schedule_info")
    return {"nrow": n.value, "dist": d.value, "nthreads": t.value}


def _sell_args(col, chunkStart, clp, val):
    """Checked views for the in-place SELL-C-sigma helpers."""
    for a, dt in ((col, np.int32), (chunkStart, np.int32), (clp, np.int32)):
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
            raise TypeError("col, chunkStart and chunk lengths must be contiguous int32 arrays")
    if not (isinstance(val, np.ndarray) and val.dtype in (np.float64, np.float32) and val.flags.c_contiguous):
        raise TypeError("val must be a contiguous float64 or float32 array")
    if len(val) != len(col):
        raise ValueError("col and val differ in length")
    return FP64 if val.dtype == np.float64 else FP32


def simdify_resolve(simdWidth, C, nrows, col, chunkStart, clp, val, diagFirst=False):
    """Distance-2 SIMD rearrangement of a SELL-C-sigma matrix in place (host
    only; the padding step of Interface.simdify not included).  True on
    success, False where a SIMD group cannot be resolved."""
    prec = _sell_args(col, chunkStart, clp, val)
    st = L.load().race_simdify_resolve(int(simdWidth), int(C), int(nrows), _ip(col), _ip(chunkStart), _ip(clp),
                                       val.ctypes.data, prec, int(bool(diagFirst)))
    if st == 6:  # RACE_ERR_INCOMPATIBILITY
        return False
    L.check(st, "simdify")
    return True


def simdifyD1(simdWidth, C, nrows, col, chunkStart, rl, clp, val, dist2Compatible=False):
    """Interface::simdifyD1 in place (host only)."""
    prec = _sell_args(col, chunkStart, clp, val)
    L.check(L.load().race_simdify_d1(int(simdWidth), int(C), int(nrows), _ip(col), _ip(chunkStart), _ip(clp),
                                     val.ctypes.data, prec, int(bool(dist2Compatible))), "simdifyD1")
    return True


def lb_split(levelRow, levelNnz, nthreads, eff, dist, d2=TWO_BLOCK, lb_target=LB_NNZ):
    """LB::splitZones on a level histogram through the C ABI (host only)."""
    lr, ln = _i32(levelRow), _i32(levelNnz)
    cap = 4 * len(lr) + 8
    nb, nt = ctypes.c_int(), ctypes.c_int()
    zr = np.zeros(cap, np.int32)
    lp = np.zeros(cap, np.int32)
    sc = np.zeros(cap, np.int32)
    L.check(L.load().race_lb_split(_ip(lr), _ip(ln), len(lr), int(nthreads), float(eff), int(dist), int(d2),
                                   int(lb_target), cap, ctypes.byref(nb), _ip(zr), _ip(lp), ctypes.byref(nt),
                                   _ip(sc)), "race_lb_split")
    return dict(zone_rows=zr[:nb.value].tolist(), levelPtr=lp[:nb.value + 1].tolist(),
                scale=sc[:nt.value].tolist(), threads=nt.value)


def split_threads(weights, total, parent_threads):
    """LevelRecursion::calculateIdealNthreads' thread split through the C
    ABI (host only)."""
    w = _i32(weights)
    out = np.zeros(max(len(w), 1), np.int32)
    L.check(L.load().race_split_threads(_ip(w), len(w), int(total), int(parent_threads), _ip(out)),
            "race_split_threads")
    return out[:len(w)].tolist()


def mpk_hopeless_regions(violated, k, total_level):
    """mtxPower::identifyHopelessRegions through the C ABI (host only)."""
    v = _i32(violated)
    cap = 2 * len(v) + 4
    out = np.zeros(cap, np.int32)
    n = ctypes.c_int()
    L.check(L.load().race_mpk_hopeless_regions(_ip(v), len(v), int(k), int(total_level), _ip(out), cap,
                                               ctypes.byref(n)), "race_mpk_hopeless_regions")
    return out[:n.value].tolist()


def mpk_consolidate(level_ptr, elem, k, cache_elem, regions):
    """mtxPower::consolidatePartition through the C ABI (host only):
    (group row pointers, group first levels, group of each region)."""
    lp = _i32(level_ptr)
    el = np.ascontiguousarray(elem, np.float64)
    rg = _i32(regions if len(regions) else [0])
    cap = len(lp) + 4
    g, gl, h = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    ng, nh = ctypes.c_int(), ctypes.c_int()
    L.check(L.load().race_mpk_consolidate(_ip(lp), el.ctypes.data_as(L.f64p), len(el), int(k), float(cache_elem),
                                          _ip(rg), len(regions), _ip(g), _ip(gl), _ip(h), cap, ctypes.byref(ng),
                                          ctypes.byref(nh)), "race_mpk_consolidate")
    return g[:ng.value + 1].tolist(), gl[:ng.value + 1].tolist(), h[:nh.value].tolist()
