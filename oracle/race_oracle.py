"""CPU oracle for the RACE hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product (race_amd/) never does.  It is the checker for the
MI355X path, not something that is measured or shipped.

This file restates in plain Python the level-granular parts of
RRZE-HPC/RACE (work over BFS levels, not rows), and drives the row-granular
C restatement in race_oracle.c through ctypes:

  split_zones          LB::splitZones / calcZonePtr / Stat::calculate
                       src/lb.cpp:188-667, 694-707, 773-795
  find_neighbour       LB::findNeighbour  src/lb.cpp:136-172
  move_one_step        LB::moveOneStep    src/lb.cpp:174-186
  ZoneTree             ZoneTree::spawnChild/findKeyChild/updateTreeEffRow/
                       updateTreeNThreads src/zone_tree.cpp:91-206,243-305
  level_recursion      LevelRecursion ctor/levelBalancing/recursivePartition/
                       calculateIdealNthreads src/level_recursion.cpp:28-99,
                       149-376
  leaf_calls           recursiveCall (OpenMP variant, serialised)
                       src/functionManager.cpp:199-265, Run :1453-1503
  mpk_partition        mtxPowerRecursive::findPartition (stage 0) +
                       mtxPower::findPartition/consolidatePartition
                       src/matrixPowerRecursive.cpp:340-445,
                       src/matrixPower.cpp:88-140,505-604,685-940
  sell_c_sigmize       sell_c_sigmize_Kernel run over the leaf calls
                       src/sell_c_sigmize.cpp:26-62
  simdify_resolve      simdifyTemplate after its sell_c_sigmize call
                       include/simdify.h:70-357
  simdify_d1           simdifyD1Template include/simdify.h:389-447

Pinning (the reference cannot be built in this image -- its build needs the
GAP and hwloc packages that are fetched from the network plus a generated
config.h):
  * traversal: closed forms for stencils, BFS invariants (race_oracle.c).
  * split_zones / ZoneTree / level_recursion: invariants only -- zones cover
    the level range contiguously, every zone has >= minGap levels, same-colour
    zones are distance-k independent (brute force), thread counts add up,
    effRow equals the critical path recomputed from the leaves; one known
    answer fixed by symmetry (equal nnz on every level, a level count that
    is a multiple of the blocks: the split stays uniform).  Beyond that case
    the exact zone boundaries chosen by the variance heuristic are PARITY
    UNPINNED against the reference: nothing in the reference tree fixes
    them.
  * leaf_calls: with one thread the order is the serial order (textbook
    Gauss-Seidel/Kaczmarz); with T threads every pair of concurrently
    scheduled leaves is checked independent.
  * mpk_partition (stage 0): every group respects the cache bound, groups
    tile the level range, and MPK through the groups equals repeated SpMV.
  * run_calls kernels (race_oracle.c, examples/src/helpers/kernels.cpp):
    SpMV and SpMTV against dense products A x and A^T x; SymmSpMV against
    the full SpMV; GS/KACZ with one thread against textbook serial sweeps;
    mpk_serial against repeated SpMV; mpk_highest_serial bit-equal to the
    last two powers of mpk_serial.
  * sell_c_sigmize / simdify_resolve / simdify_d1: the 4x4 all-equal block
    of the reference's testSimdify (src/simdify.cpp:41-77) traced by hand
    to its XOR Latin square; every row keeps its (column, value) multiset;
    after simdify no two rows of a SIMD group share a column at a position
    whenever it reports success; after simdify_d1 the entry on the row of
    lane m sits at position cl-1-m (cl-1-(i-1-m) when dist2Compatible), the
    picture of include/simdify.h:367-381; padding columns equal leaf start
    + offset mod width.
"""
import ctypes
import os
import subprocess

import numpy as np

ONE, TWO, POWER = 1, 2, 3
TWO_BLOCK, THREE_BLOCK = 0, 1
LB_ROW, LB_NNZ = 0, 1
K_SYMMSPMV, K_GS, K_KACZ, K_SPMV, K_SPMTV = 0, 1, 2, 3, 5

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "librace_oracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile race_oracle.c into oracle/librace_oracle.so."""
    src = os.path.join(_HERE, "race_oracle.c")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-o", _LIB, src, "-lm"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.or_ws_create.restype = ctypes.c_void_p
        L.or_ws_create.argtypes = [ctypes.c_int]
        L.or_ws_destroy.argtypes = [ctypes.c_void_p]
        L.or_traverse.restype = ctypes.c_int
        L.or_traverse.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p]
        L.or_traverse_boundary.restype = ctypes.c_int
        L.or_traverse_boundary.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p,
                                           _i32p, _i32p, _i32p, ctypes.c_int]
        L.or_symmetrize.restype = ctypes.c_long
        L.or_symmetrize.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_long]
        L.or_run_calls.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                   ctypes.c_int, _i32p, ctypes.c_int]
        L.or_run_phases.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                    ctypes.c_int, _i32p, _i32p, ctypes.c_int, ctypes.c_int]
        L.or_max_threads.restype = ctypes.c_int
        L.or_mpk_serial.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, ctypes.c_int]
        L.or_mpk_highest_serial.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, ctypes.c_int]
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(_i32p)


def _dp(a):
    return a.ctypes.data_as(_f64p)


# --------------------------------------------------------------------------
# helpers mirroring include/utility.h and include/macros.h
# --------------------------------------------------------------------------
def round_double(v):
    """roundDouble include/utility.h:194-198 (truncate, then +1 if frac>=0.5)."""
    vi = int(v)
    return vi if (v - vi) < 0.5 else vi + 1


def wrap32(v):
    """C int arithmetic wraps at 32 bits on the reference's targets."""
    return ((int(v) + 2 ** 31) % 2 ** 32) - 2 ** 31


def block_per_thread(dist, d2=TWO_BLOCK):
    # getBlockPerThread include/macros.h:29-30
    if dist == POWER:
        return 1
    if dist == ONE:
        return 2
    return 2 if d2 == TWO_BLOCK else 3


def min_gap(dist, d2=TWO_BLOCK):
    # getMinGap include/macros.h:33-34
    if dist == POWER:
        return 1
    if dist == ONE:
        return 1
    return 2 if d2 == TWO_BLOCK else 1


def stable_argsort_desc(vals):
    """std::stable_sort(perm, [](a,b){return arr[a] > arr[b];})"""
    return sorted(range(len(vals)), key=lambda i: -vals[i])


# --------------------------------------------------------------------------
# Graph + traversal (row-granular work in C)
# --------------------------------------------------------------------------
class Graph:
    """RACE::Graph (SoA, RACE_PERMUTE_ON_FLY): borrowed CRS in original
    numbering plus totalPerm/totalInvPerm (src/graph_SoA.cpp:48-216,437-477)."""

    def __init__(self, rowPtr, col, symm_hint=False, init_perm=None, init_inv_perm=None, max_dense=None):
        rowPtr = np.ascontiguousarray(rowPtr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        n = len(rowPtr) - 1
        self.n = n
        self.nfull = n
        self.nnz_input = int(rowPtr[-1])  # Graph::NNZ = rowPtr[NROW] of the input
        if not symm_hint:
            cap = int(rowPtr[-1]) * 2 + 16
            rp2 = np.zeros(n + 1, np.int32)
            c2 = np.zeros(cap, np.int32)
            nnz = lib().or_symmetrize(n, _ip(rowPtr), _ip(col), _ip(rp2), _ip(c2), cap)
            assert nnz >= 0
            rowPtr, col = rp2, c2[:nnz].copy()
        self.rowPtr, self.col = rowPtr, col
        if init_perm is not None:
            self.perm = np.ascontiguousarray(init_perm, dtype=np.int32).copy()
            self.inv = np.ascontiguousarray(init_inv_perm, dtype=np.int32).copy()
        else:
            self.perm = np.arange(n, dtype=np.int32)
            self.inv = np.arange(n, dtype=np.int32)
        self.graph_rows = None
        self.dense_rows = []
        if max_dense is not None:
            self._split_dense(float(max_dense))
        self.ws = lib().or_ws_create(self.n)

    def _split_dense(self, max_dense):
        """Graph::getStatistics + permuteAndRemoveSerialPart
        (src/graph_SoA.cpp:321-435, the build without RACE_PERMUTE_ON_FLY):
        rows longer than max_dense x (NNZ / NROW, integer division) move, in
        the current order, to a serial tail; the graph keeps the other rows
        in order and drops every entry pointing to a tail row."""
        n = self.n
        mean = self.nnz_input // n
        lens = np.diff(self.rowPtr)
        cur = self.perm  # current row r = original row perm[r]
        dense = [int(o) for o in cur if lens[o] > max_dense * mean]
        if not dense:
            return
        keep = [int(o) for o in cur if not lens[o] > max_dense * mean]
        new_id = np.full(n, -1, np.int64)
        new_id[keep] = np.arange(len(keep))
        rp, cl = [0], []
        for o in keep:
            for c in self.col[self.rowPtr[o]:self.rowPtr[o + 1]]:
                if new_id[c] >= 0:
                    cl.append(int(new_id[c]))
            rp.append(len(cl))
        self.rowPtr = np.array(rp, np.int32)
        self.col = np.array(cl, np.int32)
        self.graph_rows = np.array(keep, np.int32)
        self.dense_rows = dense
        self.n = len(keep)
        self.perm = np.arange(self.n, dtype=np.int32)
        self.inv = np.arange(self.n, dtype=np.int32)

    def full_perm(self):
        """perm[new] = old over all rows: the graph rows in their RACE
        order, then the dense tail"""
        if self.graph_rows is None:
            return self.perm.copy()
        return np.concatenate([self.graph_rows[self.perm], np.array(self.dense_rows, np.int32)])

    def __del__(self):
        try:
            lib().or_ws_destroy(self.ws)
        except Exception:
            pass

    def row_len(self, orig_row):
        return int(self.rowPtr[orig_row + 1] - self.rowPtr[orig_row])

    def traverse(self, dist, lo, hi, want_dist=False):
        """Traverse(graph,dist,lo,hi).calculateDistance(); returns
        (levelRow, levelNnz[, dist]) and composes the permutation."""
        cap = 3 * (hi - lo) + 2  # distance-2: up to three levels per row
        lr = np.zeros(cap, np.int32)
        ln = np.zeros(cap, np.int32)
        dd = np.zeros(max(hi - lo, 1), np.int32)
        tl = lib().or_traverse(self.ws, self.n, _ip(self.rowPtr), _ip(self.col), _ip(self.perm),
                               _ip(self.inv), dist, lo, hi, _ip(lr), _ip(ln),
                               _ip(dd) if want_dist else None)
        if tl < 0:
            raise RuntimeError("traversal lost rows")
        out = (lr[:tl].copy(), ln[:tl].copy())
        return out + ((dd[:hi - lo].copy(),) if want_dist else ())

    def traverse_boundary(self, lo, hi, bounds):
        """POWER traversal of [lo,hi) with boundary regions bounds = [(blo,
        bhi), ...] (the recursion stage of mtxPowerRecursive); returns
        (levelRow, levelNnz, [(levelRow_b, levelNnz_b) per region]) and
        composes the permutation of the range and of every region."""
        nb = len(bounds)
        cap = (hi - lo) + sum(b - a for a, b in bounds) + 2
        lr = np.zeros(cap, np.int32)
        ln = np.zeros(cap, np.int32)
        blr = np.zeros(max(nb, 1) * cap, np.int32)
        bln = np.zeros(max(nb, 1) * cap, np.int32)
        blo = np.array([a for a, _ in bounds] or [0], np.int32)
        bhi = np.array([b for _, b in bounds] or [0], np.int32)
        tl = lib().or_traverse_boundary(self.ws, self.n, _ip(self.rowPtr), _ip(self.col), _ip(self.perm),
                                        _ip(self.inv), lo, hi, nb, _ip(blo), _ip(bhi), _ip(lr), _ip(ln),
                                        _ip(blr), _ip(bln), cap)
        if tl < 0:
            raise RuntimeError("traversal lost rows")
        regs = [(blr[b * cap:b * cap + tl].copy(), bln[b * cap:b * cap + tl].copy()) for b in range(nb)]
        return lr[:tl].copy(), ln[:tl].copy(), regs


# --------------------------------------------------------------------------
# LB
# --------------------------------------------------------------------------
class _Stat:
    # Stat::calculate src/lb.cpp:773-795
    def __init__(self, arr, num_partitions, scale):
        L = len(arr)
        self.mean = [0.0] * num_partitions
        self.var = [0.0] * num_partitions
        self.acquire = [0.0] * L
        self.give = [0.0] * L
        self.weight = [0.0] * L
        for p in range(num_partitions):
            m = 0.0
            ctr = 0
            for i in range(p, L, num_partitions):
                m = m + arr[i] / float(scale[ctr])
                ctr += 1
            m = (num_partitions * m) / L
            self.mean[p] = m
            v = 0.0
            ctr = 0
            for i in range(p, L, num_partitions):
                diff = (arr[i] / float(scale[ctr])) - m
                v = v + diff * diff
                self.acquire[i] = diff
                self.give[i] = -diff
                self.weight[i] = abs(diff)
                ctr += 1
            self.var[p] = v


def split_zones(levelRow, levelNnz, nthreads, eff, dist, d2=TWO_BLOCK, lb_target=LB_NNZ):
    """LB::balance -> splitZones. Returns dict with levelPtr (blocks), scale,
    zone row sizes, numBlocks and subLevelPtr."""
    levelRow = [int(v) for v in levelRow]
    levelNnz = [int(v) for v in levelNnz]
    total_level = len(levelRow)
    target = levelNnz if lb_target == LB_NNZ else levelRow
    target_n = sum(target)  # LevelData::nnz / nrow are plain int sums
    eff_ratio = [wrap32(target[i] * nthreads) / float(target_n) for i in range(total_level)]

    mg = min_gap(dist, d2)
    bpt = block_per_thread(dist, d2)
    min_gap_per_thread = mg * bpt
    acc = 0.0
    mg_ctr = 0
    curr_rem = 0.0
    scale = []
    init_lp = [0]
    i = 0
    while i < total_level:
        acc += eff_ratio[i]
        mg_ctr += 1
        if mg_ctr >= min_gap_per_thread:
            if round_double(acc) != 0:
                if ((1 - abs(acc - round_double(acc))) * 100) > eff:
                    local_rem = acc - round_double(acc)
                    new_local_rem = local_rem
                    min_local_rem = local_rem
                    rounded = round_double(acc)
                    temp_acc = acc
                    min_temp_acc = acc
                    j = i + 1
                    best = i
                    while j < total_level and abs(new_local_rem) <= abs(local_rem):
                        temp_acc += eff_ratio[j]
                        new_local_rem = temp_acc - rounded
                        if abs(min_local_rem) > abs(new_local_rem):
                            min_local_rem = new_local_rem
                            min_temp_acc = temp_acc
                            best = j
                        elif (abs(min_local_rem) - abs(new_local_rem)) < 1e-6:
                            if (min_local_rem + curr_rem) > (new_local_rem + curr_rem):
                                min_local_rem = new_local_rem
                                min_temp_acc = temp_acc
                                best = j
                        j += 1
                    mg_ctr += best - i
                    i = best
                    acc = min_temp_acc
                    curr_rem = curr_rem + (acc - round_double(acc))
                    scale.append(round_double(acc))
                    init_lp.append(i + 1)
                    acc = 0.0
                    mg_ctr = 0
        if i >= total_level - 5:
            last = acc
            for j in range(i + 1, total_level):
                last += eff_ratio[j]
            if round_double(last) != 0:
                scale.append(round_double(last))
                init_lp.append(total_level)
            else:
                init_lp[-1] = total_level
            break
        i += 1

    thread_ctr = 0
    for k in range(len(scale)):
        if scale[k] == 0:
            scale[k] = 1
        thread_ctr += scale[k]
    curr_lvl_threads = 0
    if scale:
        if thread_ctr != nthreads:
            diff = nthreads - thread_ctr
            if abs(diff) == 1:
                if diff == 1:
                    scale[-1] += 1
                else:
                    if scale[-1] == 1:
                        scale.pop()
                        init_lp.pop()
                        init_lp[-1] = total_level
                    else:
                        scale[-1] -= 1
        curr_lvl_threads = len(scale)

    total_blocks = bpt * curr_lvl_threads
    lp = [0] * (total_blocks + 1)
    for t in range(curr_lvl_threads):
        for j in range(1, bpt + 1):
            lp[t * bpt + j] = init_lp[t] + round_double(j * (init_lp[t + 1] - init_lp[t]) / float(bpt))

    def chunk_sum(ptr):
        return [sum(target[ptr[b]:ptr[b + 1]]) for b in range(total_blocks)]

    def find_neighbour(st, acquire):
        order = stable_argsort_desc(st.acquire if acquire else st.give)
        if acquire:
            for idx in order:
                if (lp[idx + 1] - lp[idx]) > mg:
                    return idx
            return -1
        return order[0]

    def move_one_step(to_idx, from_idx):
        if to_idx < from_idx:
            for k in range(to_idx + 1, from_idx + 1):
                lp[k] += 1
        else:
            for k in range(to_idx, from_idx, -1):
                lp[k] -= 1

    if scale:
        exit_flag = False
        while not exit_flag:
            cs = chunk_sum(lp)
            st = _Stat(cs, bpt, scale)
            var = sum(st.var)
            new_var = var
            rank = stable_argsort_desc(st.weight)
            curr_rank = 0
            old_lp = list(lp)
            while new_var >= var:
                r_idx = rank[curr_rank]
                lp[:] = old_lp
                my_mean = st.mean[r_idx % bpt]
                fail = False
                move_possible = True
                first = True
                while first or (move_possible and abs(var - new_var) < 1e-5):
                    first = False
                    if (cs[r_idx] / float(scale[r_idx // bpt])) < my_mean:
                        a_idx = find_neighbour(st, True)
                        if a_idx == -1 or a_idx == r_idx:
                            fail = True
                            move_possible = False
                        move_one_step(r_idx, a_idx)
                    else:
                        if (lp[r_idx + 1] - lp[r_idx]) > mg:
                            g_idx = find_neighbour(st, False)
                            if g_idx == -1 or r_idx == g_idx:
                                move_possible = False
                                fail = True
                            move_one_step(g_idx, r_idx)
                        else:
                            move_possible = False
                    if not fail:
                        new_var = sum(_Stat(chunk_sum(lp), bpt, scale).var)
                if curr_rank == total_blocks - 1 and new_var >= var:
                    exit_flag = True
                    lp[:] = old_lp
                    break
                curr_rank += 1

    # sub-blocking (cache blocking is disabled: cacheSize=INT_MAX, :580-640)
    blocked = (2 ** 31 - 1) // (8 * 2 + 27 * 12)
    blocked = max(blocked, max(levelRow) if levelRow else 0)
    zone_rows = [sum(levelRow[lp[b]:lp[b + 1]]) for b in range(total_blocks)]
    num_blocks = [0] * total_blocks
    sub_lp = []
    for b in range(total_blocks):
        num_blocks[b] += 1
        sub_lp.append(lp[b])
        t = int(b / float(bpt))
        if zone_rows[b] > blocked and scale[t] > 1 and (lp[b + 1] - lp[b]) > 1:
            cur = levelRow[lp[b]]
            for j in range(lp[b] + 1, lp[b + 1]):
                cur += levelRow[j]
                if cur > blocked:
                    num_blocks[b] += 1
                    sub_lp.append(j)
                    cur = levelRow[j]
    sub_lp.append(lp[total_blocks])
    sub_rows = [sum(levelRow[sub_lp[s]:sub_lp[s + 1]]) for s in range(len(sub_lp) - 1)]
    return dict(levelPtr=lp, scale=scale, zone_rows=zone_rows, num_blocks=num_blocks,
                sub_rows=sub_rows, bpt=bpt, threads=len(scale))


# --------------------------------------------------------------------------
# Zone tree
# --------------------------------------------------------------------------
class Leaf:
    def __init__(self, lo, hi, parent):
        # ZoneLeaf(rangeLo, rangeHi, parent) src/zone_tree.cpp:30-34
        self.valueZ = [lo, hi]
        self.children = []
        self.nthreadsZ = 1
        self.idealNthreadsZ = 1
        self.totalSubBlocks = 1
        self.parentZ = parent
        self.effRowZ = hi - lo
        self.reachedLimit = False
        self.stage = 0


class ZoneTree:
    def __init__(self, dist, d2=TWO_BLOCK, lb_target=LB_NNZ):
        self.dist, self.d2, self.lb_target = dist, d2, lb_target
        self.bpt = block_per_thread(dist, d2)
        self.tree = []

    def find_key_child_effrow(self, p):
        ch = self.tree[p].children
        bpt = self.bpt
        max_row = [0] * bpt
        for sb in range(len(ch) // 2):
            tmax = [0] * bpt
            for b in range(bpt):
                for i in range(ch[2 * sb] + b, ch[2 * sb + 1], bpt):
                    nr = self.tree[i].effRowZ
                    if nr > tmax[b]:
                        tmax[b] = nr
            for b in range(bpt):
                max_row[b] += tmax[b]
        return sum(max_row)

    def update_eff_row(self, p):
        self.tree[p].effRowZ = self.find_key_child_effrow(p)
        nxt = self.tree[p].parentZ
        if nxt != -1:
            self.update_eff_row(nxt)

    def update_nthreads(self, p):
        ch = self.tree[p].children
        bpt = self.bpt
        nthr = 0
        for sb in range(len(ch) // 2):
            sub = 0
            nt = (ch[2 * sb + 1] - ch[2 * sb]) // bpt
            for i in range(nt):
                m = 0
                for b in range(bpt):
                    m = max(self.tree[ch[2 * sb] + bpt * i + b].nthreadsZ, m)
                sub += m
            nthr = max(nthr, sub)
        self.tree[p].nthreadsZ = nthr
        nxt = self.tree[p].parentZ
        if nxt != -1:
            self.update_nthreads(nxt)

    def spawn_child(self, p, psub, request, levelRow, levelNnz, eff):
        res = split_zones(levelRow, levelNnz, request, eff, self.dist, self.d2, self.lb_target)
        base_len = len(self.tree)
        base = self.tree[p].valueZ[psub]
        stage = self.tree[p].stage
        zp = [base]
        for r in res["zone_rows"]:
            zp.append(zp[-1] + r)
        szp = [base]
        for r in res["sub_rows"]:
            szp.append(szp[-1] + r)
        self.tree[p].children.append(base_len)
        ctr = 0
        for b in range(len(res["zone_rows"])):
            leaf = Leaf(zp[b], zp[b + 1], p)
            nb = res["num_blocks"][b]
            vz = []
            for s in range(nb):
                vz.append(szp[ctr])
                ctr += 1
            vz.append(szp[ctr])
            leaf.valueZ = vz
            leaf.totalSubBlocks = nb
            leaf.stage = stage + 1
            self.tree.append(leaf)
        self.tree[p].children.append(base_len + len(res["zone_rows"]))
        if psub == self.tree[p].totalSubBlocks - 1:
            self.update_eff_row(p)
            self.update_nthreads(p)
        return res["threads"] == request


def split_threads(weights, total, parent_threads):
    """The thread split of LevelRecursion::calculateIdealNthreads
    (src/level_recursion.cpp:165-245): child group i (the blocks of one
    thread of the parent's split) gets floor(weights[i] / total *
    parent_threads) threads, weights being row counts or nnz (ints, total =
    the parent range's); a group that got 0 ranks with remainder 1, the
    others with their fractional part, and the leftover threads go one each
    to the groups in descending rank order (stable sort, ties keep group
    order)."""
    rem = []
    out = []
    remaining = parent_threads
    for w in weights:
        my_d = (int(w) / float(total)) * parent_threads
        my = int(my_d)
        out.append(my)
        rem.append((my_d - my) if my != 0 else 1.0)
        remaining -= my
    if remaining > 0:
        order = stable_argsort_desc(rem)
        r = 0
        while remaining != 0:
            out[order[r]] += 1
            remaining -= 1
            r += 1
    return out


def level_recursion(graph, nthreads, dist, d2=TWO_BLOCK, lb_target=LB_NNZ, eff_list=None,
                    lvl_threads=None):
    """LevelRecursion(graph, T, dist).levelBalancing(); eff_list mirrors
    RACE_EFFICIENCY (appended after the default 40), lvl_threads RACE_THREADS."""
    eff_vec = [40.0] + [float(int(e)) for e in (eff_list or [])]  # atoi semantics
    lvl_threads = list(lvl_threads or [])
    for lvl in range(1, len(lvl_threads) + 1):
        while len(eff_vec) <= lvl:
            eff_vec.append(eff_vec[0])
        eff_vec[lvl] = 0.0

    def efficiency(lvl):
        return eff_vec[lvl] if lvl < len(eff_vec) else eff_vec[0]

    def lvl_thr(lvl):
        return lvl_threads[lvl - 1] if 0 <= lvl - 1 < len(lvl_threads) else -1

    zt = ZoneTree(dist, d2, lb_target)
    bpt = zt.bpt
    root = Leaf(0, graph.n, -1)
    root.idealNthreadsZ = nthreads
    root.totalSubBlocks = 1
    zt.tree.append(root)

    def row_nnz_sum(lo, hi):
        pr = graph.perm[lo:hi]
        return int(np.sum(graph.rowPtr[pr + 1] - graph.rowPtr[pr]))

    def ideal_nthreads(p, psub, curr_lvl):
        T = zt.tree
        ch = T[p].children
        nt = (ch[2 * psub + 1] - ch[2 * psub]) // bpt
        if lvl_thr(curr_lvl) == -1:
            parent_threads = T[p].idealNthreadsZ
            s, e = T[p].valueZ[psub], T[p].valueZ[psub + 1]
            parent_row = e - s
            parent_nnz = row_nnz_sum(s, e) if lb_target == LB_NNZ else 0
            weights = []
            for i in range(nt):
                first = T[ch[2 * psub] + i * bpt].valueZ
                last = T[ch[2 * psub] + bpt * (i + 1) - 1].valueZ
                ss, ee = first[0], last[-1]
                weights.append(row_nnz_sum(ss, ee) if lb_target == LB_NNZ else ee - ss)
            total = parent_nnz if lb_target == LB_NNZ else parent_row
            for i, my in enumerate(split_threads(weights, total, parent_threads)):
                for b in range(bpt):
                    T[ch[2 * psub] + bpt * i + b].idealNthreadsZ = my
        else:
            my = lvl_thr(curr_lvl)
            for i in range(nt):
                for b in range(bpt):
                    zt.tree[ch[2 * psub] + bpt * i + b].idealNthreadsZ = my

    def recursive_partition(p, psub, curr_level):
        ideal_nthreads(p, psub, curr_level)
        ch = list(zt.tree[p].children)
        for ci in range(ch[2 * psub], ch[2 * psub + 1]):
            nt = zt.tree[ci].idealNthreadsZ
            rng = list(zt.tree[ci].valueZ)
            if nt > 1:
                for j in range(len(rng) - 1):
                    if rng[j + 1] > rng[j]:
                        lr, ln = graph.traverse(dist, rng[j], rng[j + 1])
                        ok = zt.spawn_child(ci, j, nt, lr, ln, efficiency(curr_level))
                        if not ok:
                            cch = zt.tree[ci].children
                            sub = (cch[2 * j + 1] - cch[2 * j]) // bpt
                            if sub == 1:
                                zt.tree[ci].reachedLimit = True
                            if sub > 1:
                                recursive_partition(ci, j, curr_level + 1)
                    else:
                        zt.tree[ci].reachedLimit = True

    lr, ln = graph.traverse(dist, 0, graph.n)
    lvl_one = nthreads if lvl_thr(1) == -1 else lvl_thr(1)
    zt.spawn_child(0, 0, lvl_one, lr, ln, efficiency(1))
    if zt.tree[0].nthreadsZ != nthreads:
        recursive_partition(0, 0, 2)
    zt.available_threads = zt.tree[0].nthreadsZ
    zt.max_stage = max(l.stage for l in zt.tree)
    return zt


def leaf_calls(zt, rev=False, serial_part=None):
    """Serialised order of recursiveCall's leaf invocations (start,end).
    Concurrent leaves (same parent, same colour block) are independent, so
    any serialisation gives the reference's results.  serial_part (lo, hi):
    the dense-row tail, called last in both directions
    (FuncManager::Run src/functionManager.cpp:1493-1500; reversed here as
    (hi-1, lo-1), the reverse-leaf convention of recursiveCall -- the
    reference passes (hi, lo), one row off)."""
    out = []
    bpt = zt.bpt

    def rec(p):
        leaf = zt.tree[p]
        ch = leaf.children
        for sb in range(leaf.totalSubBlocks):
            n = (ch[2 * sb + 1] - ch[2 * sb]) // bpt if ch else 0
            if n <= 1:
                lo, hi = leaf.valueZ[0], leaf.valueZ[1]
                out.append((hi - 1, lo - 1) if rev else (lo, hi))
            else:
                blocks = range(bpt - 1, -1, -1) if rev else range(bpt)
                for b in blocks:
                    for tid in range(n):
                        rec(ch[2 * sb] + bpt * tid + b)
    rec(0)
    if serial_part is not None and serial_part[1] > serial_part[0]:
        lo, hi = serial_part
        out.append((hi - 1, lo - 1) if rev else (lo, hi))
    return out


def phases(zt, rev=False):
    """Leaves grouped into barrier-separated phases: the children of a node
    that run in one colour block are merged phase by phase."""
    bpt = zt.bpt

    def rec(p):
        leaf = zt.tree[p]
        ch = leaf.children
        ph = []
        for sb in range(leaf.totalSubBlocks):
            n = (ch[2 * sb + 1] - ch[2 * sb]) // bpt if ch else 0
            if n <= 1:
                lo, hi = leaf.valueZ[0], leaf.valueZ[1]
                ph.append([(hi - 1, lo - 1) if rev else (lo, hi)])
            else:
                blocks = range(bpt - 1, -1, -1) if rev else range(bpt)
                for b in blocks:
                    merged = []
                    for tid in range(n):
                        sub = rec(ch[2 * sb] + bpt * tid + b)
                        for k, s in enumerate(sub):
                            if k == len(merged):
                                merged.append([])
                            merged[k].extend(s)
                    ph.extend(merged)
        return ph
    return rec(0)


def run_calls(kind, rowPtr, col, val, x, b, calls, nrhs=1):
    """Execute kernel `kind` over leaf calls in order (x, b modified in place)."""
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    c = np.ascontiguousarray(np.asarray(calls, np.int32).reshape(-1))
    n = len(rowPtr) - 1
    lib().or_run_calls(kind, n, _ip(rowPtr), _ip(col), _dp(val), _dp(x), _dp(b),
                       len(c) // 2, _ip(c), nrhs)


def run_phases(kind, rowPtr, col, val, x, b, phase_ptr, calls, nrhs=1, nthreads=None):
    """Execute kernel `kind` phase by phase, the leaf calls of a phase on
    OpenMP threads (x, b modified in place); calls is a flat (start, end)
    array, phase p owns calls phase_ptr[p] .. phase_ptr[p+1]."""
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    c = np.ascontiguousarray(np.asarray(calls, np.int32).reshape(-1))
    pp = np.ascontiguousarray(phase_ptr, np.int32)
    n = len(rowPtr) - 1
    nt = int(nthreads or lib().or_max_threads())
    lib().or_run_phases(kind, n, _ip(rowPtr), _ip(col), _dp(val), _dp(x), _dp(b), len(pp) - 1, _ip(pp),
                        _ip(c), nrhs, nt)
    return nt


def phase_calls(zt, rev=False):
    """phases(zt) flattened for run_phases: (phase_ptr, calls)."""
    ph = phases(zt, rev)
    pp = [0]
    flat = []
    for p in ph:
        for a, b in p:
            flat += [a, b]
        pp.append(len(flat) // 2)
    return np.array(pp, np.int32), np.array(flat, np.int32)


def mpk_serial(rowPtr, col, val, x0, k):
    n = len(rowPtr) - 1
    x = np.zeros((k + 1) * n, np.float64)
    x[:n] = x0
    lib().or_mpk_serial(n, _ip(np.ascontiguousarray(rowPtr, np.int32)),
                        _ip(np.ascontiguousarray(col, np.int32)),
                        _dp(np.ascontiguousarray(val, np.float64)), _dp(x), k)
    return x.reshape(k + 1, n)


def mpk_highest_serial(rowPtr, col, val, x0, k):
    """Two-slot matrix powers (MAT_SPMV_KERNEL_only_highest); returns the
    (2, n) slots: power p in slot p % 2."""
    n = len(rowPtr) - 1
    x = np.zeros(2 * n, np.float64)
    x[:n] = x0
    lib().or_mpk_highest_serial(n, _ip(np.ascontiguousarray(rowPtr, np.int32)),
                                _ip(np.ascontiguousarray(col, np.int32)),
                                _dp(np.ascontiguousarray(val, np.float64)), _dp(x), k)
    return x.reshape(2, n)


# --------------------------------------------------------------------------
# MPK partition: stage 0 and the recursion into hopeless regions
# --------------------------------------------------------------------------
INT_MAX = 2 ** 31 - 1


def _region(regs, i):
    """getHopelessStartEnd src/matrixPower.cpp:674-683: (start, end) of
    region i of a flat [start, end, ...] list, (-1, -1) outside it."""
    if i >= 0 and 2 * i < len(regs) - 1:
        return regs[2 * i], regs[2 * i + 1]
    return -1, -1


def identify_hopeless_regions(violated, k, total_level):
    """mtxPower::identifyHopelessRegions src/matrixPower.cpp:147-502 for one
    node.  violated: ascending levels whose cache factor exceeds the cut-off.
    Only runs of at least two consecutive violated levels form a region
    (:149-217); regions closer than k-1 levels merge (:225-250); regions are
    kept k-1 levels away from the first and last level, or dropped when
    nothing is left (:268-398).  Returns the flat [start, end, ...] list
    (inclusive level ranges)."""
    cvl = list(violated)
    if len(cvl) > 1:
        cvl.append(cvl[-1] + 2)
    flag = [0] * len(cvl)
    prev, block = -2, 1
    for i, lv in enumerate(cvl):
        if lv == prev + 1:
            flag[i - 1] = block
            flag[i] = block
        else:
            block += 1
        prev = lv
    regs = []
    pf = 0
    for i, f in enumerate(flag):
        if f != pf:
            if f > 0 and pf == 0:
                regs.append(cvl[i])
            elif f > 0 and pf > 0:
                regs += [cvl[i - 1], cvl[i]]
            elif f == 0:
                regs.append(cvl[i - 1])
        pf = f
    # merge regions less than k-1 levels apart
    n_h = len(regs) // 2
    merged = []
    cs, ce = _region(regs, 0)
    ns, ne = _region(regs, 1)
    i = 0
    while i < n_h:
        if cs != -1:
            while (ns - ce - 1) < (k - 1) and ns != -1:
                ce = ne
                ns, ne = _region(regs, i + 2)
                i += 1
            merged += [cs, ce]
            cs, ce = _region(regs, i + 1)
            ns, ne = _region(regs, i + 2)
        i += 1
    # one node [0, total_level): split at its ends (nothing to split), then
    # keep k-1 levels to the first and the last level
    lvl_start, lvl_end = 0, total_level
    split = []
    for i in range(len(merged) // 2):
        cs, ce = _region(merged, i)
        if cs >= lvl_start:
            if cs < lvl_end:
                split += [cs, lvl_end - 1] if ce > lvl_end else [cs, ce]
        else:
            if ce > lvl_end:
                split += [lvl_start, lvl_end - 1]
            elif ce >= lvl_start:
                split += [lvl_start, ce]
    drop = []
    cs, ce = _region(split, 0)
    if cs != -1 and (cs - lvl_start - 1) < (k - 1):
        new_cs = lvl_start + (k - 1)
        if ce < new_cs:
            drop.append(0)
        else:
            split[0] = new_cs
    last = len(split) // 2 - 1
    cs, ce = _region(split, last)
    if ce != -1 and ((lvl_end - 1) - ce - 1) < (k - 1):
        new_ce = lvl_end - (k - 1) - 1
        if cs > new_ce:
            drop.append(last)
        else:
            split[2 * last + 1] = new_ce
    out = []
    for i in range(len(split) // 2):
        if i not in drop:
            out += [split[2 * i], split[2 * i + 1]]
    return out


def consolidate_levels(level_ptr, elem, k, cache_elem, regions):
    """mtxPower::consolidatePartition src/matrixPower.cpp:685-880 for one
    node: consecutive levels merge into a group until the running sum of
    elem reaches cache_elem; around a hopeless region every level from
    k-1 before its start to k-1 after its end is a group of its own, and the
    region's levels form one group.  Returns (group row pointers, group
    first levels, group index of each hopeless region)."""
    total_level = len(level_ptr) - 1
    bl = k - 1
    groups = [level_ptr[0]]
    glevel = [0]
    hid = []
    ctr = 0
    hc = 0
    hs, he = _region(regions, 0)
    sum_elem = 0.0
    cur = 0
    for level in range(total_level):
        e = elem[level]
        sum_elem += e
        hopeless = False
        force = False
        if hs != -1:
            if hs - bl <= level <= hs:
                force = True
            elif he < level <= he + bl:
                force = True
            elif hs < level <= he:
                hopeless = True
            if level == he + bl + 1:
                hc += 1
                hs, he = _region(regions, hc)
                force = True
            if level == hs:
                hid.append(ctr + 1)
        if cur > 0 and not hopeless:
            if sum_elem >= cache_elem or force:
                sum_elem = e
                groups.append(level_ptr[level])
                glevel.append(level)
                ctr += 1
        cur += 1
    groups.append(level_ptr[total_level])
    glevel.append(total_level)
    return groups, glevel, hid


def mpk_partition(graph, k, cache_bytes, safety=2.0, cutoff=50, mtx_type="N", cutoff1=INT_MAX):
    """mtxPowerRecursive::findPartition src/matrixPowerRecursive.cpp:340-445
    with mtxPower::findPartition src/matrixPower.cpp:505-604.  Stage 0:
    traversal (or mtx_type "L" / "U": row i is level i, or n-1-i, no BFS,
    src/traverse_GAP_graphSoA.cpp:316-335), levels whose elem exceeds the
    cache by more than `cutoff` times, hopeless regions, level groups.  For
    mtx_type "N", every hopeless region's group is traversed again with its
    boundary groups (recursivePartition :156-338, working radius
    ceil(k/2) - 1) and grouped with cut-off `cutoff1` (a further stage is
    not restated: NotImplementedError).  Returns dict(levelPtr, groupPtr,
    groupLevel, violated, hopelessRegions, hopelessGroups, children, ...)."""
    if mtx_type in ("L", "U"):
        if mtx_type == "U":
            graph.perm = graph.perm[::-1].copy()
            graph.inv[graph.perm] = np.arange(graph.n, dtype=np.int32)
        lr = np.ones(graph.n, np.int32)
        ln = (graph.rowPtr[graph.perm + 1] - graph.rowPtr[graph.perm]).astype(np.int32)
    else:
        lr, ln = graph.traverse(POWER, 0, graph.n)
    total_level = len(lr)
    level_ptr = [0]
    for r in lr:
        level_ptr.append(level_ptr[-1] + int(r))
    nnzr = graph.nnz_input / float(graph.n)
    byte_per_nnz = (1 + k / nnzr) * 8 + (1 + 1 / nnzr) * 4
    cache_elem = cache_bytes / byte_per_nnz

    def violated_levels(elem, cut):
        out = []
        for L, e in enumerate(elem):
            if e > cache_elem and int(np.ceil(e / cache_elem)) > cut:
                out.append(L)
        return out

    elem = [safety * (k + 1) * int(v) for v in ln]
    violated = violated_levels(elem, cutoff)
    regions = identify_hopeless_regions(violated, k, total_level)
    groups, glevel, hid = consolidate_levels(level_ptr, elem, k, cache_elem, regions)
    children = []
    if mtx_type == "N":
        wbl = int(np.ceil(k / 2.0)) - 1
        for h in hid:
            lo, hi = groups[h], groups[h + 1]
            bounds = []
            for r in range(-wbl, wbl + 1):
                if r != 0 and groups[h + r + 1] > groups[h + r]:
                    bounds.append((groups[h + r], groups[h + r + 1]))
            clr, cln, blds = graph.traverse_boundary(lo, hi, bounds)
            cptr = [lo]
            for r in clr:
                cptr.append(cptr[-1] + int(r))
            celem = []
            for L in range(len(clr)):
                nz = int(cln[L]) + sum(int(b[1][L]) for b in blds)
                celem.append(safety * (k + 1) * nz)
            cviol = violated_levels(celem, cutoff1)
            cregs = identify_hopeless_regions(cviol, k, len(clr))
            if cregs:
                raise NotImplementedError("hopeless regions inside a recursion stage")
            cgroups, _, _ = consolidate_levels(cptr, celem, k, cache_elem, [])
            children.append(dict(lo=lo, hi=hi, group=h, boundaries=bounds, groupPtr=cgroups,
                                 levelRow=clr, levelNnz=cln))
    return dict(levelPtr=level_ptr, groupPtr=groups, groupLevel=glevel, violated=violated, hopeless=violated,
                hopelessRegions=regions, hopelessGroups=hid, children=children,
                levelRow=lr, levelNnz=ln, cache_elem=cache_elem, byte_per_nnz=byte_per_nnz)


# ---------------------------------------------------------------------------
# SELL-C-sigma helpers.  Entry j of row r at chunkStart[r // C] + r % C + j C.

def sell_c_sigmize(calls, simdWidth, C, col, chunkStart, rl, clp):
    """src/sell_c_sigmize.cpp:26-62 run as a RACE callback: calls are the
    (start, end) leaf calls of the forward colouring (leaf_calls)."""
    for start, end in calls:
        for row in range(start, end):
            chunk = row // C
            base = chunkStart[chunk] + row % C
            for j in range(rl[row], clp[chunk]):
                col[base + j * C] = start + ((row - start) % simdWidth)


def simdify_resolve(simdWidth, C, nrows, col, chunkStart, clp, val, diagFirst=False):
    """include/simdify.h:70-357 without the sell_c_sigmize call (:78);
    returns False where the reference returns false."""
    nChunk = nrows // C
    simdInChunk = C // simdWidth
    startingIdx = 1 if diagFirst else 0
    simdCol = [0] * simdWidth
    for chunk in range(nChunk + 1):
        if chunk * simdInChunk * simdWidth >= nrows:
            break
        for simdIdx in range(simdInChunk):
            base = chunkStart[chunk] + simdIdx * simdWidth
            for j in range(startingIdx, clp[chunk]):
                row = chunk * simdInChunk * simdWidth + simdIdx * simdWidth
                i = 0
                while i < simdWidth and row < nrows:
                    pos = j                      # currCol points at position pos, lane i
                    ctr = 0
                    wrapped = False
                    repeat = True
                    while repeat:
                        repeat = any(col[base + C * pos + i] == simdCol[k] for k in range(i))
                        if wrapped:
                            here = col[base + C * j + i]
                            for k in range(simdWidth):
                                if k != i and here == col[base + C * pos + k]:
                                    repeat = True
                                    break
                        if repeat:
                            if ctr < clp[chunk]:
                                if pos != clp[chunk] - 1:
                                    pos += 1
                                else:
                                    pos = startingIdx
                                    wrapped = True
                                ctr += 1
                            else:
                                return False
                    simdCol[i] = col[base + C * pos + i]
                    if pos != j:
                        a, b = base + C * j + i, base + C * pos + i
                        col[a], col[b] = col[b], col[a]
                        val[a], val[b] = val[b], val[a]
                    i += 1
                    row += 1
    return True


def simdify_d1(simdWidth, C, nrows, col, chunkStart, cl, val, dist2Compatible=False):
    """include/simdify.h:389-447, including its row counter that carries
    over between SIMD groups."""
    nChunk = nrows // C
    simdInChunk = C // simdWidth
    row = 0
    for chunk in range(nChunk + 1):
        for simdIdx in range(simdInChunk):
            i = 1
            while i < simdWidth and row < nrows:
                row = chunk * simdInChunk * simdWidth + simdIdx * simdWidth + i
                dep = [0] * i
                for k in range(i):
                    dep[k if dist2Compatible else i - k - 1] = row - (k + 1)
                lane = chunkStart[chunk] + simdIdx * simdWidth + i
                for j in range(cl[chunk]):
                    for k in range(i):
                        if col[lane + C * j] == dep[k]:
                            e = cl[chunk] - (k + 1)
                            if e != j:
                                a, b = lane + C * j, lane + C * e
                                col[a], col[b] = col[b], col[a]
                                val[a], val[b] = val[b], val[a]
                i += 1
                row += 1
    return True


# ---------------------------------------------------------------------------
# The layout tuner's MLP (no counterpart in the reference): 16-64-64-12,
# ReLU hidden layers, masked mean-squared error.  Parameters concatenated
# W1 [16][64], c1, W2 [64][64], c2, W3 [64][12], c3 (row-major).

def mlp_unpack(p, n_in=16, n_h=64, n_out=12):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((n_in, n_h), (n_h,), (n_h, n_h), (n_h,), (n_h, n_out), (n_out,)):
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    assert o == len(p)
    return out


def mlp_forward(p, X):
    W1, c1, W2, c2, W3, c3 = mlp_unpack(p)
    h1 = np.maximum(np.asarray(X, np.float64) @ W1 + c1, 0.0)
    h2 = np.maximum(h1 @ W2 + c2, 0.0)
    return h2 @ W3 + c3


def mlp_loss_grad(p, X, T, mask):
    """sum(mask (y - t)^2) / sum(mask) and its gradient (back-propagation)."""
    W1, c1, W2, c2, W3, c3 = mlp_unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ W1 + c1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + c2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3 + c3
    ms = float(np.sum(mask))
    d = (y - T) * mask
    loss = float(np.sum(d * d) / ms)
    dy = 2.0 * d / ms
    gW3, gc3 = h2.T @ dy, dy.sum(axis=0)
    d2 = (dy @ W3.T) * (h2 > 0)
    gW2, gc2 = h1.T @ d2, d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    gW1, gc1 = X.T @ d1, d1.sum(axis=0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gc1, gW2, gc2, gW3, gc3)])
    return loss, g

