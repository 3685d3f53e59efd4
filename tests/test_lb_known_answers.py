"""Known answers for the level-granular scheduling rules, derived by hand.

Every expected value below follows from the rules of the reference, traced
step by step on paper for the given histogram (the trace is summarised in
each docstring); none of them comes from running the oracle or the product.
Each case runs both the oracle (oracle/race_oracle.py) and the product's
load balancer (race_lb_split / race_split_threads in librace_amd.so):

  LB::splitZones         src/lb.cpp:188-391 (zones, thread fix, blocks)
  variance balancing     src/lb.cpp:408-566 with Stat src/lb.cpp:773-795
  calculateIdealNthreads src/level_recursion.cpp:149-263
  getBytePerNNZ, the level grouping of consolidatePartition
                         src/matrixPower.cpp:134-140, 685-880

Histograms are chosen so that every effRatio = levelNnz * T / total is a
dyadic fraction: the accumulations are exact and no comparison sits on a
rounding boundary.
"""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

ONE, TWO = 1, 2
TWO_BLOCK, THREE_BLOCK = 0, 1
LB_ROW, LB_NNZ = 0, 1


def both(oracle, lr, ln, T, eff, dist, d2=TWO_BLOCK, target=LB_NNZ):
    a = oracle.split_zones(lr, ln, T, eff, dist, d2, target)
    b = R.lb_split(lr, ln, T, eff, dist, d2, target)
    return a, b


def check(oracle, lr, ln, T, eff, dist, lp, scale, d2=TWO_BLOCK, target=LB_NNZ):
    a, b = both(oracle, lr, ln, T, eff, dist, d2, target)
    for name, r in (("oracle", a), ("product", b)):
        assert list(r["levelPtr"]) == lp, (name, r["levelPtr"])
        assert list(r["scale"]) == scale, (name, r["scale"])
        rows = [int(sum(lr[lp[k]:lp[k + 1]])) for k in range(len(lp) - 1)]
        assert list(r["zone_rows"]) == rows, name


@pytest.mark.parametrize("L", [1, 2, 3, 7, 10, 25, 64])
@pytest.mark.parametrize("dist,d2,bpt", [(ONE, TWO_BLOCK, 2), (TWO, TWO_BLOCK, 2), (TWO, THREE_BLOCK, 3)])
@pytest.mark.parametrize("eff", [0.0, 40.0, 95.0])
def test_one_thread_cuts_blocks_at_rounded_fractions(oracle, L, dist, d2, bpt, eff):
    """T = 1: the effRatios add up to 1, so the only zone is accepted at the
    last level (the lookahead runs to the closest point to 1, which is the
    end; or the last-5-levels rule takes everything), scale = [1], and the
    blocks are cut at roundDouble(j L / bpt), j = 1..bpt-1 (src/lb.cpp:
    384-391).  Each colour partition then holds one block, so every
    variance is 0 and no move is kept (src/lb.cpp:554-560)."""
    rng = np.random.default_rng(L * 7 + bpt)
    ln = rng.integers(1, 50, L)
    lr = rng.integers(1, 9, L)
    lp = [0] + [int(np.floor(j * L / bpt + 0.5)) for j in range(1, bpt)] + [L]
    check(oracle, lr, ln, 1, eff, dist, lp, [1], d2)


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("T", [2, 3, 8, 64])
@pytest.mark.parametrize("dist", [ONE, TWO])
def test_at_most_five_levels_go_to_one_thread(oracle, L, T, dist):
    """totalLevel <= 5: at i = 0 the counter is below minGap * blocks (>= 2),
    so no zone is accepted, and i >= totalLevel - 5 already holds: the
    last-5-levels rule (src/lb.cpp:289-314) puts all levels into one zone
    whose scale is roundDouble(sum effRatio) = T.  One thread with T nested
    threads, its two blocks cut at roundDouble(L / 2)."""
    ln = np.arange(1, L + 1) * 3
    lr = np.arange(1, L + 1)
    lp = [0, int(np.floor(L / 2 + 0.5)), L]
    check(oracle, lr, ln, T, 40.0, dist, lp, [T])


def test_integer_prefix_sums_fix_the_zone_cuts(oracle):
    """T = 3, distance 1, efficiency 40, levelNnz [4,4,4,4,8,8,4,4,4,4]
    (total 48): effRatio = levelNnz / 16 = [.25 x4, .5, .5, .25 x4].
    i = 1: acc .5 is accepted (50 > 40); the lookahead improves to 0 at
    level 3 and stops after level 5 (|new| > .5): zone [0,4).  i = 5: acc
    1.0, zone [4,6).  i = 5 >= 10 - 5: the rest (1.0) is the last zone.
    scale [1,1,1]; blocks halve each zone: [0,2,4,5,6,8,10].  Every block
    then carries 8 nnz (zero variance); no move lowers it."""
    ln = np.array([4, 4, 4, 4, 8, 8, 4, 4, 4, 4])
    lr = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
    check(oracle, lr, ln, 3, 40.0, ONE, [0, 2, 4, 5, 6, 8, 10], [1, 1, 1])


def test_row_target_reads_level_rows(oracle):
    """lbTarget ROW balances levelRow instead of levelNnz
    (LB::calcEffectiveRatio src/lb.cpp:114-131): the histogram of the
    previous case given as rows, with an unrelated nnz histogram, gives the
    same split."""
    lr = np.array([4, 4, 4, 4, 8, 8, 4, 4, 4, 4])
    ln = np.array([9, 1, 7, 3, 2, 8, 5, 5, 1, 30])
    check(oracle, lr, ln, 3, 40.0, ONE, [0, 2, 4, 5, 6, 8, 10], [1, 1, 1], target=LB_ROW)


LN_EFF = np.array([3, 3, 4, 1, 1, 1, 1, 1, 1])  # T = 2, total 16: effRatio = levelNnz / 8


def test_efficiency_just_below_the_first_zone(oracle):
    """T = 2, efficiency 74, effRatio [.375,.375,.5,.125 x6].
    i = 1: acc .75, (1 - .25) 100 = 75 > 74: accepted; the lookahead's
    +.25 at level 2 only ties |-.25| and the earlier point is kept
    (strictly-closer rule; the 1e-6 branch needs new < min, impossible for
    growing sums): zone [0,2), currRem -.25.  i = 3: acc .625 (62.5, no);
    i = 4: acc .75, accepted, lookahead to 1.0 at level 6: zone [2,7);
    i = 6 >= 4: the rest .25 rounds to 0 and joins the last zone.
    scale [1,1], blocks [0,1,2,6,9], chunk sums [3,3,7,3].
    Variance (src/lb.cpp:408-566): colour 0 has (3,7), mean 5, var 8;
    block 0 (below its mean) acquires from block 2 (highest acquire weight
    with more than minGap levels): [0,2,3,6,9], chunk sums [6,4,3,3],
    variance 5 < 8, kept.  From there every single move raises the
    variance (8, 8, -, 5 then 34), so the final split is [0,2,3,6,9]."""
    lr = np.ones(9, int)
    check(oracle, lr, LN_EFF, 2, 74.0, ONE, [0, 2, 3, 6, 9], [1, 1])


@pytest.mark.parametrize("eff", [75.0, 76.0])
def test_efficiency_at_or_above_the_first_zone(oracle, eff):
    """The same histogram with efficiency 75 or 76: (1 - .25) 100 = 75 is
    not > eff at i = 1 or 2, i = 3 gives 62.5 and i = 4 gives 50 (acc 1.5
    rounds to 2); i = 4 >= 9 - 5, so the last-5-levels rule takes all levels
    as one zone of scale roundDouble(2.0) = 2.  One thread, blocks cut at
    roundDouble(4.5) = 5; one block per colour, zero variance."""
    lr = np.ones(9, int)
    check(oracle, lr, LN_EFF, 2, eff, ONE, [0, 5, 9], [2])


def test_thread_fix_adds_one_to_the_last_zone(oracle):
    """T = 4, distance 2 (two blocks, minGap 2), levelNnz [11 x8, 10 x4]
    (total 128): effRatio 11/32 and 10/32.  Zones close after 4 levels
    (acc 1.375 -> 1, the lookahead only moves away): [0,4), [4,8); at
    i = 7 >= 12 - 5 the last zone takes 1.25 -> 1.  scale [1,1,1] sums to
    T - 1, so the last zone gets one more (src/lb.cpp:330-340): [1,1,2].
    Every block has exactly minGap levels, so the variance step can neither
    give nor acquire: [0,2,4,6,8,10,12]."""
    ln = np.array([11] * 8 + [10] * 4)
    lr = np.arange(12) + 1
    check(oracle, lr, ln, 4, 40.0, TWO, [0, 2, 4, 6, 8, 10, 12], [1, 1, 2])


def test_thread_fix_takes_one_from_the_last_zone(oracle):
    """T = 3, distance 2, levelNnz [3,3,3,3,8,2,1,1] (total 24): effRatio
    = levelNnz / 8.  i = 3: acc 1.5 -> 2, the lookahead reaches only |.5|
    (ties keep the earlier point): zone [0,4); i = 3 >= 8 - 5: the rest
    (1.5 -> 2) is the last zone.  scale [2,2] sums to T + 1 and the last is
    not 1, so it becomes [2,1] (src/lb.cpp:341-356).  Blocks of minGap
    levels: no variance move; [0,2,4,6,8]."""
    ln = np.array([3, 3, 3, 3, 8, 2, 1, 1])
    lr = np.ones(8, int)
    check(oracle, lr, ln, 3, 40.0, TWO, [0, 2, 4, 6, 8], [2, 1])


def test_thread_fix_drops_a_last_zone_of_one(oracle):
    """T = 4, distance 2, levelNnz [3,3,3,3,8,4,2,2,1,1,1,1] (total 32):
    zones [0,4) (1.5 -> 2), [4,8) (2.0 -> 2), last [8,12) (.5 -> 1): the sum
    5 = T + 1 with a last scale of 1 drops that zone and merges its levels
    into the previous one (src/lb.cpp:343-350): scale [2,2], zones [0,4),
    [4,12), blocks [0,2,4,8,12] with chunk sums [6,6,16,4].
    Variance: colour 0 has (3, 8) per thread, mean 5.5, var 12.5; colour 1
    (3, 2), var .5.  Block 0 acquires from block 2: [0,3,5,8,12], chunk
    sums [9,11,8,4], variance .125 + 6.125 = 6.25 < 13, kept.  Next round:
    block 1 cannot give (minGap levels), block 3 acquiring from block 0
    gives 8, block 0 giving to block 3 gives 8, block 2 acquiring from
    block 0 restores 13: no improvement, final [0,3,5,8,12]."""
    ln = np.array([3, 3, 3, 3, 8, 4, 2, 2, 1, 1, 1, 1])
    lr = np.ones(12, int)
    check(oracle, lr, ln, 4, 40.0, TWO, [0, 3, 5, 8, 12], [2, 2])


# ---------------------------------------------------------------------------
# calculateIdealNthreads
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("weights,total,T,expect", [
    # equal weights: 2.5 each, floor 2, remainders .5 tie -> stable order
    ([5, 5, 5, 5], 20, 10, [3, 3, 2, 2]),
    # 0.7, 1.4, 2.1, 2.8 -> floor [0,1,2,2]; the zero share ranks first
    # (remainder 1), then .8, .4, .1: two leftover threads to groups 0, 3
    ([10, 20, 30, 40], 100, 7, [1, 1, 2, 3]),
    # exact shares: no leftover
    ([1, 3], 4, 8, [2, 6]),
    # every share below one thread: all rank 1, leftover goes in group order
    ([1, 1, 1, 1, 1], 5, 3, [1, 1, 1, 0, 0]),
    # 4.5, 2.7, 1.8 -> [4,2,1]; remainders .5, .7, .8: the two leftover
    # threads go to groups 2 and 1 (largest fractions), not to group 0
    # (largest share)
    ([50, 30, 20], 100, 9, [4, 3, 2]),
    # weights not summing to the total (nnz outside the groups' own rows):
    # 6 x 40/100 = 2.4 each -> [2,2], leftover 2 -> [3,3]
    ([40, 40], 100, 6, [3, 3]),
])
def test_ideal_nthreads_known_answers(oracle, weights, total, T, expect):
    assert oracle.split_threads(weights, total, T) == expect
    assert R.split_threads(weights, total, T) == expect


def test_ideal_nthreads_is_what_the_recursion_assigns(oracle):
    """Every partitioned node of the oracle's zone tree hands its child
    groups the thread split of their nnz (the wiring of
    calculateIdealNthreads into LevelRecursion::recursivePartition,
    src/level_recursion.cpp:265-275)."""
    rp, col = M.stencil27(10)
    for T in (12, 40, 100):
        G = oracle.Graph(rp, col, symm_hint=True)
        zt = oracle.level_recursion(G, T, ONE)
        lens = np.diff(G.rowPtr)[G.perm]
        checked = 0
        for p, leaf in enumerate(zt.tree):
            if p == 0 or not leaf.children or leaf.idealNthreadsZ <= 1:
                continue
            ch = leaf.children
            for sb in range(len(ch) // 2):
                nt = (ch[2 * sb + 1] - ch[2 * sb]) // zt.bpt
                s, e = leaf.valueZ[sb], leaf.valueZ[sb + 1]
                w = []
                for i in range(nt):
                    ss = zt.tree[ch[2 * sb] + i * zt.bpt].valueZ[0]
                    ee = zt.tree[ch[2 * sb] + zt.bpt * (i + 1) - 1].valueZ[-1]
                    w.append(int(lens[ss:ee].sum()))
                got = [zt.tree[ch[2 * sb] + zt.bpt * i].idealNthreadsZ for i in range(nt)]
                if got == [1] * nt and sum(oracle.split_threads(w, int(lens[s:e].sum()), leaf.idealNthreadsZ)) != nt:
                    continue  # a node below the recursion's last split keeps the defaults
                assert got == oracle.split_threads(w, int(lens[s:e].sum()), leaf.idealNthreadsZ)
                checked += 1
        assert checked > 0 or T == 12


# ---------------------------------------------------------------------------
# MPK byte model and level grouping
# ---------------------------------------------------------------------------
def path_graph(n):
    rp = [0]
    col = []
    for r in range(n):
        row = [r] + [c for c in (r - 1, r + 1) if 0 <= c < n]
        col += row
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32)


def test_mpk_byte_model_and_groups_by_hand(oracle):
    """Path of 10 rows (tridiagonal, 28 entries, nnzr 2.8), k = 2, safety 2:
    bytePerNNZ = (1 + 2/2.8) 8 + (1 + 1/2.8) 4 = 12 + 20/2.8 = 19.142857
    (src/matrixPower.cpp:134-140); 1168 bytes hold 61.015 elements.
    Every BFS level from row 0 is one row, levelNnz [2,3,...,3,2] (a row
    counts its diagonal, src/graph_SoA.cpp:92-95), so elem = 2 (k+1) nnz =
    [12,18,...,18,12].  Consolidation (src/matrixPower.cpp:795-858) closes
    a group before the level whose running sum reaches cacheElem:
    12, 30, 48, 66 -> cut at 3; 18, 36, 54, 72 -> cut at 6; 18, 36, 54, 66
    -> cut at 9: groups [0,3,6,9,10].  No level exceeds the cache, so there
    are no hopeless levels."""
    rp, col = path_graph(10)
    G = oracle.Graph(rp, col, symm_hint=True)
    part = oracle.mpk_partition(G, 2, 1168.0)
    assert abs(part["byte_per_nnz"] - (12 + 20 / 2.8)) < 1e-12
    assert abs(part["cache_elem"] - 1168.0 / (12 + 20 / 2.8)) < 1e-9
    assert list(part["levelNnz"]) == [2] + [3] * 8 + [2]
    assert part["groupLevel"] == [0, 3, 6, 9, 10]
    assert part["groupPtr"] == [0, 3, 6, 9, 10]
    assert part["hopeless"] == []


def test_mpk_hopeless_level_by_hand(oracle):
    """A star: row 0 joined to rows 1..40 (121 pattern entries over 41
    rows, nnzr 121/41), k = 3, safety 2.  Levels from row 0: {0} with 41
    entries, {1..40} with 80.  bytePerNNZ = (1 + 3 41/121) 8 + (1 + 41/121)
    4 = 21.488; 100 bytes hold cacheElem = 4.654 elements; elem = 2 (k+1)
    nnz = 328 and 640, factors ceil(70.5) = 71 and ceil(137.5) = 138
    (src/matrixPower.cpp:540-565): both exceed the default cut-off 50, so
    both levels are hopeless; with cut-off 100 only level 1 is."""
    n = 41
    rp = [0]
    col = list(range(n))
    rp.append(n)
    for r in range(1, n):
        col += [0, r]
        rp.append(len(col))
    rp, col = np.array(rp, np.int32), np.array(col, np.int32)
    G = oracle.Graph(rp, col, symm_hint=True)
    bpn = (1 + 3 * 41 / 121) * 8 + (1 + 41 / 121) * 4
    part = oracle.mpk_partition(G, 3, 100.0)
    assert abs(part["byte_per_nnz"] - bpn) < 1e-12
    assert list(part["levelNnz"]) == [41, 80]
    assert part["violated"] == [0, 1]
    # the run [0, 1] lies within k-1 = 2 levels of the first level: the
    # clamp (src/matrixPower.cpp:320-334) drops it, so no hopeless region
    assert part["hopelessRegions"] == [] and part["hopelessGroups"] == []
    assert oracle.mpk_partition(G, 3, 100.0, cutoff=100)["violated"] == [1]


# ---------------------------------------------------------------------------
# hopeless regions and the level grouping around them
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("violated,k,L,expect", [
    # a single violated level is no region (only runs of consecutive levels
    # get a flag, src/matrixPower.cpp:149-217)
    ([7], 3, 20, []),
    ([3, 7, 12], 3, 20, []),
    # a run of two
    ([5, 6], 3, 20, [5, 6]),
    # two runs 1 level apart merge when k-1 > 1 (src/matrixPower.cpp:225-250)
    ([5, 6, 8, 9], 4, 30, [5, 9]),
    ([5, 6, 8, 9], 2, 30, [5, 6, 8, 9]),
    # adjacent runs with one gap level: kept apart for k = 2, merged for 3
    ([3, 4, 6, 7], 2, 20, [3, 4, 6, 7]),
    ([3, 4, 6, 7], 3, 20, [3, 7]),
    # isolated levels between runs are ignored
    ([2, 5, 6, 9, 12, 13, 14], 2, 30, [5, 6, 12, 14]),
    # start moved to k-1 (src/matrixPower.cpp:320-334) ...
    ([1, 2, 3, 4], 4, 20, [3, 4]),
    # ... or the region dropped when nothing is left of it
    ([1, 2], 4, 20, []),
    # end moved to L-1-(k-1) (src/matrixPower.cpp:336-352)
    ([15, 16, 17, 18], 3, 20, [15, 17]),
    # both ends
    (list(range(1, 19)), 3, 20, [2, 17]),
])
def test_hopeless_regions_known_answers(oracle, violated, k, L, expect):
    assert oracle.identify_hopeless_regions(violated, k, L) == expect
    assert R.mpk_hopeless_regions(violated, k, L) == expect


@pytest.mark.parametrize("cache,groups,hid", [
    # a cache no sum of levels reaches: only the forced cuts.  k = 3 (k-1 =
    # 2), region [5, 6]: every level of [3, 5] and of [7, 8] starts a group,
    # level 9 = end + k starts one (and ends the region's rules), levels
    # 5..6 are one group (src/matrixPower.cpp:769-858): group 3 is hopeless
    (1e9, [0, 3, 4, 5, 7, 8, 9, 12], [3]),
    # elem 1 per level, cacheElem 2.5: a cut also where the running sum
    # reaches 2.5 (before level 2 and level 11); the region is group 4
    (2.5, [0, 2, 3, 4, 5, 7, 8, 9, 11, 12], [4]),
])
def test_consolidation_around_a_hopeless_region(oracle, cache, groups, hid):
    lp = list(range(13))  # one row per level, 12 levels
    elem = [1.0] * 12
    g, gl, h = oracle.consolidate_levels(lp, elem, 3, cache, [5, 6])
    assert g == groups and gl == groups and h == hid
    assert R.mpk_consolidate(lp, elem, 3, cache, [5, 6]) == (groups, groups, hid)
