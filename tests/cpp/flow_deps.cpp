// Host-only check of the one-launch sweeps' dependency tables
// (race_amd/csrc/bundle.cpp: chainDeps, flowExt), built with g++ by
// tests/test_host_logic.py.  On a 5-point grid coloured by lines (Gauss-
// Seidel: lines by parity; Kaczmarz: lines mod 3), with bundles small enough
// that every line becomes a chain of several bundles, forward and backward:
//  * every dependency points to an earlier chain and none lies in a phase;
//  * every pair of chains that conflict, computed directly from the matrix
//    (GS: one writes a row the other reads; KACZ: both touch a column), is
//    ordered by the dependencies (the earlier one is an ancestor);
//  * the coherent-read bit marks exactly the columns an earlier chain, or an
//    earlier bundle of the same chain, writes (GS) or touches (KACZ);
//  * with one dependency removed the check reports a violation (the check
//    is not vacuous);
//  * the longest dependency paths (chainLongestPaths) equal a fixpoint of
//    path[a] >= cost[a] + path[c] over every dependency a of c, relaxed in
//    no particular order, and a chain no other depends on costs its own
//    bundles only.
// Prints "flow deps ok" and exits 0 on success.
#include <algorithm>
#include <cstdio>
#include <set>
#include <vector>

#include "bundle.hpp"
#include "race_amd.h"

using namespace race;

namespace {

struct Grid {
    int h, w, n;
    std::vector<int> rp, col;
    std::vector<double> val;
};

// 5-point stencil, diagonal first in every row
Grid grid(int h, int w) {
    Grid g{h, w, h * w, {0}, {}, {}};
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            const int r = i * w + j;
            g.col.push_back(r);
            g.val.push_back(4.0);
            const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            for (int q = 0; q < 4; ++q) {
                const int a = i + di[q], b = j + dj[q];
                if (a < 0 || a >= h || b < 0 || b >= w) continue;
                g.col.push_back(a * w + b);
                g.val.push_back(-1.0);
            }
            g.rp.push_back((int)g.col.size());
        }
    return g;
}

// leaves = grid lines, phases = line index modulo ncol, in colour order
// (reversed for the backward sweep)
Schedule lines(const Grid& g, int ncol, bool rev) {
    Schedule s;
    s.rev = rev;
    s.phasePtr.push_back(0);
    for (int q = 0; q < ncol; ++q) {
        const int p = rev ? ncol - 1 - q : q;
        for (int i = 0; i < g.h; ++i)
            if (i % ncol == p) {
                s.leafLo.push_back(i * g.w);
                s.leafHi.push_back((i + 1) * g.w);
            }
        s.phasePtr.push_back((int)s.leafLo.size());
    }
    return s;
}

int failures = 0;
#define EXPECT(c, ...)                   \
    do {                                 \
        if (!(c)) {                      \
            ++failures;                  \
            if (failures < 20) {         \
                std::printf(__VA_ARGS__); \
                std::printf("\n");       \
            }                            \
        }                                \
    } while (0)

// the conflicting chain pairs not ordered by the dependencies
long unordered(const Grid& g, const BundleTable& t, int kind, const std::vector<int>& dp,
               const std::vector<int>& dd) {
    const int nc = (int)t.chainPtr.size() - 1;
    std::vector<std::set<int>> W(nc), R(nc);
    for (int c = 0; c < nc; ++c)
        for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b)
            for (int s = t.slot[b]; s < t.slot[b + 1]; ++s) {
                const int r = t.rowId[s];
                for (int j = g.rp[r]; j < g.rp[r + 1]; ++j) {
                    const int cl = g.col[j];
                    if (kind == RACE_K_KACZ) {
                        W[c].insert(cl);
                        R[c].insert(cl);
                    } else if (cl != r) {
                        R[c].insert(cl);
                    }
                }
                if (kind != RACE_K_KACZ) W[c].insert(r);
            }
    auto meets = [](const std::set<int>& a, const std::set<int>& b) {
        for (int v : a)
            if (b.count(v)) return true;
        return false;
    };
    long bad = 0;
    for (int b = 0; b < nc; ++b) {
        // ancestors of b through the dependencies
        std::vector<char> anc(nc, 0);
        std::vector<int> stack(dd.begin() + dp[b], dd.begin() + dp[b + 1]);
        while (!stack.empty()) {
            const int a = stack.back();
            stack.pop_back();
            if (anc[a]) continue;
            anc[a] = 1;
            for (int j = dp[a]; j < dp[a + 1]; ++j) stack.push_back(dd[j]);
        }
        for (int a = 0; a < b; ++a)
            if ((meets(W[a], R[b]) || meets(R[a], W[b]) || meets(W[a], W[b])) && !anc[a]) ++bad;
    }
    return bad;
}

void check(const Grid& g, int kind, bool rev) {
    const int ncol = kind == RACE_K_KACZ ? 3 : 2;
    const Schedule s = lines(g, ncol, rev);
    // at most 4 rows per bundle: every line of 10 rows is a chain of bundles
    const BundleTable t = buildBundles(s, kind, g.n, g.rp.data(), g.col.data(), g.val.data(), 4, 4096, 1, 256);
    EXPECT(t.ok, "bundle table refused");
    const int nc = (int)t.chainPtr.size() - 1;
    int longChains = 0;
    for (int c = 0; c < nc; ++c) longChains += t.chainPtr[c + 1] - t.chainPtr[c] > 1;
    EXPECT(longChains > 0, "no chain of several bundles");
    std::vector<int> dp, dd;
    const long same = chainDeps(t, g.n, nullptr, &dp, &dd);
    EXPECT(same == 0, "kind %d rev %d: %ld dependencies inside a phase", kind, (int)rev, same);
    for (int c = 0; c < nc; ++c)
        for (int j = dp[c]; j < dp[c + 1]; ++j) EXPECT(dd[j] < c, "dependency %d of chain %d not earlier", dd[j], c);
    const long bad = unordered(g, t, kind, dp, dd);
    EXPECT(bad == 0, "kind %d rev %d: %ld conflicting chain pairs not ordered", kind, (int)rev, bad);
    // one dependency removed: the check must see it
    int victim = -1;
    for (int c = 0; c < nc && victim < 0; ++c)
        if (dp[c + 1] > dp[c]) victim = c;
    std::vector<int> dp2(dp), dd2;
    for (int c = 0; c < nc; ++c)
        for (int j = dp[c]; j < dp[c + 1]; ++j)
            if (!(c == victim && j == dp[c])) dd2.push_back(dd[j]);
    for (int c = victim; c < nc; ++c) --dp2[c + 1];
    EXPECT(unordered(g, t, kind, dp2, dd2) > 0, "kind %d: a removed dependency went unnoticed", kind);
    // longest paths against a fixpoint relaxation (no use of chain order)
    const int fixed = 4;
    const std::vector<long> lp = chainLongestPaths(t, dp, dd, fixed);
    EXPECT((int)lp.size() == nc, "longest paths: %d for %d chains", (int)lp.size(), nc);
    std::vector<long> cost(nc, 0), fp(nc, 0);
    std::vector<char> hasSucc(nc, 0);
    for (int c = 0; c < nc; ++c) {
        for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b) cost[c] += t.wave[b + 1] - t.wave[b] + fixed;
        fp[c] = cost[c];
    }
    for (bool changed = true; changed;) {
        changed = false;
        for (int c = nc - 1; c >= 0; --c)  // against the chain order on purpose
            for (int j = dp[c]; j < dp[c + 1]; ++j) {
                const int a = dd[j];
                hasSucc[a] = 1;
                if (cost[a] + fp[c] > fp[a]) {
                    fp[a] = cost[a] + fp[c];
                    changed = true;
                }
            }
    }
    long longest = 0;
    for (int c = 0; c < nc && c < (int)lp.size(); ++c) {
        EXPECT(lp[c] == fp[c], "kind %d rev %d: chain %d path %ld, fixpoint %ld", kind, (int)rev, c, lp[c], fp[c]);
        if (!hasSucc[c]) EXPECT(lp[c] == cost[c], "chain %d without successors: path %ld, cost %ld", c, lp[c], cost[c]);
        longest = std::max(longest, lp[c]);
    }
    EXPECT(longest > cost[0], "kind %d rev %d: no path longer than one chain", kind, (int)rev);
    // coherent-read bits
    const std::vector<int> fx = flowExt(t, g.n, nullptr);
    std::vector<int> chainOf(t.slot.size() - 1), owner(g.n, -1);
    for (int c = 0; c < nc; ++c)
        for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b) {
            chainOf[b] = c;
            for (int q = t.slot[b]; q < t.slot[b + 1]; ++q) owner[t.rowId[q]] = b;
        }
    std::vector<int> firstTouch(g.n, 1 << 30);
    for (int b = 0; b + 1 < (int)t.extPtr.size(); ++b)
        for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) firstTouch[t.ext[e]] = std::min(firstTouch[t.ext[e]], b);
    for (int b = 0; b + 1 < (int)t.extPtr.size(); ++b)
        for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) {
            const int cl = t.ext[e];
            const bool marked = ((unsigned)fx[e] >> 31) != 0;
            EXPECT((fx[e] & 0x7fffffff) == cl, "column id changed");
            bool need;
            if (kind == RACE_K_KACZ)
                need = firstTouch[cl] < b;  // an earlier bundle touched it
            else
                need = owner[cl] >= 0 && chainOf[owner[cl]] <= chainOf[b];  // written before or by this chain
            EXPECT(marked == need, "kind %d rev %d: column %d of bundle %d marked %d, expected %d", kind, (int)rev, cl,
                   b, (int)marked, (int)need);
        }
}

}  // namespace

int main() {
    const Grid g = grid(9, 10);
    for (int kind : {RACE_K_GS, RACE_K_KACZ})
        for (bool rev : {false, true}) check(g, kind, rev);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("flow deps ok\n");
    return 0;
}
