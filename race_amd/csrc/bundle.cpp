// bundle.cpp -- host construction of the leaf-bundle layout.
#include "bundle.hpp"

#include <algorithm>

#include "../../include/race_amd.h"

namespace race {

namespace {

// Dependency depth of every row inside its leaf, for the leaf's sequential
// order (forward: lo..hi-1, reverse: hi-1..lo).
//  GS   : row r writes x[r] and reads x[c], c != r.
//  KACZ : row r reads and writes x[c] for every column c of the row.
// Rows of equal depth in a leaf are independent; running the depths in
// order reproduces the sequential result exactly.
void leafDepths(const Schedule& s, int kind, int n, const int* rp, const int* col,
                std::vector<int>& depth) {
    depth.assign(n, 0);
    std::vector<int> W(n, 0), R(n, 0), stamp(n, -1);
    for (size_t l = 0; l < s.leafLo.size(); ++l) {
        const int lo = s.leafLo[l], hi = s.leafHi[l];
        const int tag = (int)l;
        for (int k = 0; k < hi - lo; ++k) {
            const int r = s.rev ? hi - 1 - k : lo + k;
            int d = 0;
            if (kind == RACE_K_KACZ) {
                for (int j = rp[r]; j < rp[r + 1]; ++j)
                    if (stamp[col[j]] == tag) d = std::max(d, W[col[j]]);
                ++d;
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    stamp[col[j]] = tag;
                    W[col[j]] = d;
                }
            } else {
                // true dependency on earlier writers of what r reads, anti
                // dependency on earlier readers of what r writes
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if (c != r && stamp[c] == tag) d = std::max(d, W[c]);
                }
                if (stamp[r] == tag) d = std::max(d, R[r]);
                ++d;
                if (stamp[r] != tag) {
                    stamp[r] = tag;
                    R[r] = 0;
                }
                W[r] = d;
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if (c == r) continue;
                    if (stamp[c] != tag) {
                        stamp[c] = tag;
                        W[c] = 0;
                        R[c] = 0;
                    }
                    R[c] = std::max(R[c], d);
                }
            }
            depth[r] = d;
        }
    }
}

}  // namespace

BundleTable buildBundles(const Schedule& s, int kind, int n, const int* rp, const int* col,
                         const double* val, int cap, int localCap, int tpr, int laneCap) {
    BundleTable t;
    t.kind = kind;
    t.cap = cap;
    t.tpr = tpr;
    const bool var = tpr == 0;
    // lanes of a row in the variable layout (entries without the GS diagonal)
    auto rowLanes = [&](int r) {
        const int cnt = rp[r + 1] - rp[r] - (kind != RACE_K_KACZ ? 1 : 0);
        return std::max(1, (cnt + kVarE - 1) / kVarE);
    };
    if (var) t.lanePtr.push_back(0);
    localCap = std::min(localCap, 65535);
    const bool gs = kind != RACE_K_KACZ;
    std::vector<int> depth;
    leafDepths(s, kind, n, rp, col, depth);

    // owner[r]: bundle whose rows include r; seen[c] tags columns already in
    // the local space of the bundle being filled (tags are never reset)
    std::vector<int> owner(n, -1), seen(n, -1), localOf(n, -1);
    t.slot.push_back(0);
    t.wave.push_back(0);
    t.extPtr.push_back(0);

    // a bundle is filled with row ranges: whole leaves, or segments of a
    // leaf too large for one bundle
    std::vector<std::pair<int, int>> cur;
    std::vector<int> rows, extCols;
    int curRows = 0, curLocal = 0, curLanes = 0;
    int bundleId = 0;  // id of the bundle being filled

    // local columns the rows [lo, hi) add to the bundle tagged `tag`
    auto countNew = [&](int lo, int hi, int tag) {
        int add = 0;
        for (int r = lo; r < hi; ++r)
            for (int j = rp[r]; j < rp[r + 1]; ++j) {
                const int c = col[j];
                if (gs && c >= lo && c < hi) continue;  // one of these rows
                if (seen[c] != tag) {
                    seen[c] = tag;
                    ++add;
                }
            }
        return add + (gs ? hi - lo : 0);
    };
    // drop the tags of rows that did not fit, so they can be recounted
    auto untag = [&](int lo, int hi, int tag) {
        for (int r = lo; r < hi; ++r)
            for (int j = rp[r]; j < rp[r + 1]; ++j)
                if (seen[col[j]] == tag) seen[col[j]] = -2;
    };

    auto flush = [&]() {
        if (cur.empty()) return;
        const int b = bundleId;
        const int s0 = t.slot.back();
        rows.clear();
        for (const auto& rg : cur)
            for (int r = rg.first; r < rg.second; ++r) rows.push_back(r);
        // slots ordered by (depth, row); one wave per depth present
        std::sort(rows.begin(), rows.end(), [&](int a_, int b_) {
            return depth[a_] != depth[b_] ? depth[a_] < depth[b_] : a_ < b_;
        });
        const std::vector<int>& ordered = rows;
        const int nb = (int)ordered.size();
        int nw = 0;
        for (int i = 0; i < nb; ++i)
            if (i == 0 || depth[ordered[i]] != depth[ordered[i - 1]]) {
                t.waveSlot.push_back(s0 + i);
                ++nw;
            }
        t.wave.push_back((int)t.waveSlot.size());
        // local column space (sorted outside columns: coalesced gathers)
        const int tagL = -(b + 3);  // distinct from the packing tags
        extCols.clear();
        if (gs) {
            for (int i = 0; i < nb; ++i) {
                owner[ordered[i]] = b;
                localOf[ordered[i]] = i;
            }
            for (int r : ordered)
                for (int j = rp[r] + 1; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if (owner[c] != b && seen[c] != tagL) {
                        seen[c] = tagL;
                        extCols.push_back(c);
                    }
                }
            std::sort(extCols.begin(), extCols.end());
            for (size_t e = 0; e < extCols.size(); ++e) localOf[extCols[e]] = nb + (int)e;
        } else {
            for (int r : ordered)
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if (seen[c] != tagL) {
                        seen[c] = tagL;
                        extCols.push_back(c);
                    }
                }
            std::sort(extCols.begin(), extCols.end());
            for (size_t e = 0; e < extCols.size(); ++e) localOf[extCols[e]] = (int)e;
        }
        t.ext.insert(t.ext.end(), extCols.begin(), extCols.end());
        t.extPtr.push_back((int)t.ext.size());
        const int local = (gs ? nb : 0) + (int)extCols.size();
        // entry block: lane l of slot i at (lane0(i) + l), column-major
        // over the bundle's lanes
        std::vector<int> lane0(nb + 1, 0);
        for (int i = 0; i < nb; ++i) lane0[i + 1] = lane0[i] + (var ? rowLanes(ordered[i]) : tpr);
        const long lanes = lane0[nb];
        int width = 0;
        for (int i = 0; i < nb; ++i) {
            const int cnt = (rp[ordered[i] + 1] - rp[ordered[i]]) - (gs ? 1 : 0);
            const int nl = lane0[i + 1] - lane0[i];
            width = std::max(width, (cnt + nl - 1) / nl);
        }
        t.len.push_back(width);
        t.maxLen = std::max(t.maxLen, width);
        const long off = (long)t.sidx.size();
        t.off.push_back(off);
        t.sidx.resize(off + (long)width * lanes);
        t.sval.resize(off + (long)width * lanes);
        if (var) {
            for (int i = 0; i < nb; ++i)
                for (int l = lane0[i]; l < lane0[i + 1]; ++l) t.laneSlot.push_back((uint16_t)i);
            t.lanePtr.push_back((int)t.laneSlot.size());
        }
        for (int i = 0; i < nb; ++i) {
            const int r = ordered[i];
            t.rowId.push_back(r);
            const int j0 = gs ? rp[r] + 1 : rp[r];
            const int cntRow = rp[r + 1] - j0;
            const int nl = lane0[i + 1] - lane0[i];
            double nrm = 0;
            for (int k = 0; k < width * nl; ++k) {
                const long pos = off + (long)(k / nl) * lanes + lane0[i] + (k % nl);
                if (k < cntRow) {
                    const int j = j0 + k;
                    t.sidx[pos] = (uint16_t)localOf[col[j]];
                    t.sval[pos] = val[j];
                    nrm += val[j] * val[j];
                } else {
                    // padding: zero value on local column 0 (never written through)
                    t.sidx[pos] = 0;
                    t.sval[pos] = 0.0;
                    ++t.paddedEntries;
                }
            }
            if (gs)
                t.diag.push_back(val[rp[r]]);
            else
                t.rowNorm.push_back(nrm);
        }
        t.slot.push_back(s0 + nb);
        t.phRows.back() = std::max(t.phRows.back(), nb);
        t.phLocal.back() = std::max(t.phLocal.back(), local);
        t.phWaves.back() = std::max(t.phWaves.back(), nw);
        t.phExt.back() = std::max(t.phExt.back(), (int)extCols.size());
        cur.clear();
        curRows = 0;
        curLocal = 0;
        curLanes = 0;
        ++bundleId;
    };
    // Every bundle starts a chain, except inside a large leaf, whose
    // bundles all belong to the chain started by its first one.
    bool inLeafChain = false, chainOpen = false;
    long chainWaves = 0;
    auto markChain = [&]() {
        // called after each flush: the bundle just written is slot.size()-2
        if (!chainOpen) t.chainPtr.push_back((int)t.slot.size() - 2);
        if (inLeafChain) {
            chainOpen = true;
            chainWaves += t.wave.back() - t.wave[t.wave.size() - 2];
        }
    };
    auto beginPhase = [&]() {
        t.phaseChain.push_back((int)t.chainPtr.size());
        t.phaseBundle.push_back((int)t.slot.size() - 1);
        t.phRows.push_back(0);
        t.phLocal.push_back(0);
        t.phWaves.push_back(0);
        t.phExt.push_back(0);
    };
    auto flushMark = [&]() {
        const bool had = !cur.empty();
        flush();
        if (had) markChain();
    };
    auto endPhase = [&]() {
        flushMark();
        t.critical += t.phWaves.back();
    };
    // add [lo, hi) to the bundle being filled, flushing it first when the
    // rows or local columns would exceed the limits; false if [lo, hi)
    // alone does not fit
    auto place = [&](int lo, int hi) {
        const int rowsL = hi - lo;
        if (rowsL > cap) return false;
        int lanesL = 0;
        if (var) {
            for (int r = lo; r < hi; ++r) lanesL += rowLanes(r);
            if (lanesL > laneCap) return false;
        }
        int add = countNew(lo, hi, bundleId);
        if (!cur.empty() && (curRows + rowsL > cap || curLocal + add > localCap ||
                             (var && curLanes + lanesL > laneCap))) {
            untag(lo, hi, bundleId);
            flushMark();
            add = countNew(lo, hi, bundleId);
        }
        if (add > localCap) {
            untag(lo, hi, bundleId);
            return false;
        }
        cur.push_back({lo, hi});
        curRows += rowsL;
        curLocal += add;
        curLanes += lanesL;
        return true;
    };

    std::vector<int> big;
    for (int p = 0; p < s.numPhases(); ++p) {
        beginPhase();
        big.clear();
        for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l)
            if (!place(s.leafLo[l], s.leafHi[l])) big.push_back(l);
        endPhase();
        if (big.empty()) continue;
        // Leaves too large for one bundle: their rows in (depth, row) order
        // (depth in the leaf's execution order, as for the waves), cut into
        // bundles as large as the limits allow.  A row of depth d depends
        // only on rows of smaller depth -- in an earlier bundle or an
        // earlier wave of its own -- so running the bundles in this order
        // reproduces the sequential leaf.  The bundles of a leaf form a
        // chain that one workgroup runs in order; the chains of all such
        // leaves of the phase run in one extra phase.
        beginPhase();
        long deepest = 0;
        std::vector<int> ord;
        for (size_t q = 0; q < big.size(); ++q) {
            const int lo = s.leafLo[big[q]], hi = s.leafHi[big[q]];
            ord.resize(hi - lo);
            for (int r = lo; r < hi; ++r) ord[r - lo] = r;
            std::sort(ord.begin(), ord.end(), [&](int a_, int b_) {
                return depth[a_] != depth[b_] ? depth[a_] < depth[b_] : a_ < b_;
            });
            inLeafChain = true;
            chainOpen = false;
            chainWaves = 0;
            for (int r : ord)
                if (!place(r, r + 1)) {
                    t.ok = false;  // a single row exceeds the local column space
                    return t;
                }
            flushMark();
            inLeafChain = false;
            chainOpen = false;
            deepest = std::max(deepest, chainWaves);
        }
        t.critical += deepest;  // a chain's waves run one after the other
    }
    t.phaseChain.push_back((int)t.chainPtr.size());
    t.chainPtr.push_back((int)t.slot.size() - 1);
    t.phaseBundle.push_back((int)t.slot.size() - 1);
    t.waveSlot.push_back(t.slot.back());
    return t;
}

BundleTable concatBundles(const std::vector<const BundleTable*>& tabs, std::vector<int>* bundleTable) {
    BundleTable t;
    bundleTable->clear();
    if (tabs.empty()) return t;
    t.kind = tabs[0]->kind;
    t.tpr = tabs[0]->tpr;
    t.cap = 0;
    int P = 0;
    for (const BundleTable* a : tabs) {
        P = std::max(P, a->numPhases());
        t.cap = std::max(t.cap, a->cap);
        t.maxLen = std::max(t.maxLen, a->maxLen);
        t.paddedEntries += a->paddedEntries;
        t.ok = t.ok && a->ok;
    }
    const bool gs = t.kind != RACE_K_KACZ;
    t.slot.push_back(0);
    t.wave.push_back(0);
    t.extPtr.push_back(0);
    if (t.tpr == 0) t.lanePtr.push_back(0);
    for (int p = 0; p < P; ++p) {
        t.phaseBundle.push_back((int)t.slot.size() - 1);
        t.phaseChain.push_back((int)t.chainPtr.size());
        int rows = 0, local = 0, waves = 0, extMax = 0;
        for (size_t m = 0; m < tabs.size(); ++m) {
            const BundleTable& a = *tabs[m];
            if (p >= a.numPhases()) continue;
            rows = std::max(rows, a.phRows[p]);
            extMax = std::max(extMax, a.phExt[p]);
            local = std::max(local, a.phLocal[p]);
            waves = std::max(waves, a.phWaves[p]);
            // chains keep their bundles, renumbered
            const int base = (int)t.slot.size() - 1;
            for (int c = a.phaseChain[p]; c < a.phaseChain[p + 1]; ++c)
                t.chainPtr.push_back(base + a.chainPtr[c] - a.phaseBundle[p]);
            for (int q = a.phaseBundle[p]; q < a.phaseBundle[p + 1]; ++q) {
                const int s0 = a.slot[q], nb = a.slot[q + 1] - s0;
                const int newS0 = t.slot.back();
                for (int w = a.wave[q]; w < a.wave[q + 1]; ++w) t.waveSlot.push_back(newS0 + a.waveSlot[w] - s0);
                t.wave.push_back((int)t.waveSlot.size());
                t.rowId.insert(t.rowId.end(), a.rowId.begin() + s0, a.rowId.begin() + s0 + nb);
                if (gs)
                    t.diag.insert(t.diag.end(), a.diag.begin() + s0, a.diag.begin() + s0 + nb);
                else
                    t.rowNorm.insert(t.rowNorm.end(), a.rowNorm.begin() + s0, a.rowNorm.begin() + s0 + nb);
                t.slot.push_back(newS0 + nb);
                const long lanesQ = a.tpr == 0 ? (long)(a.lanePtr[q + 1] - a.lanePtr[q]) : (long)nb * a.tpr;
                const long block = (long)a.len[q] * lanesQ;
                t.off.push_back((long)t.sidx.size());
                t.len.push_back(a.len[q]);
                if (a.tpr == 0) {
                    t.laneSlot.insert(t.laneSlot.end(), a.laneSlot.begin() + a.lanePtr[q],
                                      a.laneSlot.begin() + a.lanePtr[q + 1]);
                    t.lanePtr.push_back((int)t.laneSlot.size());
                }
                t.sidx.insert(t.sidx.end(), a.sidx.begin() + a.off[q], a.sidx.begin() + a.off[q] + block);
                t.sval.insert(t.sval.end(), a.sval.begin() + a.off[q], a.sval.begin() + a.off[q] + block);
                t.ext.insert(t.ext.end(), a.ext.begin() + a.extPtr[q], a.ext.begin() + a.extPtr[q + 1]);
                t.extPtr.push_back((int)t.ext.size());
                bundleTable->push_back((int)m);
            }
        }
        t.phRows.push_back(rows);
        t.phLocal.push_back(local);
        t.phWaves.push_back(waves);
        t.phExt.push_back(extMax);
        t.critical += waves;
    }
    t.phaseChain.push_back((int)t.chainPtr.size());
    t.chainPtr.push_back((int)t.slot.size() - 1);
    t.phaseBundle.push_back((int)t.slot.size() - 1);
    t.waveSlot.push_back(t.slot.back());
    return t;
}

}  // namespace race

namespace race {

namespace {
// key of the row / column id v of bundle b
inline long keyOf(const std::vector<long>* colBase, int b, int v) { return (colBase ? (*colBase)[b] : 0) + v; }
// chain of every bundle
std::vector<int> chainOfBundle(const BundleTable& t) {
    std::vector<int> cb(t.slot.size() > 0 ? t.slot.size() - 1 : 0, 0);
    for (int c = 0; c + 1 < (int)t.chainPtr.size(); ++c)
        for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b) cb[b] = c;
    return cb;
}
}  // namespace

long chainDeps(const BundleTable& t, long nkeys, const std::vector<long>* colBase, std::vector<int>* depPtr,
               std::vector<int>* dep) {
    const int nchain = (int)t.chainPtr.size() - 1;
    std::vector<int> phaseOf(std::max(nchain, 0), 0);
    for (int p = 0; p < t.numPhases(); ++p)
        for (int c = t.phaseChain[p]; c < t.phaseChain[p + 1]; ++c) phaseOf[c] = p;
    // (waiting chain, chain it waits for)
    std::vector<std::pair<int, int>> edge;
    long same = 0;
    auto add = [&](int waiter, int first) {
        if (phaseOf[waiter] == phaseOf[first]) ++same;
        edge.push_back({waiter, first});
    };
    if (t.kind != RACE_K_KACZ) {
        // owner[key]: chain that writes that row
        std::vector<int> owner(nkeys, -1);
        for (int c = 0; c < nchain; ++c)
            for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b)
                for (int s = t.slot[b]; s < t.slot[b + 1]; ++s) owner[keyOf(colBase, b, t.rowId[s])] = c;
        for (int c = 0; c < nchain; ++c)
            for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b)
                for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) {
                    const int o = owner[keyOf(colBase, b, t.ext[e])];
                    if (o < 0 || o == c) continue;  // not written in the sweep, or by this chain
                    if (o < c)
                        add(c, o);  // c reads what o wrote
                    else
                        add(o, c);  // o overwrites what c read
                }
    } else {
        // every column a chain touches is read and written: each chain waits
        // for the previous toucher of each of its columns
        std::vector<int> last(nkeys, -1);
        for (int c = 0; c < nchain; ++c) {
            for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b)
                for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) {
                    const int o = last[keyOf(colBase, b, t.ext[e])];
                    if (o >= 0 && o != c) add(c, o);
                }
            for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b)
                for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) last[keyOf(colBase, b, t.ext[e])] = c;
        }
    }
    std::sort(edge.begin(), edge.end());
    edge.erase(std::unique(edge.begin(), edge.end()), edge.end());
    depPtr->assign(nchain + 1, 0);
    dep->resize(edge.size());
    for (size_t k = 0; k < edge.size(); ++k) {
        ++(*depPtr)[edge[k].first + 1];
        (*dep)[k] = edge[k].second;
    }
    for (int c = 0; c < nchain; ++c) (*depPtr)[c + 1] += (*depPtr)[c];
    return same;
}

std::vector<int> flowExt(const BundleTable& t, long nkeys, const std::vector<long>* colBase) {
    const std::vector<int> cb = chainOfBundle(t);
    const int nb = (int)cb.size();
    std::vector<int> out(t.ext);
    auto mark = [&](int e) { out[e] = (int)((unsigned)t.ext[e] | 0x80000000u); };
    if (t.kind != RACE_K_KACZ) {
        std::vector<int> owner(nkeys, -1);
        for (int b = 0; b < nb; ++b)
            for (int s = t.slot[b]; s < t.slot[b + 1]; ++s) owner[keyOf(colBase, b, t.rowId[s])] = cb[b];
        for (int b = 0; b < nb; ++b)
            for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) {
                const int o = owner[keyOf(colBase, b, t.ext[e])];
                if (o >= 0 && o <= cb[b]) mark(e);
            }
    } else {
        // bundles in order: a column is fresh once any bundle before has touched it
        std::vector<char> touched(nkeys, 0);
        for (int b = 0; b < nb; ++b) {
            for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e)
                if (touched[keyOf(colBase, b, t.ext[e])]) mark(e);
            for (int e = t.extPtr[b]; e < t.extPtr[b + 1]; ++e) touched[keyOf(colBase, b, t.ext[e])] = 1;
        }
    }
    return out;
}

std::vector<long> chainLongestPaths(const BundleTable& t, const std::vector<int>& depPtr,
                                    const std::vector<int>& dep, int fixedPerBundle) {
    const int nc = (int)t.chainPtr.size() - 1;
    std::vector<long> path(std::max(nc, 0), 0);
    // successors of every chain, all later than it: a backward pass sees
    // them finished
    std::vector<std::vector<int>> succ(std::max(nc, 0));
    for (int c = 0; c < nc; ++c)
        for (int j = depPtr[c]; j < depPtr[c + 1]; ++j) succ[dep[j]].push_back(c);
    for (int c = nc - 1; c >= 0; --c) {
        long cost = 0;
        for (int b = t.chainPtr[c]; b < t.chainPtr[c + 1]; ++b) cost += t.wave[b + 1] - t.wave[b] + fixedPerBundle;
        long tail = 0;
        for (int s : succ[c]) tail = std::max(tail, path[s]);
        path[c] = cost + tail;
    }
    return path;
}

}  // namespace race
