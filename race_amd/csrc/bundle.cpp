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
                         const double* val, int cap, int localCap, int tpr) {
    BundleTable t;
    t.kind = kind;
    t.cap = cap;
    t.tpr = tpr;
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

    std::vector<int> curLeaves, rows, cnt, extCols;
    int curRows = 0, curLocal = 0;
    int bundleId = 0;  // id of the bundle being filled

    // local columns a leaf adds to the bundle tagged `tag`
    auto countNew = [&](int l, int tag) {
        int add = 0;
        const int lo = s.leafLo[l], hi = s.leafHi[l];
        for (int r = lo; r < hi; ++r)
            for (int j = rp[r]; j < rp[r + 1]; ++j) {
                const int c = col[j];
                if (gs && c >= lo && c < hi) continue;  // a row of this leaf
                if (seen[c] != tag) {
                    seen[c] = tag;
                    ++add;
                }
            }
        return add + (gs ? hi - lo : 0);
    };
    // drop the tags of a leaf that did not fit, so it can be recounted
    auto untag = [&](int l, int tag) {
        for (int r = s.leafLo[l]; r < s.leafHi[l]; ++r)
            for (int j = rp[r]; j < rp[r + 1]; ++j)
                if (seen[col[j]] == tag) seen[col[j]] = -2;
    };

    auto flush = [&]() {
        if (curLeaves.empty()) return;
        const int b = bundleId;
        const int s0 = t.slot.back();
        rows.clear();
        int maxD = 0;
        for (int l : curLeaves)
            for (int r = s.leafLo[l]; r < s.leafHi[l]; ++r) {
                rows.push_back(r);
                maxD = std::max(maxD, depth[r]);
            }
        // slots ordered by (depth, row)
        std::sort(rows.begin(), rows.end());
        cnt.assign(maxD + 2, 0);
        for (int r : rows) cnt[depth[r] + 1]++;
        for (int d = 1; d <= maxD + 1; ++d) cnt[d] += cnt[d - 1];
        std::vector<int> ordered(rows.size());
        for (int r : rows) ordered[cnt[depth[r]]++] = r;
        const int nb = (int)ordered.size();
        {
            int i = 0;
            for (int d = 1; d <= maxD; ++d) {
                t.waveSlot.push_back(s0 + i);
                while (i < nb && depth[ordered[i]] == d) ++i;
            }
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
        // entry block
        int rowMax = 0;
        for (int r : ordered) rowMax = std::max(rowMax, (rp[r + 1] - rp[r]) - (gs ? 1 : 0));
        const int width = (rowMax + tpr - 1) / tpr;
        t.len.push_back(width);
        t.maxLen = std::max(t.maxLen, width);
        const long off = (long)t.sidx.size();
        const long lanes = (long)nb * tpr;
        t.off.push_back(off);
        t.sidx.resize(off + (long)width * lanes);
        t.sval.resize(off + (long)width * lanes);
        for (int i = 0; i < nb; ++i) {
            const int r = ordered[i];
            t.rowId.push_back(r);
            const int j0 = gs ? rp[r] + 1 : rp[r];
            const int cntRow = rp[r + 1] - j0;
            double nrm = 0;
            for (int k = 0; k < width * tpr; ++k) {
                const long pos = off + (long)(k / tpr) * lanes + (long)i * tpr + (k % tpr);
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
        t.phWaves.back() = std::max(t.phWaves.back(), maxD);
        curLeaves.clear();
        curRows = 0;
        curLocal = 0;
        ++bundleId;
    };

    for (int p = 0; p < s.numPhases(); ++p) {
        t.phaseBundle.push_back((int)t.slot.size() - 1);
        t.phRows.push_back(0);
        t.phLocal.push_back(0);
        t.phWaves.push_back(0);
        for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l) {
            const int rowsL = s.leafHi[l] - s.leafLo[l];
            if (rowsL > cap) {
                t.ok = false;
                return t;
            }
            int add = countNew(l, bundleId);
            if (!curLeaves.empty() && (curRows + rowsL > cap || curLocal + add > localCap)) {
                untag(l, bundleId);
                flush();
                add = countNew(l, bundleId);
            }
            if (add > localCap) {
                t.ok = false;
                return t;
            }
            curLeaves.push_back(l);
            curRows += rowsL;
            curLocal += add;
        }
        flush();
        t.critical += t.phWaves.back();
    }
    t.phaseBundle.push_back((int)t.slot.size() - 1);
    t.waveSlot.push_back(t.slot.back());
    return t;
}

}  // namespace race
