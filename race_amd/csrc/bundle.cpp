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
                         const double* val, int cap, int innerCap, int tpr) {
    BundleTable t;
    t.kind = kind;
    t.cap = cap;
    t.tpr = tpr;
    // Kaczmarz holds every touched column in LDS: bound it by the entries
    if (kind == RACE_K_KACZ) innerCap = std::min(innerCap, 6144);
    std::vector<int> depth;
    leafDepths(s, kind, n, rp, col, depth);
    const bool gs = kind != RACE_K_KACZ;

    std::vector<int> bundleOf(n, -1), slotOf(n, -1);
    std::vector<int> colLocal(n, -1), colStamp(n, -1);
    std::vector<int> rows, cnt;
    t.slot.push_back(0);
    t.wave.push_back(0);
    t.innerPtr.reserve(n + 1);

    // staged entries of one leaf: GS -> couplings inside the leaf, KACZ -> all
    auto leafStaged = [&](int lo, int hi) {
        long c = 0;
        for (int r = lo; r < hi; ++r) {
            if (gs) {
                for (int j = rp[r] + 1; j < rp[r + 1]; ++j)
                    if (col[j] >= lo && col[j] < hi && col[j] != r) ++c;
            } else {
                c += rp[r + 1] - rp[r];
            }
        }
        return c;
    };
    auto leafWidth = [&](int lo, int hi) {
        int w = 0;
        for (int r = lo; r < hi; ++r) w = std::max(w, rp[r + 1] - rp[r]);
        return w;
    };
    int curWidth = 0;

    std::vector<int> bCols;  // columns held in LDS per bundle (Kaczmarz)
    std::vector<int> curLeaves;
    long curStaged = 0;
    int curRows = 0;

    auto flush = [&]() {
        if (curLeaves.empty()) return;
        const int b = (int)t.slot.size() - 1;
        const int s0 = t.slot.back();
        rows.clear();
        int maxD = 0;
        for (int l : curLeaves)
            for (int r = s.leafLo[l]; r < s.leafHi[l]; ++r) {
                rows.push_back(r);
                maxD = std::max(maxD, depth[r]);
            }
        // order by (depth, row)
        cnt.assign(maxD + 2, 0);
        for (int r : rows) cnt[depth[r] + 1]++;
        for (int d = 1; d <= maxD + 1; ++d) cnt[d] += cnt[d - 1];
        std::vector<int> ordered(rows.size());
        std::sort(rows.begin(), rows.end());
        for (int r : rows) ordered[cnt[depth[r]]++] = r;
        const int nb = (int)ordered.size();
        for (int i = 0; i < nb; ++i) {
            bundleOf[ordered[i]] = b;
            slotOf[ordered[i]] = i;
        }
        // wave boundaries (depths are contiguous 1..maxD)
        {
            int i = 0;
            for (int d = 1; d <= maxD; ++d) {
                t.waveSlot.push_back(s0 + i);
                while (i < nb && depth[ordered[i]] == d) ++i;
            }
        }
        t.wave.push_back((int)t.waveSlot.size());
        t.maxWaves = std::max(t.maxWaves, maxD);
        // KACZ: the set of columns touched by the bundle, held in LDS
        int ncols = 0;
        if (!gs) {
            t.colSetPtr.push_back((int)t.colSet.size());
            for (int r : ordered)
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if (colStamp[c] != b) {
                        colStamp[c] = b;
                        colLocal[c] = ncols++;
                        t.colSet.push_back(c);
                    }
                }
            t.maxCols = std::max(t.maxCols, ncols);
        }
        bCols.push_back(ncols);
        // per-lane register width: off-diagonal entries (GS) or all entries (KACZ)
        int rowMax = 0;
        for (int r : ordered) rowMax = std::max(rowMax, (rp[r + 1] - rp[r]) - (gs ? 1 : 0));
        const int width = (rowMax + tpr - 1) / tpr;
        t.len.push_back(width);
        t.maxLen = std::max(t.maxLen, width);
        const long off = (long)t.scol.size();
        const long lanes = (long)nb * tpr;
        t.off.push_back(off);
        t.scol.resize(off + (long)width * lanes);
        t.sval.resize(off + (long)width * lanes);
        int staged = 0;
        for (int i = 0; i < nb; ++i) {
            const int r = ordered[i];
            t.rowId.push_back(r);
            t.innerPtr.push_back(staged);
            const int j0 = gs ? rp[r] + 1 : rp[r];
            const int cntRow = rp[r + 1] - j0;
            double nrm = 0;
            for (int k = 0; k < width * tpr; ++k) {
                const long pos = off + (long)(k / tpr) * lanes + (long)i * tpr + (k % tpr);
                if (k < cntRow) {
                    const int j = j0 + k;
                    const int c = col[j];
                    int enc;
                    if (gs) {
                        enc = (bundleOf[c] == b) ? ~slotOf[c] : c;
                        if (enc < 0) ++staged;
                    } else {
                        enc = ~colLocal[c];
                        ++staged;
                        nrm += val[j] * val[j];
                    }
                    t.scol[pos] = enc;
                    t.sval[pos] = val[j];
                } else {
                    // padding: zero value; GS reads global column 0, KACZ local
                    // column 0 (the kernel never writes through a zero value)
                    t.scol[pos] = gs ? 0 : ~0;
                    t.sval[pos] = 0.0;
                    ++t.paddedEntries;
                }
            }
            if (gs)
                t.diag.push_back(val[rp[r]]);
            else
                t.rowNorm.push_back(nrm);
        }
        t.innerEnd.push_back(staged);
        t.maxInner = std::max(t.maxInner, staged);
        t.maxRows = std::max(t.maxRows, nb);
        t.slot.push_back(s0 + nb);
        curLeaves.clear();
        curRows = 0;
        curStaged = 0;
    };

    for (int p = 0; p < s.numPhases(); ++p) {
        t.phaseBundle.push_back((int)t.slot.size() - 1);
        for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l) {
            const int rowsL = s.leafHi[l] - s.leafLo[l];
            const long st = leafStaged(s.leafLo[l], s.leafHi[l]);
            if (rowsL > cap || st > innerCap) {
                t.ok = false;
                return t;
            }
            // Kaczmarz stages the padded SELL block: rows x widest row
            const int wL = gs ? 0 : leafWidth(s.leafLo[l], s.leafHi[l]);
            if (!gs && (long)rowsL * wL > innerCap) {
                t.ok = false;
                return t;
            }
            if (curRows + rowsL > cap || curStaged + st > innerCap ||
                (!gs && (long)(curRows + rowsL) * std::max(curWidth, wL) > innerCap)) {
                flush();
                curWidth = 0;
            }
            curLeaves.push_back(l);
            curRows += rowsL;
            curStaged += st;
            curWidth = std::max(curWidth, wL);
        }
        flush();
        // depth of the phase = deepest bundle; LDS needs of the phase
        int deepest = 0, mr = 0, mi = 0, mc = 0, me = 0;
        for (int b = t.phaseBundle.back(); b + 1 < (int)t.slot.size(); ++b) {
            const int rows_b = t.slot[b + 1] - t.slot[b];
            deepest = std::max(deepest, t.wave[b + 1] - t.wave[b]);
            mr = std::max(mr, rows_b);
            mi = std::max(mi, t.innerEnd[b]);
            me = std::max(me, t.len[b] * rows_b * tpr);
            mc = std::max(mc, bCols[b]);
        }
        t.critical += deepest;
        t.phRows.push_back(mr);
        t.phInner.push_back(mi);
        t.phWaves.push_back(deepest);
        t.phEntries.push_back(me);
        t.phCols.push_back(mc);
    }
    t.phaseBundle.push_back((int)t.slot.size() - 1);
    t.waveSlot.push_back(t.slot.back());
    if (!gs) t.colSetPtr.push_back((int)t.colSet.size());
    return t;
}

}  // namespace race
