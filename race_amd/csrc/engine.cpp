// engine.cpp -- host scheduling logic of the MI355X RACE engine.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>

namespace race {

std::vector<int> envIntList(const char* name) {
    std::vector<int> out;
    const char* v = std::getenv(name);
    if (!v) return out;
    std::string s(v);
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t e = s.find(',', pos);
        if (e == std::string::npos) e = s.size();
        std::string tok = s.substr(pos, e - pos);
        if (!tok.empty()) out.push_back(std::atoi(tok.c_str()));
        pos = e + 1;
    }
    return out;
}

// indices 0..len-1 stably ordered by descending key
static std::vector<int> rankDescending(const std::vector<double>& key) {
    std::vector<int> idx(key.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] > key[b]; });
    return idx;
}

// Per-colour mean and variance of the scaled chunk sums
// (Stat::calculate src/lb.cpp:773-795).
struct ZoneStat {
    std::vector<double> mean, var, acquire, give, weight;
    ZoneStat(const std::vector<long>& chunk, int parts, const std::vector<int>& scale) {
        int len = (int)chunk.size();
        mean.assign(parts, 0.0);
        var.assign(parts, 0.0);
        acquire.assign(len, 0.0);
        give.assign(len, 0.0);
        weight.assign(len, 0.0);
        for (int p = 0; p < parts; ++p) {
            double m = 0.0;
            int t = 0;
            for (int i = p; i < len; i += parts, ++t) m = m + chunk[i] / (double)scale[t];
            m = (parts * m) / len;
            mean[p] = m;
            double v = 0.0;
            t = 0;
            for (int i = p; i < len; i += parts, ++t) {
                double d = (chunk[i] / (double)scale[t]) - m;
                v = v + d * d;
                acquire[i] = d;
                give[i] = -d;
                weight[i] = std::fabs(d);
            }
            var[p] = v;
        }
    }
    double totalVar() const {
        double s = 0;
        for (double v : var) s += v;
        return s;
    }
};

LBResult splitZones(const LevelData& ld, int nThreads, double efficiency, int dist, int d2,
                    int lbTarget) {
    const int L = ld.totalLevel();
    const std::vector<int>& target = (lbTarget == LB_ROW) ? ld.levelRow : ld.levelNnz;
    long targetTotal = 0;
    for (int v : target) targetTotal += v;
    // LevelData::nnz/nrow are int sums in the reference
    int targetTotalI = (int)targetTotal;

    std::vector<double> effRatio(L);
    for (int i = 0; i < L; ++i) {
        // int product, wrapping like the reference's int arithmetic
        int32_t prod = (int32_t)((uint32_t)target[i] * (uint32_t)nThreads);
        effRatio[i] = prod / (double)targetTotalI;
    }

    const int gap = minGap(dist, d2);
    const int bpt = blockPerThread(dist, d2);
    const int gapPerThread = gap * bpt;

    std::vector<int> scale;
    std::vector<int> firstLevel{0};
    double acc = 0.0, carried = 0.0;
    int levelsSinceCut = 0;
    for (int i = 0; i < L; ++i) {
        acc += effRatio[i];
        ++levelsSinceCut;
        if (levelsSinceCut >= gapPerThread && roundDouble(acc) != 0 &&
            (1 - std::fabs(acc - roundDouble(acc))) * 100 > efficiency) {
            // accepted; look ahead for a cut closer to the rounded value
            const double rem0 = acc - roundDouble(acc);
            const int rounded = roundDouble(acc);
            double remNew = rem0, remBest = rem0, accProbe = acc, accBest = acc;
            int best = i;
            for (int j = i + 1; j < L && std::fabs(remNew) <= std::fabs(rem0); ++j) {
                accProbe += effRatio[j];
                remNew = accProbe - rounded;
                if (std::fabs(remBest) > std::fabs(remNew)) {
                    remBest = remNew;
                    accBest = accProbe;
                    best = j;
                } else if ((std::fabs(remBest) - std::fabs(remNew)) < 1e-6) {
                    if (remBest + carried > remNew + carried) {
                        remBest = remNew;
                        accBest = accProbe;
                        best = j;
                    }
                }
            }
            levelsSinceCut += best - i;
            i = best;
            acc = accBest;
            carried = carried + (acc - roundDouble(acc));
            scale.push_back(roundDouble(acc));
            firstLevel.push_back(i + 1);
            acc = 0;
            levelsSinceCut = 0;
        }
        if (i >= L - 5) {
            // the tail goes to the last thread
            double tail = acc;
            for (int j = i + 1; j < L; ++j) tail += effRatio[j];
            if (roundDouble(tail) != 0) {
                scale.push_back(roundDouble(tail));
                firstLevel.push_back(L);
            } else {
                firstLevel.back() = L;
            }
            break;
        }
    }

    int threadCount = 0;
    for (int& s : scale) {
        if (s == 0) s = 1;
        threadCount += s;
    }
    int threads = 0;
    if (!scale.empty()) {
        int diff = nThreads - threadCount;
        if (diff == 1) {
            scale.back() += 1;
        } else if (diff == -1) {
            if (scale.back() == 1) {
                scale.pop_back();
                firstLevel.pop_back();
                firstLevel.back() = L;
            } else {
                scale.back() -= 1;
            }
        }
        threads = (int)scale.size();
    }

    const int blocks = bpt * threads;
    std::vector<int> lp(blocks + 1, 0);
    for (int t = 0; t < threads; ++t)
        for (int j = 1; j <= bpt; ++j)
            lp[t * bpt + j] =
                firstLevel[t] + roundDouble(j * (firstLevel[t + 1] - firstLevel[t]) / (double)bpt);

    auto chunkSums = [&](const std::vector<int>& ptr) {
        std::vector<long> cs(blocks, 0);
        for (int b = 0; b < blocks; ++b)
            for (int l = ptr[b]; l < ptr[b + 1]; ++l) cs[b] += target[l];
        return cs;
    };

    if (!scale.empty()) {
        bool done = false;
        while (!done) {
            std::vector<long> cs = chunkSums(lp);
            ZoneStat st(cs, bpt, scale);
            const double var = st.totalVar();
            double newVar = var;
            const std::vector<int> rank = rankDescending(st.weight);
            const std::vector<int> acquireRank = rankDescending(st.acquire);
            const std::vector<int> giveRank = rankDescending(st.give);
            const std::vector<int> saved = lp;
            int r = 0;
            while (newVar >= var) {
                const int me = rank[r];
                lp = saved;
                const double myMean = st.mean[me % bpt];
                bool fail = false, movable = true, first = true;
                while (first || (movable && std::fabs(var - newVar) < 1e-5)) {
                    first = false;
                    int from = -1, to = -1;
                    bool doMove = false;
                    if ((cs[me] / (double)scale[me / bpt]) < myMean) {
                        // the most loaded block that can spare a level
                        int donor = -1;
                        for (int cand : acquireRank)
                            if (lp[cand + 1] - lp[cand] > gap) {
                                donor = cand;
                                break;
                            }
                        if (donor == -1 || donor == me) {
                            fail = true;
                            movable = false;
                        }
                        to = me;
                        from = donor;
                        doMove = true;
                    } else if (lp[me + 1] - lp[me] > gap) {
                        const int taker = giveRank[0];
                        if (taker == -1 || taker == me) {
                            fail = true;
                            movable = false;
                        }
                        to = taker;
                        from = me;
                        doMove = true;
                    } else {
                        movable = false;
                    }
                    if (doMove) {
                        // shift every boundary between the two blocks by one level
                        if (to < from) {
                            for (int k = to + 1; k <= from; ++k) lp[k] += 1;
                        } else {
                            for (int k = to; k >= from + 1; --k) lp[k] -= 1;
                        }
                    }
                    if (!fail) newVar = ZoneStat(chunkSums(lp), bpt, scale).totalVar();
                }
                if (r == blocks - 1 && newVar >= var) {
                    done = true;
                    lp = saved;
                    break;
                }
                ++r;
            }
        }
    }

    LBResult res;
    res.levelPtr = lp;
    res.scale = scale;
    res.zoneRows.assign(blocks, 0);
    for (int b = 0; b < blocks; ++b)
        for (int l = lp[b]; l < lp[b + 1]; ++l) res.zoneRows[b] += ld.levelRow[l];

    // cache sub-blocking: the reference sets its cache size to INT_MAX, so a
    // zone is only split when it exceeds INT_MAX/(16+27*12) rows
    int blocked = std::numeric_limits<int>::max() / (8 * 2 + 27 * 12);
    int maxLevelRows = 0;
    for (int v : ld.levelRow) maxLevelRows = std::max(maxLevelRows, v);
    blocked = std::max(blocked, maxLevelRows);
    std::vector<int> subLevel;
    res.numBlocks.assign(blocks, 0);
    for (int b = 0; b < blocks; ++b) {
        res.numBlocks[b] += 1;
        subLevel.push_back(lp[b]);
        int t = (int)(b / (double)bpt);
        if (res.zoneRows[b] > blocked && scale[t] > 1 && (lp[b + 1] - lp[b]) > 1) {
            int cur = ld.levelRow[lp[b]];
            for (int j = lp[b] + 1; j < lp[b + 1]; ++j) {
                cur += ld.levelRow[j];
                if (cur > blocked) {
                    res.numBlocks[b] += 1;
                    subLevel.push_back(j);
                    cur = ld.levelRow[j];
                }
            }
        }
    }
    subLevel.push_back(lp[blocks]);
    for (size_t s = 0; s + 1 < subLevel.size(); ++s) {
        int rows = 0;
        for (int l = subLevel[s]; l < subLevel[s + 1]; ++l) rows += ld.levelRow[l];
        res.subRows.push_back(rows);
    }
    return res;
}

// ---------------------------------------------------------------------------
ZoneTree::ZoneTree(int dist_, int d2_, int lbTarget_)
    : dist(dist_), d2(d2_), lbTarget(lbTarget_), bpt(blockPerThread(dist_, d2_)) {}

int ZoneTree::maxStage() const {
    int m = 0;
    for (const auto& l : tree) m = std::max(m, l.stage);
    return m;
}

int ZoneTree::keyChildEffRow(int p) const {
    const std::vector<int>& ch = tree[p].children;
    std::vector<int> maxRow(bpt, 0);
    for (size_t sb = 0; sb < ch.size() / 2; ++sb) {
        for (int b = 0; b < bpt; ++b) {
            int m = 0;
            for (int i = ch[2 * sb] + b; i < ch[2 * sb + 1]; i += bpt) m = std::max(m, tree[i].effRowZ);
            maxRow[b] += m;
        }
    }
    int s = 0;
    for (int v : maxRow) s += v;
    return s;
}

void ZoneTree::updateEffRow(int p) {
    for (int q = p; q != -1; q = tree[q].parentZ) tree[q].effRowZ = keyChildEffRow(q);
}

void ZoneTree::updateNThreads(int p) {
    for (int q = p; q != -1; q = tree[q].parentZ) {
        const std::vector<int>& ch = tree[q].children;
        int best = 0;
        for (size_t sb = 0; sb < ch.size() / 2; ++sb) {
            int nt = (ch[2 * sb + 1] - ch[2 * sb]) / bpt;
            int sum = 0;
            for (int i = 0; i < nt; ++i) {
                int m = 0;
                for (int b = 0; b < bpt; ++b) m = std::max(m, tree[ch[2 * sb] + bpt * i + b].nthreadsZ);
                sum += m;
            }
            best = std::max(best, sum);
        }
        tree[q].nthreadsZ = best;
    }
}

bool ZoneTree::spawnChild(int parent, int parentSub, int request, const LevelData& ld, double eff) {
    LBResult lb = splitZones(ld, request, eff, dist, d2, lbTarget);
    const int first = (int)tree.size();
    const int base = tree[parent].valueZ[parentSub];
    const int stage = tree[parent].stage;
    const int blocks = (int)lb.zoneRows.size();
    tree[parent].children.push_back(first);
    int row = base, subRow = base, sub = 0;
    for (int b = 0; b < blocks; ++b) {
        ZoneLeaf leaf;
        leaf.parentZ = parent;
        leaf.effRowZ = lb.zoneRows[b];
        leaf.totalSubBlocks = lb.numBlocks[b];
        leaf.stage = stage + 1;
        leaf.valueZ.clear();
        for (int s = 0; s < lb.numBlocks[b]; ++s) {
            leaf.valueZ.push_back(subRow);
            subRow += lb.subRows[sub++];
        }
        leaf.valueZ.push_back(subRow);
        row += lb.zoneRows[b];
        tree.push_back(leaf);
    }
    (void)row;
    tree[parent].children.push_back(first + blocks);
    if (parentSub == tree[parent].totalSubBlocks - 1) {
        updateEffRow(parent);
        updateNThreads(parent);
    }
    return lb.threads() == request;
}

// ---------------------------------------------------------------------------
LevelRecursion::LevelRecursion(Traverser* t, int n_, int request_, int dist_, int d2_,
                               int lbTarget_, const std::vector<double>& effList,
                               const std::vector<int>& lvlThreads_)
    : trav(t), n(n_), request(request_), dist(dist_), d2(d2_), lbTarget(lbTarget_),
      lvlThreadsVec(lvlThreads_) {
    effVec.push_back(40.0);  // default efficiency (src/level_recursion.cpp:48)
    for (double e : effList) effVec.push_back(e);
    // an explicit per-stage thread count drops that stage's efficiency to 0
    for (size_t lvl = 1; lvl <= lvlThreadsVec.size(); ++lvl) {
        while (effVec.size() <= lvl) effVec.push_back(effVec[0]);
        effVec[lvl] = 0;
    }
}

double LevelRecursion::efficiency(int lvl) const {
    return (lvl >= 0 && lvl < (int)effVec.size()) ? effVec[lvl] : effVec[0];
}

int LevelRecursion::lvlThreads(int lvl) const {
    return (lvl - 1 >= 0 && lvl - 1 < (int)lvlThreadsVec.size()) ? lvlThreadsVec[lvl - 1] : -1;
}

long LevelRecursion::nnzSum(int lo, int hi) const {
    const std::vector<int>& p = trav->perm();
    long s = 0;
    for (int r = lo; r < hi; ++r) s += trav->rowLen(p[r]);
    return s;
}

void LevelRecursion::idealNthreads(int parent, int parentSub, int currLvl) {
    auto& T = zt->tree;
    const int bpt = zt->bpt;
    const int c0 = T[parent].children[2 * parentSub];
    const int nt = (T[parent].children[2 * parentSub + 1] - c0) / bpt;
    if (lvlThreads(currLvl) != -1) {
        int my = lvlThreads(currLvl);
        for (int i = 0; i < nt; ++i)
            for (int b = 0; b < bpt; ++b) T[c0 + bpt * i + b].idealNthreadsZ = my;
        return;
    }
    const int parentThreads = T[parent].idealNthreadsZ;
    const int s = T[parent].valueZ[parentSub], e = T[parent].valueZ[parentSub + 1];
    // the reference accumulates nnz in int
    const int total = (lbTarget == LB_NNZ) ? (int)nnzSum(s, e) : e - s;
    std::vector<int> weights(nt);
    for (int i = 0; i < nt; ++i) {
        const int ss = T[c0 + i * bpt].valueZ.front();
        const int ee = T[c0 + bpt * (i + 1) - 1].valueZ.back();
        weights[i] = (lbTarget == LB_NNZ) ? (int)nnzSum(ss, ee) : ee - ss;
    }
    const std::vector<int> my = splitThreads(weights, total, parentThreads);
    for (int i = 0; i < nt; ++i)
        for (int b = 0; b < bpt; ++b) T[c0 + bpt * i + b].idealNthreadsZ = my[i];
}

// The thread split of LevelRecursion::calculateIdealNthreads
// (src/level_recursion.cpp:165-245): group i gets floor(weight_i / total *
// parentThreads); groups left with 0 rank first (remainder 1), the others by
// their fractional part, and the leftover threads go one each down that
// ranking (stable: ties keep group order).
std::vector<int> splitThreads(const std::vector<int>& weights, int total, int parentThreads) {
    const int nt = (int)weights.size();
    std::vector<int> my(nt);
    std::vector<double> frac(nt);
    int remaining = parentThreads;
    for (int i = 0; i < nt; ++i) {
        const double myD = (weights[i] / (double)total) * parentThreads;
        my[i] = (int)myD;
        frac[i] = (my[i] != 0) ? (myD - my[i]) : 1.0;
        remaining -= my[i];
    }
    if (remaining > 0) {
        std::vector<int> order = rankDescending(frac);
        for (int r = 0; remaining != 0; ++r, --remaining) my[order[r]] += 1;
    }
    return my;
}

int LevelRecursion::recursivePartition(int parent, int parentSub, int currLvl) {
    idealNthreads(parent, parentSub, currLvl);
    const std::vector<int> ch = zt->tree[parent].children;
    // The ranges traversed below are fixed here (valueZ and idealNthreadsZ
    // of these children change no more) and disjoint, and a traversal
    // depends only on the rows of its range, so all of them are traversed
    // up front as one batch; the loop then consumes the results in order.
    std::vector<std::pair<int, int>> ranges;
    for (int ci = ch[2 * parentSub]; ci < ch[2 * parentSub + 1]; ++ci) {
        if (zt->tree[ci].idealNthreadsZ <= 1) continue;
        const std::vector<int>& rng = zt->tree[ci].valueZ;
        for (size_t j = 0; j + 1 < rng.size(); ++j)
            if (rng[j + 1] > rng[j]) ranges.push_back({rng[j], rng[j + 1]});
    }
    std::vector<LevelData> lds;
    int berr = trav->traverseBatch(dist, ranges, &lds);
    if (berr != OK) return berr;
    size_t next = 0;
    for (int ci = ch[2 * parentSub]; ci < ch[2 * parentSub + 1]; ++ci) {
        const int nt = zt->tree[ci].idealNthreadsZ;
        const std::vector<int> rng = zt->tree[ci].valueZ;
        if (nt <= 1) continue;
        for (size_t j = 0; j + 1 < rng.size(); ++j) {
            if (rng[j + 1] > rng[j]) {
                const LevelData& ld = lds[next++];
                int err;
                bool ok = zt->spawnChild(ci, (int)j, nt, ld, efficiency(currLvl));
                if (!ok) {
                    const std::vector<int>& cc = zt->tree[ci].children;
                    int sub = (cc[2 * j + 1] - cc[2 * j]) / zt->bpt;
                    if (sub == 1) zt->tree[ci].reachedLimit = true;
                    if (sub > 1) {
                        err = recursivePartition(ci, (int)j, currLvl + 1);
                        if (err != OK) return err;
                    }
                }
            } else {
                zt->tree[ci].reachedLimit = true;
            }
        }
    }
    return OK;
}

int LevelRecursion::levelBalancing(ZoneTree* tree) {
    zt = tree;
    ZoneLeaf root;
    root.valueZ = {0, n};
    root.parentZ = -1;
    root.effRowZ = n;
    root.idealNthreadsZ = request;
    root.totalSubBlocks = 1;
    zt->tree.clear();
    zt->tree.push_back(root);
    LevelData ld;
    int err = trav->traverse(dist, 0, n, &ld);
    if (err != OK) return err;
    int stage1 = lvlThreads(1) != -1 ? lvlThreads(1) : request;
    zt->spawnChild(0, 0, stage1, ld, efficiency(1));
    if (zt->tree[0].nthreadsZ != request) {
        err = recursivePartition(0, 0, 2);
        if (err != OK) return err;
    }
    availableThreads = zt->tree[0].nthreadsZ;
    return OK;
}

// ---------------------------------------------------------------------------
// Phases of the colouring executor.  A node whose sub-block has at most one
// thread is executed as one sequential call over valueZ[0..1) (the reference
// uses the first two boundaries for every sub-block of such a node); a node
// with more threads runs its children colour by colour, all threads of a
// colour concurrently, with a barrier after each colour.  Concurrent children
// are merged phase by phase, which only adds barriers.
static void flattenNode(const ZoneTree& zt, int p, bool rev,
                        std::vector<std::vector<std::pair<int, int>>>& out) {
    const ZoneLeaf& leaf = zt.tree[p];
    const std::vector<int>& ch = leaf.children;
    const int bpt = zt.bpt;
    for (int sb = 0; sb < leaf.totalSubBlocks; ++sb) {
        int nt = ((int)ch.size() < 2 * sb + 2) ? 0 : (ch[2 * sb + 1] - ch[2 * sb]) / bpt;
        if (nt <= 1) {
            out.push_back({{leaf.valueZ[0], leaf.valueZ[1]}});
            continue;
        }
        for (int k = 0; k < bpt; ++k) {
            const int b = rev ? bpt - 1 - k : k;
            std::vector<std::vector<std::pair<int, int>>> merged;
            for (int tid = 0; tid < nt; ++tid) {
                std::vector<std::vector<std::pair<int, int>>> sub;
                flattenNode(zt, ch[2 * sb] + bpt * tid + b, rev, sub);
                if (merged.size() < sub.size()) merged.resize(sub.size());
                for (size_t q = 0; q < sub.size(); ++q)
                    merged[q].insert(merged[q].end(), sub[q].begin(), sub[q].end());
            }
            for (auto& m : merged) out.push_back(std::move(m));
        }
    }
}

Schedule flattenSchedule(const ZoneTree& zt, bool rev, int nrow) {
    std::vector<std::vector<std::pair<int, int>>> ph;
    if (zt.tree.empty()) {
        ph.push_back({{0, nrow}});
    } else {
        flattenNode(zt, 0, rev, ph);
    }
    Schedule s;
    s.rev = rev;
    s.phasePtr.push_back(0);
    for (auto& p : ph) {
        for (auto& lh : p) {
            if (lh.second <= lh.first) continue;  // empty leaves do nothing
            s.leafLo.push_back(lh.first);
            s.leafHi.push_back(lh.second);
        }
        if ((int)s.leafLo.size() > s.phasePtr.back()) s.phasePtr.push_back((int)s.leafLo.size());
    }
    return s;
}

// ---------------------------------------------------------------------------
namespace {
// getHopelessStartEnd src/matrixPower.cpp:674-683
void regionAt(const std::vector<int>& r, int i, int* s, int* e) {
    *s = -1;
    *e = -1;
    if (i >= 0 && 2 * i < (int)r.size() - 1) {
        *s = r[2 * i];
        *e = r[2 * i + 1];
    }
}
}  // namespace

std::vector<int> identifyHopelessRegions(const std::vector<int>& violated, int k, int totalLevel) {
    // runs of consecutive violated levels (a single violated level is no
    // region): flag each run, then read off its first and last level
    std::vector<int> cvl = violated;
    if (cvl.size() > 1) cvl.push_back(cvl.back() + 2);
    std::vector<int> flag(cvl.size(), 0);
    int prev = -2, block = 1;
    for (size_t i = 0; i < cvl.size(); ++i) {
        if (cvl[i] == prev + 1) {
            flag[i - 1] = block;
            flag[i] = block;
        } else {
            ++block;
        }
        prev = cvl[i];
    }
    std::vector<int> regs;
    for (size_t i = 0, pf = 0; i < flag.size(); pf = flag[i], ++i) {
        const int f = flag[i];
        if (f == (int)pf) continue;
        if (f > 0 && pf == 0) {
            regs.push_back(cvl[i]);
        } else if (f > 0) {
            regs.push_back(cvl[i - 1]);
            regs.push_back(cvl[i]);
        } else {
            regs.push_back(cvl[i - 1]);
        }
    }
    // regions closer than k-1 levels become one
    const int nh = (int)regs.size() / 2;
    std::vector<int> merged;
    int cs, ce, ns, ne;
    regionAt(regs, 0, &cs, &ce);
    regionAt(regs, 1, &ns, &ne);
    for (int i = 0; i < nh; ++i) {
        if (cs == -1) continue;
        while ((ns - ce - 1) < (k - 1) && ns != -1) {
            ce = ne;
            regionAt(regs, i + 2, &ns, &ne);
            ++i;
        }
        merged.push_back(cs);
        merged.push_back(ce);
        regionAt(regs, i + 1, &cs, &ce);
        regionAt(regs, i + 2, &ns, &ne);
    }
    // one node [0, totalLevel): keep k-1 levels to its first and last level
    const int lvlStart = 0, lvlEnd = totalLevel;
    std::vector<int> split;
    for (int i = 0; i < (int)merged.size() / 2; ++i) {
        regionAt(merged, i, &cs, &ce);
        if (cs >= lvlStart) {
            if (cs < lvlEnd) {
                split.push_back(cs);
                split.push_back(ce > lvlEnd ? lvlEnd - 1 : ce);
            }
        } else if (ce > lvlEnd) {
            split.push_back(lvlStart);
            split.push_back(lvlEnd - 1);
        } else if (ce >= lvlStart) {
            split.push_back(lvlStart);
            split.push_back(ce);
        }
    }
    std::vector<int> drop;
    regionAt(split, 0, &cs, &ce);
    if (cs != -1 && (cs - lvlStart - 1) < (k - 1)) {
        const int newStart = lvlStart + (k - 1);
        if (ce < newStart)
            drop.push_back(0);
        else
            split[0] = newStart;
    }
    const int last = (int)split.size() / 2 - 1;
    regionAt(split, last, &cs, &ce);
    if (ce != -1 && ((lvlEnd - 1) - ce - 1) < (k - 1)) {
        const int newEnd = lvlEnd - (k - 1) - 1;
        if (cs > newEnd)
            drop.push_back(last);
        else
            split[2 * last + 1] = newEnd;
    }
    std::vector<int> out;
    for (int i = 0; i < (int)split.size() / 2; ++i)
        if (std::find(drop.begin(), drop.end(), i) == drop.end()) {
            out.push_back(split[2 * i]);
            out.push_back(split[2 * i + 1]);
        }
    return out;
}

void consolidateLevels(const std::vector<int>& levelPtr, const std::vector<double>& elem, int k,
                       double cacheElem, const std::vector<int>& regions, std::vector<int>* groups,
                       std::vector<int>* glevel, std::vector<int>* hid) {
    const int L = (int)levelPtr.size() - 1;
    const int bl = k - 1;
    groups->assign(1, levelPtr[0]);
    glevel->assign(1, 0);
    hid->clear();
    int ctr = 0, hc = 0, hs, he;
    regionAt(regions, 0, &hs, &he);
    double sum = 0;
    for (int l = 0; l < L; ++l) {
        const double e = elem[l];
        sum += e;
        bool hopeless = false, force = false;
        if (hs != -1) {
            // k-1 levels on each side of a region are groups of their own,
            // the region itself is one group
            if (l >= hs - bl && l <= hs)
                force = true;
            else if (l > he && l <= he + bl)
                force = true;
            else if (l > hs && l <= he)
                hopeless = true;
            if (l == he + bl + 1) {
                regionAt(regions, ++hc, &hs, &he);
                force = true;
            }
            if (l == hs) hid->push_back(ctr + 1);
        }
        if (l > 0 && !hopeless && (sum >= cacheElem || force)) {
            sum = e;
            groups->push_back(levelPtr[l]);
            glevel->push_back(l);
            ++ctr;
        }
    }
    groups->push_back(levelPtr[L]);
    glevel->push_back(L);
}

int mpkPartition(Traverser* t, int n, long nnzInput, int k, double cacheBytes, double safety,
                 int cutoff0, int cutoff1, char mtxType, MPKPartition* out) {
    LevelData ld;
    int err = (mtxType == 'L' || mtxType == 'U') ? t->orderByRows(mtxType == 'U', &ld)
                                                 : t->traverse(DIST_POWER, 0, n, &ld);
    if (err != OK) return err;
    const int L = ld.totalLevel();
    out->highestPower = k;
    out->levelPtr.assign(L + 1, 0);
    for (int l = 0; l < L; ++l) out->levelPtr[l + 1] = out->levelPtr[l] + ld.levelRow[l];
    out->levelNnz = ld.levelNnz;
    // bytes per graph entry: matrix value + (k vector entries)/nnzr + index
    // + row pointer/nnzr (mtxPower::getBytePerNNZ src/matrixPower.cpp:134-140)
    const double nnzr = nnzInput / (double)n;
    out->bytePerNnz = (1 + k / nnzr) * sizeof(double) + (1 + 1 / nnzr) * sizeof(int);
    out->cacheElem = cacheBytes / out->bytePerNnz;
    auto violatedOf = [&](const std::vector<double>& elem, int cutoff) {
        std::vector<int> v;
        for (int l = 0; l < (int)elem.size(); ++l)
            if (elem[l] > out->cacheElem && (int)std::ceil(elem[l] / out->cacheElem) > cutoff) v.push_back(l);
        return v;
    };
    std::vector<double> elem(L);
    for (int l = 0; l < L; ++l) elem[l] = safety * (k + 1) * ld.levelNnz[l];
    out->violated = violatedOf(elem, cutoff0);
    out->hopelessRegions = identifyHopelessRegions(out->violated, k, L);
    consolidateLevels(out->levelPtr, elem, k, out->cacheElem, out->hopelessRegions, &out->groupPtr,
                      &out->groupLevel, &out->hopeless);
    out->children.clear();
    if (mtxType != 'N') return out->hopeless.empty() ? OK : ERR_NOT_IMPLEMENTED;
    // recursion stage: every hopeless group is traversed again together with
    // the groups within the working radius ceil(k/2) - 1 around it
    // (mtxPowerRecursive::recursivePartition src/matrixPowerRecursive.cpp:156-338)
    const int wbl = (int)std::ceil(k / 2.0) - 1;
    const std::vector<int>& g = out->groupPtr;
    for (int h : out->hopeless) {
        if (h - wbl < 0 || h + wbl + 1 >= (int)g.size()) return ERR_INTERNAL;
        MPKChild c;
        c.lo = g[h];
        c.hi = g[h + 1];
        c.group = h;
        for (int r = -wbl; r <= wbl; ++r)
            if (r != 0 && g[h + r + 1] > g[h + r]) c.boundaries.push_back({g[h + r], g[h + r + 1]});
        LevelData cl;
        std::vector<LevelData> bl;
        if ((err = t->traverseBoundary(c.lo, c.hi, c.boundaries, &cl, &bl)) != OK) return err;
        const int CL = cl.totalLevel();
        std::vector<int> cptr(CL + 1, c.lo);
        for (int l = 0; l < CL; ++l) cptr[l + 1] = cptr[l] + cl.levelRow[l];
        // elem counts the boundary regions' entries of the level too
        // (mtxPower::getElemUpperLimit src/matrixPower.cpp:88-99)
        std::vector<double> celem(CL);
        for (int l = 0; l < CL; ++l) {
            long nz = cl.levelNnz[l];
            for (const LevelData& b : bl) nz += b.levelNnz[l];
            celem[l] = safety * (k + 1) * nz;
        }
        if (!identifyHopelessRegions(violatedOf(celem, cutoff1), k, CL).empty()) return ERR_NOT_IMPLEMENTED;
        std::vector<int> gl, hid;
        consolidateLevels(cptr, celem, k, out->cacheElem, {}, &c.groupPtr, &gl, &hid);
        out->children.push_back(std::move(c));
    }
    return OK;
}

}  // namespace race
