// rcm.cpp -- reverse Cuthill-McKee ordering, the initial permutation the
// reference's harness hands to RACE (sparsemat::doRCM
// examples/src/helpers/sparsemat.cpp:715-760 calls SpMP's
// CSR::getRCMPermutation, then passes the result as initPerm/initInvPerm,
// :762-771).  SpMP is not part of the reference tree; this is the
// published algorithm:
//   per connected component (in order of its smallest unvisited row), a
//   pseudo-peripheral root by the George-Liu iteration (BFS from the
//   current root; move to a minimum-degree row of the last level while the
//   eccentricity grows), then a Cuthill-McKee BFS from that root that
//   visits the unvisited neighbours of each row in increasing degree
//   (ties: increasing index); the final order is reversed.
// Host-only (no device work): it runs before the Interface is created.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/race_amd.h"
#include "engine.hpp"

namespace race {
namespace {

// BFS from root over unvisited rows (mark != stamp): returns the level
// sets in order; marks with stamp
int bfsLevels(int root, const int* rp, const int* col, std::vector<int>& mark, int stamp,
              std::vector<int>& order, std::vector<int>& levelPtr) {
    order.clear();
    levelPtr.assign(1, 0);
    order.push_back(root);
    mark[root] = stamp;
    size_t lo = 0;
    while (lo < order.size()) {
        const size_t hi = order.size();  // level [lo, hi), non-empty
        levelPtr.push_back((int)hi);
        for (size_t q = lo; q < hi; ++q) {
            const int r = order[q];
            for (int j = rp[r]; j < rp[r + 1]; ++j) {
                const int c = col[j];
                if (mark[c] != stamp && mark[c] != -1) {
                    mark[c] = stamp;
                    order.push_back(c);
                }
            }
        }
        lo = hi;
    }
    return (int)levelPtr.size() - 1;
}

}  // namespace
}  // namespace race

using namespace race;

int race_rcm(int n, const int* rowPtr, const int* col, int* perm, int* inv_perm) {
    if (n <= 0 || !rowPtr || !col || !perm || !inv_perm || rowPtr[0] != 0) return ERR_INVALID_ARG;
    for (int r = 0; r < n; ++r)
        if (rowPtr[r + 1] < rowPtr[r]) return ERR_INVALID_ARG;
    const int nnz = rowPtr[n];
    for (int j = 0; j < nnz; ++j)
        if (col[j] < 0 || col[j] >= n) return ERR_INVALID_ARG;
    // the pattern must be symmetric (SpMP's RCM refuses otherwise,
    // examples/src/helpers/sparsemat.cpp:729-738): compare each row's
    // sorted off-diagonal columns with the transposed pattern
    std::vector<int> deg(n), tp(n + 1, 0), tc((size_t)nnz);
    for (int r = 0; r < n; ++r)
        for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) ++tp[col[j] + 1];
    for (int r = 0; r < n; ++r) tp[r + 1] += tp[r];
    {
        std::vector<int> pos(tp.begin(), tp.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) tc[pos[col[j]]++] = r;
    }
    std::vector<int> a, b;
    for (int r = 0; r < n; ++r) {
        a.assign(col + rowPtr[r], col + rowPtr[r + 1]);
        b.assign(tc.begin() + tp[r], tc.begin() + tp[r + 1]);
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        if (a != b) return ERR_MATRIX_SYMM;
        deg[r] = (int)a.size() - (std::binary_search(a.begin(), a.end(), r) ? 1 : 0);
    }
    // neighbour lists without the diagonal, each sorted by (degree, index)
    std::vector<int> np(n + 1, 0), nc;
    nc.reserve((size_t)nnz);
    for (int r = 0; r < n; ++r) {
        a.assign(col + rowPtr[r], col + rowPtr[r + 1]);
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        a.erase(std::remove(a.begin(), a.end(), r), a.end());
        std::sort(a.begin(), a.end(), [&](int u, int v) { return deg[u] != deg[v] ? deg[u] < deg[v] : u < v; });
        nc.insert(nc.end(), a.begin(), a.end());
        np[r + 1] = (int)nc.size();
    }
    // mark: -1 = placed (in an earlier component), otherwise a BFS stamp
    std::vector<int> mark(n, 0), order, lp, cm;
    cm.reserve(n);
    int stamp = 0;
    for (int s = 0; s < n; ++s) {
        if (mark[s] == -1) continue;
        // George-Liu pseudo-peripheral row of s's component
        int root = s;
        int ecc = bfsLevels(root, np.data(), nc.data(), mark, ++stamp, order, lp);
        for (int it = 0; it < n; ++it) {
            int cand = -1;
            for (int q = lp[ecc - 1]; q < lp[ecc]; ++q) {
                const int v = order[q];
                if (cand < 0 || deg[v] < deg[cand] || (deg[v] == deg[cand] && v < cand)) cand = v;
            }
            std::vector<int> o2, l2;
            const int e2 = bfsLevels(cand, np.data(), nc.data(), mark, ++stamp, o2, l2);
            if (e2 <= ecc) break;
            root = cand;
            ecc = e2;
            order.swap(o2);
            lp.swap(l2);
        }
        // Cuthill-McKee from root: neighbours in (degree, index) order
        const size_t c0 = cm.size();
        cm.push_back(root);
        mark[root] = -1;
        for (size_t h = c0; h < cm.size(); ++h) {
            const int r = cm[h];
            for (int j = np[r]; j < np[r + 1]; ++j) {
                const int c = nc[j];
                if (mark[c] != -1) {
                    mark[c] = -1;
                    cm.push_back(c);
                }
            }
        }
    }
    // reverse: new row i is old row cm[n - 1 - i]
    for (int i = 0; i < n; ++i) {
        perm[i] = cm[n - 1 - i];
        inv_perm[perm[i]] = i;
    }
    return OK;
}
