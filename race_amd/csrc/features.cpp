// features.cpp -- structural features of a sparse pattern for the layout
// tuner (host only).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/race_amd.h"

namespace {

// Level sizes of a breadth-first search from row 0; a disconnected part
// restarts from its smallest unvisited row.
std::vector<long> bfsLevelSizes(int n, const int* rp, const int* col) {
    std::vector<unsigned char> seen(n, 0);
    std::vector<int> queue(n);
    std::vector<long> sizes;
    int scan = 0, head = 0, tail = 0;
    for (;;) {
        while (scan < n && seen[scan]) ++scan;
        if (scan >= n) break;
        seen[scan] = 1;
        queue[tail++] = scan;
        // the queue holds the levels one after the other
        while (head < tail) {
            const int end = tail;
            sizes.push_back((long)(end - head));
            for (; head < end; ++head) {
                const int r = queue[head];
                for (int j = rp[r]; j < rp[r + 1]; ++j) {
                    const int c = col[j];
                    if ((unsigned)c < (unsigned)n && !seen[c]) {
                        seen[c] = 1;
                        queue[tail++] = c;
                    }
                }
            }
        }
    }
    return sizes;
}

}  // namespace

extern "C" int race_matrix_features(int n, const int* rowPtr, const int* col, double* out) {
    if (n <= 0 || !rowPtr || !col || !out) return RACE_ERR_INVALID_ARG;
    const long nnz = rowPtr[n];
    // row lengths as a histogram: moments, shares and the sorted order of
    // the Gini coefficient in one pass over n plus one over the lengths
    int mx = 0, mn = 1 << 30;
    for (int r = 0; r < n; ++r) {
        const int l = rowPtr[r + 1] - rowPtr[r];
        mx = std::max(mx, l);
        mn = std::min(mn, l);
    }
    std::vector<long> hist((size_t)mx + 1, 0);
    for (int r = 0; r < n; ++r) ++hist[rowPtr[r + 1] - rowPtr[r]];
    const double sum = (double)nnz;
    const double mean = sum / n;
    double var = 0, big = 0, over32 = 0, cum = 0, cumsum = 0;
    for (int l = 0; l <= mx; ++l) {
        const double h = (double)hist[l];
        if (h == 0) continue;
        var += h * (l - mean) * (l - mean);
        if (l > 2 * mean) big += h;
        if (l > 32) over32 += h;
        // the cumulative sums of the sorted lengths over the h rows of length l
        cumsum += h * cum + l * h * (h + 1) / 2;
        cum += h * l;
    }
    long dsum = 0, dmax = 0, near = 0;
    for (int r = 0; r < n; ++r)
        for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) {
            const long d = std::labs((long)col[j] - r);
            dsum += d;
            dmax = std::max(dmax, d);
            near += d <= 64;
        }
    const std::vector<long> lv = bfsLevelSizes(n, rowPtr, col);
    double lmax = 0, lsum = 0;
    for (long q : lv) {
        lmax = std::max(lmax, (double)q);
        lsum += q;
    }
    const double ent = std::max((double)nnz, 1.0);
    out[0] = std::log10((double)n);
    out[1] = std::log10(ent);
    out[2] = mean;
    out[3] = std::sqrt(var / n) / mean;
    out[4] = std::log10((double)mx);
    out[5] = mn;
    out[6] = big / n;
    out[7] = over32 / n;
    out[8] = nnz ? (double)dsum / nnz / n : 0.0;
    out[9] = (double)dmax / n;
    out[10] = nnz ? (double)near / nnz : 0.0;
    out[11] = std::log10((double)lv.size());
    out[12] = lmax / n;
    out[13] = (lsum / lv.size()) / lmax;
    out[14] = 1.0 - 2.0 * cumsum / (n * sum) + 1.0 / n;
    out[15] = std::log10((double)nnz / lv.size());
    return RACE_SUCCESS;
}
