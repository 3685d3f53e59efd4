// simdify.cpp -- SELL-C-sigma helpers of the RACE interface: padding
// columns that never alias inside a SIMD group (sell_c_sigmize), column
// rearrangement so that no two lanes of a SIMD group touch the same column
// in one step (simdify, distance 2) and placement of the dependency columns
// of preceding lanes at the end of each row (simdifyD1, distance 1).
//
// SELL-C-sigma addressing (as the reference's callers use it): entry j of
// row r sits at chunkStart[r / C] + r % C + j * C; rl[r] is the row length,
// clp[c] the padded length of chunk c.
#include "simdify.hpp"

#include <cstdio>
#include <vector>

#include "../../include/race_amd.h"

namespace race {

// src/sell_c_sigmize.cpp:26-62.  The reference runs the kernel as a RACE
// callback, so `start` is the first row of the leaf that covers the row:
// padding slot j >= rl[row] of row gets column start + (row - start) %
// simdWidth.  Every row lies in exactly one leaf, so the leaf order does not
// matter.
void sellCSigmize(const Schedule& s, int simdWidth, int C, int* col, const int* chunkStart, const int* rl,
                  const int* clp) {
    for (size_t l = 0; l < s.leafLo.size(); ++l) {
        const int start = s.leafLo[l], end = s.leafHi[l];
        for (int row = start; row < end; ++row) {
            const int chunk = row / C;
            const int base = chunkStart[chunk] + row % C;
            for (int j = rl[row]; j < clp[chunk]; ++j) col[base + j * C] = start + ((row - start) % simdWidth);
        }
    }
}

namespace {

template <typename T>
void swapEntries(int* col, T* val, long a, long b) {
    const int tc = col[a];
    col[a] = col[b];
    col[b] = tc;
    const T tv = val[a];
    val[a] = val[b];
    val[b] = tv;
}

}  // namespace

// include/simdify.h:70-357 after the sell_c_sigmize call (:78).  For every
// SIMD group (simdWidth consecutive rows of a chunk) and every position j
// (from 1 when the diagonal is stored first), lane i takes the first entry
// at or after j, wrapping around, whose column no earlier lane holds at j
// and, once wrapped, that does not collide with the other lanes at the
// candidate position; the entry is swapped into position j.
template <typename T>
int simdifyResolve(int simdWidth, int C, int nrows, int* col, const int* chunkStart, const int* clp, T* val,
                   bool diagFirst) {
    if (simdWidth <= 0 || C <= 0) return ERR_INVALID_ARG;
    if (C % simdWidth)
        std::fprintf(stderr, "race_amd: simdify: chunk length %d is not a multiple of the SIMD width %d\n", C,
                     simdWidth);
    const int nChunk = nrows / C;
    const int simdInChunk = C / simdWidth;
    const int startingIdx = diagFirst ? 1 : 0;
    std::vector<int> simdCol(simdWidth);
    for (int chunk = 0; chunk <= nChunk; ++chunk) {
        // rows of a chunk counted as the reference does (chunk * C when C is
        // a multiple of the SIMD width)
        const int chunkRow0 = chunk * simdInChunk * simdWidth;
        if (chunkRow0 >= nrows) break;  // no rows: nothing below would run
        const int width = clp[chunk];
        for (int simdIdx = 0; simdIdx < simdInChunk; ++simdIdx) {
            const long lane0 = (long)chunkStart[chunk] + (long)simdIdx * simdWidth;
            const int row0 = chunkRow0 + simdIdx * simdWidth;
            for (int j = startingIdx; j < width; ++j) {
                for (int i = 0; i < simdWidth && row0 + i < nrows; ++i) {
                    int exchangeIdx = j;
                    int ctr = 0;
                    bool wrapped = false;
                    bool repeat = true;
                    while (repeat) {
                        repeat = false;
                        const int cur = col[lane0 + (long)C * exchangeIdx + i];
                        for (int k = 0; k < i; ++k)
                            if (cur == simdCol[k]) {
                                repeat = true;
                                break;
                            }
                        if (wrapped) {
                            // the entry at position j would move to exchangeIdx: it
                            // must not collide there with another lane
                            const int here = col[lane0 + (long)C * j + i];
                            for (int k = 0; k < simdWidth; ++k)
                                if (k != i && here == col[lane0 + (long)C * exchangeIdx + k]) {
                                    repeat = true;
                                    break;
                                }
                        }
                        if (repeat) {
                            if (ctr >= width) {
                                std::fprintf(stderr,
                                             "race_amd: cannot simdify row %d (a row length of at least %d is "
                                             "needed; chunk length %d)\n",
                                             row0 + i, simdWidth, width);
                                return ERR_INCOMPATIBILITY;
                            }
                            if (exchangeIdx != width - 1) {
                                ++exchangeIdx;
                            } else {
                                exchangeIdx = startingIdx;
                                wrapped = true;
                            }
                            ++ctr;
                        }
                    }
                    simdCol[i] = col[lane0 + (long)C * exchangeIdx + i];
                    if (exchangeIdx != j) {
                        if ((j == 0 || exchangeIdx == 0) && diagFirst)
                            std::fprintf(stderr, "race_amd: simdify moved a diagonal entry\n");
                        swapEntries(col, val, lane0 + (long)C * j + i, lane0 + (long)C * exchangeIdx + i);
                    }
                }
            }
        }
    }
    return OK;
}

// include/simdify.h:389-447.  In lane i >= 1 of a SIMD group, an entry on
// the row of lane m < i is swapped to position cl - 1 - m (lane 0's row
// last), or to cl - 1 - (i - 1 - m) when dist2Compatible (the nearest
// lane's row last).  The row counter carries over from one SIMD group to the
// next exactly as in the reference loop (its bound test at the start of a
// group sees the previous group's last row plus one).
template <typename T>
int simdifyD1(int simdWidth, int C, int nrows, int* col, const int* chunkStart, const int* cl, T* val,
              bool dist2Compatible) {
    if (simdWidth <= 0 || C <= 0) return ERR_INVALID_ARG;
    const int nChunk = nrows / C;
    const int simdInChunk = C / simdWidth;
    int row = 0;
    std::vector<int> dependentRow;
    for (int chunk = 0; chunk <= nChunk; ++chunk) {
        for (int simdIdx = 0; simdIdx < simdInChunk; ++simdIdx) {
            for (int i = 1; i < simdWidth && row < nrows; ++i, ++row) {
                row = chunk * simdInChunk * simdWidth + simdIdx * simdWidth + i;
                dependentRow.assign(i, 0);
                for (int k = 0; k < i; ++k) dependentRow[dist2Compatible ? k : i - k - 1] = row - (k + 1);
                const long lane = (long)chunkStart[chunk] + (long)simdIdx * simdWidth + i;
                const int width = cl[chunk];
                for (int j = 0; j < width; ++j) {
                    for (int k = 0; k < i; ++k) {
                        if (col[lane + (long)C * j] == dependentRow[k]) {
                            const int exchangeIdx = width - (k + 1);
                            if (exchangeIdx != j) swapEntries(col, val, lane + (long)C * j, lane + (long)C * exchangeIdx);
                        }
                    }
                }
            }
        }
    }
    return OK;
}

template int simdifyResolve<double>(int, int, int, int*, const int*, const int*, double*, bool);
template int simdifyResolve<float>(int, int, int, int*, const int*, const int*, float*, bool);
template int simdifyD1<double>(int, int, int, int*, const int*, const int*, double*, bool);
template int simdifyD1<float>(int, int, int, int*, const int*, const int*, float*, bool);

}  // namespace race
