// simdify.hpp -- SELL-C-sigma helpers of the RACE interface.
#pragma once
#include "engine.hpp"

namespace race {

// padding columns of a SELL-C-sigma matrix by the leaves of a schedule
void sellCSigmize(const Schedule& s, int simdWidth, int C, int* col, const int* chunkStart, const int* rl,
                  const int* clp);
// distance-2 rearrangement (after sellCSigmize); OK or ERR_INCOMPATIBILITY
template <typename T>
int simdifyResolve(int simdWidth, int C, int nrows, int* col, const int* chunkStart, const int* clp, T* val,
                   bool diagFirst);
// distance-1 rearrangement
template <typename T>
int simdifyD1(int simdWidth, int C, int nrows, int* col, const int* chunkStart, const int* cl, T* val,
              bool dist2Compatible);

}  // namespace race
