// bfs.hpp -- GPU frontier BFS implementing RACE's traversal semantics.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "engine.hpp"
#include "hip_util.hpp"

namespace race {

// Level-synchronous frontier BFS on the device.  The graph is kept in its
// original numbering (RACE_PERMUTE_ON_FLY): a frontier row u is read as
// graph row perm[u] and a neighbour c is mapped to invPerm[c].
//
// Follows Traverse::TDStep / calculateDistance / findLevelData /
// permuteGraph (src/traverse_GAP_graphSoA.cpp:135-221, 236-383, 385-504,
// 506-641):
//  * the BFS root of a range is its first row;
//  * DIST_TWO lets an out-of-range neighbour of an in-range row join the
//    next frontier without a level ("bridge"), so distance-2 reachability
//    through rows outside the range is kept;
//  * a disconnected range restarts from its smallest unvisited row, one
//    empty level apart for colouring (DIST_ONE/TWO) and adjacent for POWER;
//  * the range is then stably sorted by level and the permutation composed
//    into perm/invPerm.
class GpuTraverser : public Traverser {
  public:
    GpuTraverser() {}
    ~GpuTraverser();
    // graph in original numbering (host arrays are copied to the device)
    int init(int n, const int* rowPtr, const int* col, const int* initPerm,
             const int* initInvPerm, hipStream_t stream);
    int traverse(int dist, int lo, int hi, LevelData* out) override;
    // disjoint ranges, traversed concurrently (one workgroup per range)
    int traverseBatch(int dist, const std::vector<std::pair<int, int>>& ranges,
                      std::vector<LevelData>* outs) override;
    int orderByRows(bool reverse, LevelData* out) override;
    int traverseBoundary(int lo, int hi, const std::vector<std::pair<int, int>>& bounds, LevelData* out,
                         std::vector<LevelData>* bout) override;
    const std::vector<int>& perm() const override { return hPerm; }
    const std::vector<int>& invPerm() const override { return hInv; }
    int rowLen(int origRow) const override { return hRowLen[origRow]; }
    int n = 0;
    long nnz = 0;
    // device permutation (perm[new] = old)
    int* devPerm() const { return dPerm.p; }
    int* devInvPerm() const { return dInv.p; }
    // ranges of at most this many rows are traversed by one workgroup each
    // (batched); larger ranges level by level over the whole device
    int rangeBlockMax = 1 << 17;
    double lastTraverseMs = 0;
    long traversals = 0;

  private:
    int traverseLevels(int dist, int lo, int hi, LevelData* out);
    int bfsLevels(int dist, int lo, int hi, const unsigned char* bmask, int* levels);
    int histogram(int lo, int hi, int totalLevel, LevelData* out);
    int sortByLevel(int lo, int hi, int totalLevel);
    DevBuf<unsigned char> dMask;  // boundary rows of a recursion stage's traversal
    hipStream_t st = nullptr;
    int stampNext = 0, bvisitSlots = 0;
    DevBuf<int> dBVisit, dBatchQ, dLvRow, dLvNnz, dLvStart, dPackRow, dPackNnz, dNLev;
    DevBuf<unsigned char> dJobs;
    std::vector<int> hNLev, hPackRow, hPackNnz;
    std::vector<int> hPerm, hInv, hRowLen;
    DevBuf<int> dRowPtr, dCol, dPerm, dInv, dVisit, dDist, dQ0, dQ1, dCount;
    DevBuf<int> dLevelRow, dLevelNnz, dKeys, dKeysOut, dVals, dValsOut, dTmp;
    DevBuf<unsigned char> dSortTmp;
    int* hPinned = nullptr;  // pinned host words for counters
};

}  // namespace race
