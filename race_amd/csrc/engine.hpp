// engine.hpp -- host side of the MI355X RACE engine.
//
// The level-granular scheduling logic of RACE stays on the host: the load
// balancer that groups BFS levels into zones, the zone tree, the level
// recursion that re-partitions zones until the requested parallelism is met,
// and the MPK level grouping.  Every BFS these steps need is delegated to a
// Traverser, which in this library is the GPU frontier BFS (bfs.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace race {

enum Dist { DIST_ONE = 1, DIST_TWO = 2, DIST_POWER = 3 };
enum D2Method { TWO_BLOCK = 0, THREE_BLOCK = 1 };
enum LBTarget { LB_ROW = 0, LB_NNZ = 1 };

// Error codes, value-compatible with the reference's RACE_error
// (include/error.h:29-41).
enum Err {
    OK = 0,
    ERR_INVALID_ARG = 1,
    ERR_MATRIX_SYMM = 2,
    ERR_D2_COLOR = 3,
    ERR_D1_COLOR = 4,
    ERR_GRAPH_TRAVERSAL = 5,
    ERR_INCOMPATIBILITY = 6,
    ERR_HWLOC = 7,
    ERR_PIN = 8,
    ERR_INTERNAL = 9,
    ERR_NOT_IMPLEMENTED = 10,
};

// blocks (colours) per thread and minimum levels per block
// (getBlockPerThread / getMinGap, include/macros.h:29-34)
inline int blockPerThread(int dist, int d2) {
    return dist == DIST_POWER ? 1 : dist == DIST_ONE ? 2 : (d2 == TWO_BLOCK ? 2 : 3);
}
inline int minGap(int dist, int d2) {
    return dist == DIST_POWER ? 1 : dist == DIST_ONE ? 1 : (d2 == TWO_BLOCK ? 2 : 1);
}
// roundDouble include/utility.h:194-198
inline int roundDouble(double v) {
    int vi = static_cast<int>(v);
    return ((v - vi) < 0.5) ? vi : vi + 1;
}

struct LevelData {
    std::vector<int> levelRow;  // rows per BFS level
    std::vector<int> levelNnz;  // graph entries per BFS level
    int totalLevel() const { return (int)levelRow.size(); }
};

// Result of one load-balancing step over a level histogram.
struct LBResult {
    std::vector<int> levelPtr;   // block -> first level (blocks*colours + 1)
    std::vector<int> scale;      // nested thread count of each thread
    std::vector<int> zoneRows;   // rows per block
    std::vector<int> numBlocks;  // cache sub-blocks per block
    std::vector<int> subRows;    // rows per sub-block
    int threads() const { return (int)scale.size(); }
};

// LB::splitZones (src/lb.cpp:188-667) including the variance balancing and
// the (effectively disabled) cache sub-blocking.
LBResult splitZones(const LevelData& ld, int nThreads, double efficiency, int dist, int d2,
                    int lbTarget);

// Threads per child group of a partitioned range
// (LevelRecursion::calculateIdealNthreads, src/level_recursion.cpp:165-245).
std::vector<int> splitThreads(const std::vector<int>& weights, int total, int parentThreads);

struct ZoneLeaf {
    std::vector<int> valueZ;    // row boundaries of the leaf's sub-blocks
    std::vector<int> children;  // [begin,end) leaf indices per sub-block
    int nthreadsZ = 1;
    int idealNthreadsZ = 1;
    int totalSubBlocks = 1;
    int parentZ = -2;
    int effRowZ = 0;
    int stage = 0;
    bool reachedLimit = false;
};

class ZoneTree {
  public:
    ZoneTree(int dist, int d2, int lbTarget);
    int dist, d2, lbTarget, bpt;
    std::vector<ZoneLeaf> tree;
    // ZoneTree::spawnChild src/zone_tree.cpp:243-305
    bool spawnChild(int parent, int parentSub, int request, const LevelData& ld, double eff);
    int maxStage() const;

  private:
    int keyChildEffRow(int p) const;  // findKeyChild :91-155
    void updateEffRow(int p);         // :158-170
    void updateNThreads(int p);       // :172-206
};

// Something that runs RACE traversals: BFS in [lo,hi) with distance-1/2
// semantics, the level histograms, and the stable by-level permutation of
// the range composed into the global permutation.
class Traverser {
  public:
    virtual ~Traverser() {}
    virtual int traverse(int dist, int lo, int hi, LevelData* out) = 0;  // Err
    // disjoint ranges whose rows are final: same result as traversing them
    // one after another in the given order
    virtual int traverseBatch(int dist, const std::vector<std::pair<int, int>>& ranges,
                              std::vector<LevelData>* outs) {
        outs->assign(ranges.size(), LevelData());
        for (size_t r = 0; r < ranges.size(); ++r) {
            int e = traverse(dist, ranges[r].first, ranges[r].second, &(*outs)[r]);
            if (e) return e;
        }
        return 0;
    }
    // mtxType "L" / "U" of the POWER traversal (Traverse::calculateDistance
    // src/traverse_GAP_graphSoA.cpp:316-335): no BFS, row i is level i
    // ("L") or level n-1-i ("U"); the stable sort by level then keeps or
    // reverses the current order, composed into the permutation
    virtual int orderByRows(bool reverse, LevelData* out) = 0;
    // The POWER traversal of a recursion stage (mtxPowerRecursive, a
    // hopeless region [lo,hi) with boundary regions): boundary rows reached
    // from the frontier take the level and are expanded but not counted,
    // untouched boundary rows go to one extra last level, the range and
    // every region are sorted by level (src/traverse_GAP_graphSoA.cpp:76-88,
    // 184-197, 449-575); bout: one histogram per boundary region
    virtual int traverseBoundary(int lo, int hi, const std::vector<std::pair<int, int>>& bounds,
                                 LevelData* out, std::vector<LevelData>* bout) = 0;
    virtual const std::vector<int>& perm() const = 0;     // host mirror, perm[new]=old
    virtual const std::vector<int>& invPerm() const = 0;  // host mirror
    virtual int rowLen(int origRow) const = 0;            // graph row length
};

// LevelRecursion (src/level_recursion.cpp)
class LevelRecursion {
  public:
    LevelRecursion(Traverser* t, int n, int requestThreads, int dist, int d2, int lbTarget,
                   const std::vector<double>& effList, const std::vector<int>& lvlThreads);
    int levelBalancing(ZoneTree* zt);
    int availableThreads = -1;

  private:
    Traverser* trav;
    ZoneTree* zt = nullptr;
    int n, request, dist, d2, lbTarget;
    std::vector<double> effVec;
    std::vector<int> lvlThreadsVec;
    double efficiency(int lvl) const;
    int lvlThreads(int lvl) const;
    long nnzSum(int lo, int hi) const;
    void idealNthreads(int parent, int parentSub, int currLvl);
    int recursivePartition(int parent, int parentSub, int currLvl);
};

// Flattened colour schedule: leaves grouped into barrier-separated phases.
// Leaves of one phase are distance-k independent of each other.
struct Schedule {
    std::vector<int> phasePtr;  // phase -> [first leaf, last leaf)
    std::vector<int> leafLo;    // leaf rows [lo, hi)
    std::vector<int> leafHi;
    bool rev = false;           // leaves run hi-1 .. lo and phases in reverse colour order
    int numPhases() const { return (int)phasePtr.size() - 1; }
};
// Phases of recursiveCall's execution (src/functionManager.cpp:199-265).
Schedule flattenSchedule(const ZoneTree& zt, bool rev, int nrow);

// MPK level groups (mtxPowerRecursive::findPartition / recursivePartition
// src/matrixPowerRecursive.cpp:156-445 with mtxPower::findPartition /
// identifyHopelessRegions / consolidatePartition src/matrixPower.cpp:147-940).
struct MPKChild {
    int lo = 0, hi = 0;              // rows of the hopeless group
    int group = 0;                   // its stage-0 group
    std::vector<int> groupPtr;       // the stage's own level groups over [lo,hi)
    std::vector<std::pair<int, int>> boundaries;  // boundary regions traversed with it
};
struct MPKPartition {
    std::vector<int> levelPtr;    // BFS level -> first row
    std::vector<int> groupPtr;    // level group -> first row
    std::vector<int> groupLevel;  // level group -> first BFS level
    std::vector<int> levelNnz;
    std::vector<int> violated;         // levels over the cache by more than the cut-off
    std::vector<int> hopelessRegions;  // [start, end] level pairs (inclusive)
    std::vector<int> hopeless;         // stage-0 group of each hopeless region
    std::vector<MPKChild> children;    // the recursion stage of each hopeless group
    double cacheElem = 0, bytePerNnz = 0;
    int highestPower = 1;
};
// Hopeless regions of one node from the violated levels
// (mtxPower::identifyHopelessRegions src/matrixPower.cpp:147-502).
std::vector<int> identifyHopelessRegions(const std::vector<int>& violated, int k, int totalLevel);
// Level groups around hopeless regions (mtxPower::consolidatePartition
// src/matrixPower.cpp:685-880): groups (row pointers), first level per group,
// group index of each hopeless region.
void consolidateLevels(const std::vector<int>& levelPtr, const std::vector<double>& elem, int k,
                       double cacheElem, const std::vector<int>& regions, std::vector<int>* groups,
                       std::vector<int>* glevel, std::vector<int>* hid);
// cutoff0 / cutoff1: cache-violation cut-offs of stage 0 and of the
// recursion stage; a hopeless region inside a recursion stage returns
// ERR_NOT_IMPLEMENTED
int mpkPartition(Traverser* t, int n, long nnzInput, int highestPower, double cacheBytes,
                 double safety, int cutoff0, int cutoff1, char mtxType, MPKPartition* out);

// Parse comma separated integers with atoi semantics (getEnv, utility.h:94-111).
std::vector<int> envIntList(const char* name);

}  // namespace race
