// bundle.hpp -- leaf bundles: the execution-order layout for Gauss-Seidel
// and Kaczmarz phases.
//
// The leaves of one colour phase are independent, so one workgroup runs
// several of them at once.  A bundle is a run of whole leaves of one phase
// (at most `cap` rows).  Its rows ("slots") are ordered by dependency wave.
//
// Every bundle has a local column space held in LDS:
//   Gauss-Seidel: [0, rows)      the bundle's own rows (updated by the waves)
//                 [rows, rows+e) the distinct columns outside the bundle that
//                                its rows read (constant during the phase)
//   Kaczmarz:     [0, cols)      every column the bundle's rows touch
// so a matrix entry is a 16-bit local index plus a value.  With tpr lanes per
// row, row entry k goes to lane k % tpr at register slot k / tpr and the
// bundle's block is laid out column-major over lanes:
//   entry(u, slot, lane) = off + u*(rows*tpr) + slot*tpr + lane
// so each load instruction of a workgroup reads one contiguous stream.
#pragma once
#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace race {

struct BundleTable {
    int kind = 0;                   // RACE_K_GS or RACE_K_KACZ
    int cap = 128;                  // rows per bundle
    int tpr = 1;                    // lanes per row (0: variable, see lanePtr)
    std::vector<int> phaseBundle;   // phase -> [first bundle, last bundle)
    std::vector<int> phaseChain;    // phase -> [first chain, last chain)
    std::vector<int> chainPtr;      // chain -> [first bundle, last bundle): run in order
                                    // by one workgroup (a bundle alone, or the
                                    // segments of a leaf too large for one bundle)
    std::vector<int> slot;          // bundle -> first slot
    std::vector<int> wave;          // bundle -> first wave
    std::vector<int> waveSlot;      // wave -> first slot (absolute)
    std::vector<long> off;          // bundle -> first entry of its block
    std::vector<int> len;           // bundle -> register entries per lane
    std::vector<int> extPtr;        // bundle -> first entry in ext
    std::vector<int> ext;           // global columns of the local column space
                                    // (GS: outside columns; KACZ: all columns)
    std::vector<int> rowId;         // slot -> matrix row
    std::vector<double> diag;       // slot -> diagonal value (GS)
    std::vector<double> rowNorm;    // slot -> sum of squares of the row (KACZ)
    // variable lanes per row (tpr == 0): row of slot i gets max(1,
    // ceil(len / kVarE)) lanes; entry k of the row goes to its lane k % nl
    // at register k / nl; the bundle's block is column-major over its lanes
    std::vector<int> lanePtr;       // bundle -> first lane
    std::vector<uint16_t> laneSlot; // lane -> slot inside its bundle
    std::vector<uint16_t> sidx;     // entries: local column
    std::vector<double> sval;       // entries: value
    long critical = 0;              // sum over phases of the deepest bundle
    int maxLen = 0;                 // largest per-lane width
    long paddedEntries = 0;
    bool ok = true;                 // false if a leaf exceeds the bundle limits
    // per-phase maxima (LDS sizing of each launch)
    std::vector<int> phRows, phLocal, phWaves, phExt;  // phExt: most outside columns of a bundle
    int numPhases() const { return (int)phaseBundle.size() - 1; }
};

// Build the bundle table of one sweep direction.  rowPtr/col/val: permuted
// CRS with the diagonal first in each row.  localCap bounds the local column
// space of a bundle (LDS entries, at most 65535).
// register entries per lane in the variable-lanes layout
constexpr int kVarE = 16;

// tpr == 0 selects variable lanes per row; laneCap then bounds the lanes of
// a bundle (the workgroup size) besides cap rows.
BundleTable buildBundles(const Schedule& s, int kind, int n, const int* rowPtr, const int* col,
                         const double* val, int cap, int localCap, int tpr, int laneCap = 0);

// Bundle tables of several matrices joined phase by phase: fused phase p
// holds the bundles of phase p of every table that has one, in table order.
// Slots, waves, entry blocks and outside-column lists are renumbered; row
// ids and column ids stay those of their own matrix (each bundle of a batch
// carries its own vectors).  bundleTable[k] = table of fused bundle k.
BundleTable concatBundles(const std::vector<const BundleTable*>& tabs, std::vector<int>* bundleTable);

// Dependencies between the chains of one sweep, for a launch that runs all
// of them in chain order (phase order) with a completion flag per chain
// instead of one launch per phase.  Chain c waits for the chains
// dep[depPtr[c] .. depPtr[c+1]), all earlier than c:
//   GS  : an earlier chain that writes a row c reads (an outside column of
//         c), or that reads a row c writes;
//   KACZ: for each column c touches, the previous chain touching it.
// For a fused table of several matrices, colBase[b] offsets the row and
// column ids of bundle b into one key space of nkeys entries (each matrix
// its own range); for one matrix colBase is null and nkeys = n.
// Returns the number of such pairs inside one phase (0 for a valid
// colouring: the leaves of a phase are independent).
long chainDeps(const BundleTable& t, long nkeys, const std::vector<long>* colBase, std::vector<int>* depPtr,
               std::vector<int>* dep);

// Local column ids (t.ext) for the one-launch sweep, with bit 31 set on
// every column an earlier chain of the launch (or an earlier bundle of the
// same chain) may have written: GS, a row of such a chain; KACZ, a column
// such a chain touched.  Only those must be read coherently after the
// dependency wait; the others still hold the value the launch started with.
std::vector<int> flowExt(const BundleTable& t, long nkeys, const std::vector<long>* colBase);

// Longest dependency path from every chain to the end of the sweep, for
// dependencies as chainDeps returns them (each an earlier chain): a chain
// costs its bundles' dependency waves plus fixedPerBundle per bundle, and
// its path is that cost plus the longest path among the chains that depend
// on it.
std::vector<long> chainLongestPaths(const BundleTable& t, const std::vector<int>& depPtr,
                                    const std::vector<int>& dep, int fixedPerBundle);


}  // namespace race
