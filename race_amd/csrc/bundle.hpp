// bundle.hpp -- leaf bundles: the execution-order SELL layout for
// Gauss-Seidel and Kaczmarz phases.
//
// The leaves of one colour phase are independent, so a workgroup can run
// several of them at once.  A bundle is a run of whole leaves of one phase
// (at most `cap` rows).  Its rows are re-ordered by dependency wave, and the
// bundle's matrix rows are stored column-major (SELL with the bundle as the
// chunk), so lane i of the workgroup reads entry j of slot i at
// off + j*rows + i: one coalesced stream per entry column.
//
// Column encoding: c >= 0 is a global column (constant during the phase);
// c < 0 is ~slot of a column that lives in the workgroup's LDS copy of x.
#pragma once
#include <vector>

#include "engine.hpp"

namespace race {

struct BundleTable {
    int kind = 0;                   // RACE_K_GS or RACE_K_KACZ
    int cap = 512;                  // rows per bundle
    std::vector<int> phaseBundle;   // phase -> [first bundle, last bundle)
    std::vector<int> slot;          // bundle -> first slot (rows in slot order)
    std::vector<int> wave;          // bundle -> first wave
    std::vector<int> waveSlot;      // wave -> first slot (absolute)
    std::vector<long> off;          // bundle -> first SELL entry
    std::vector<int> len;           // bundle -> register entries per lane
    int tpr = 1;                    // lanes per row
    int maxLen = 0;                 // largest per-lane width over all bundles
    std::vector<int> innerPtr;      // slot -> first staged entry, bundle-relative
    std::vector<int> innerEnd;      // bundle -> number of staged entries
    std::vector<int> rowId;         // slot -> matrix row
    std::vector<double> diag;       // slot -> diagonal value (GS)
    std::vector<double> rowNorm;    // slot -> sum of squares of the row (KACZ)
    std::vector<int> colSet;        // KACZ: bundle columns held in LDS
    std::vector<int> colSetPtr;     // KACZ: bundle -> first entry in colSet
    std::vector<int> scol;          // SELL columns (encoded)
    std::vector<double> sval;       // SELL values
    long critical = 0;              // sum over phases of the deepest bundle
    int numPhases() const { return (int)phaseBundle.size() - 1; }
    int maxRows = 0, maxInner = 0, maxWaves = 0, maxCols = 0;
    // per-phase maxima, used to size the LDS of each phase launch
    std::vector<int> phRows, phInner, phWaves, phCols, phEntries;
    long paddedEntries = 0;
    bool ok = true;                 // false if some leaf exceeds the bundle limits
};

// Build the bundle table of a schedule direction.  rowPtr/col/val: permuted
// CRS with the diagonal first in each row.
// tpr lanes per row: row entry k (0-based among the stored entries) goes to
// lane k % tpr at register position k / tpr; a bundle's block is laid out as
// off + u*(rows*tpr) + slot*tpr + lane.
BundleTable buildBundles(const Schedule& s, int kind, int n, const int* rowPtr, const int* col,
                         const double* val, int cap, int innerCap, int tpr);

}  // namespace race
