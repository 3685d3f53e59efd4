// schedule_io.hpp -- a coloured handle's schedule as a byte string.
//
// The reference has no save/load: its zone tree and MPK tree are rebuilt
// by every RACEColor call and can only be printed (ZoneTree::printTree
// src/zone_tree.cpp:307-339, mtxPowerRecursive::printTree
// src/matrixPowerRecursive.cpp:447-490).  Saving the result lets a caller
// skip the preprocessing for a matrix it has seen before (and keep layout
// and tuner decisions that depend on it).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace race {

struct SavedSchedule {
    int nrow = 0, dist = 0, d2 = 0, lbTarget = 0, requested = 1, available = 1;
    int highestPower = 1, highestSubPower = 1;
    uint64_t pattern = 0;               // patternHash of the graph it was built on
    std::vector<int> perm, inv;         // perm[new] = old
    bool hasTree = false;
    std::vector<ZoneLeaf> tree;         // colouring (dist ONE / TWO)
    bool hasMpk = false;
    MPKPartition mpk;                   // POWER
};

// FNV-1a over the row pointers and column indices
uint64_t patternHash(int n, const int* rowPtr, const int* col);

std::vector<unsigned char> encodeSchedule(const SavedSchedule& s);
// Parses and checks a byte string: framing, checksum, and that everything
// the executors index with is in range (perm a bijection with inv its
// inverse; tree children after their parent, each leaf the child of one
// parent at most, row ranges inside [0, nrow); MPK level and group pointers
// from 0 to nrow).  Returns OK or ERR_INVALID_ARG.
int decodeSchedule(const void* buf, size_t size, SavedSchedule* out);

}  // namespace race
