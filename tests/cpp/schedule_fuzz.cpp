// Host-only fuzz of the schedule decoder (race_amd/csrc/schedule_io.cpp),
// built with AddressSanitizer and UBSan by tests/test_host_logic.py: random
// field changes in a valid encoded schedule, checksum recomputed, so the
// content checks must refuse what the executors could not index with.
// Prints the accepted count; the sanitizers abort on any invalid access.
#include <cstdio>
#include <cstring>
#include <random>
#include "schedule_io.hpp"
using namespace race;
int main() {
    SavedSchedule s;
    s.nrow = 6; s.dist = 1; s.d2 = 0; s.lbTarget = 1; s.requested = 2; s.available = 2;
    s.perm = {5, 4, 3, 2, 1, 0}; s.inv = {5, 4, 3, 2, 1, 0};
    s.hasTree = true;
    ZoneLeaf root; root.valueZ = {0, 6}; root.children = {1, 5}; root.nthreadsZ = 2; root.effRowZ = 6; root.parentZ = -1;
    s.tree.push_back(root);
    int q[5] = {0, 2, 3, 5, 6};
    for (int i = 0; i < 4; ++i) { ZoneLeaf l; l.valueZ = {q[i], q[i + 1]}; l.parentZ = 0; l.effRowZ = q[i + 1] - q[i]; l.stage = 1; s.tree.push_back(l); }
    std::vector<unsigned char> b = encodeSchedule(s);
    SavedSchedule o;
    if (decodeSchedule(b.data(), b.size(), &o)) { printf("valid rejected\n"); return 1; }
    std::mt19937 rng(5);
    int acc = 0;
    for (int it = 0; it < 100000; ++it) {
        std::vector<unsigned char> c(b.begin(), b.end() - 8);
        int k = 1 + rng() % 3;
        for (int j = 0; j < k; ++j) {
            size_t pos = 12 + 4 * (rng() % ((c.size() - 12) / 4));
            int vals[] = {-1, 0, 1, 2, 3, 5, 6, 7, 1 << 20, -(1 << 30), (int)(rng() % 24) - 8};
            int v = vals[rng() % 11];
            memcpy(&c[pos], &v, 4);
        }
        // reseal: FNV-1a of the body
        uint64_t h = 1469598103934665603ull;
        for (unsigned char ch : c) { h ^= ch; h *= 1099511628211ull; }
        c.insert(c.end(), (unsigned char*)&h, (unsigned char*)&h + 8);
        if (rng() % 7 == 0) c.resize(rng() % c.size());
        SavedSchedule d;
        if (decodeSchedule(c.data(), c.size(), &d) == 0) ++acc;
    }
    printf("accepted %d of 100000\n", acc);
    return 0;
}
