// schedule_io.cpp -- byte-string form of a coloured handle's schedule.
//
// Layout (little endian): "RACESCHD", u32 version, eight i32 header fields
// (nrow, dist, d2, lbTarget, requested, available, highestPower,
// highestSubPower), u64 pattern hash, i32 vectors as (i32 count, data):
// perm, inv; i32 hasTree, then per leaf (i32 count of leaves first) seven
// i32 fields and the vectors valueZ, children; i32 hasMpk, then the vectors
// levelPtr, groupPtr, groupLevel, levelNnz, hopeless (the group of each
// hopeless region), f64 cacheElem, f64 bytePerNnz, i32 highestPower; u64
// FNV-1a checksum of all bytes before.  The recursion stage's own groups of
// a hopeless region are not stored (the permutation they produced is).
#include "schedule_io.hpp"

#include <cstring>

namespace race {

namespace {

constexpr char kMagic[8] = {'R', 'A', 'C', 'E', 'S', 'C', 'H', 'D'};
constexpr uint32_t kVersion = 1;

uint64_t fnv(const unsigned char* p, size_t n, uint64_t h = 1469598103934665603ull) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct Writer {
    std::vector<unsigned char> b;
    void raw(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    void i32(int v) { raw(&v, 4); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f64(double v) { raw(&v, 8); }
    void vec(const std::vector<int>& v) {
        i32((int)v.size());
        if (!v.empty()) raw(v.data(), v.size() * 4);
    }
};

struct Reader {
    const unsigned char* p;
    size_t n, at = 0;
    bool bad = false;
    bool raw(void* dst, size_t k) {
        if (bad || k > n - at) {
            bad = true;
            return false;
        }
        std::memcpy(dst, p + at, k);
        at += k;
        return true;
    }
    int i32() {
        int v = 0;
        raw(&v, 4);
        return v;
    }
    uint32_t u32() {
        uint32_t v = 0;
        raw(&v, 4);
        return v;
    }
    uint64_t u64() {
        uint64_t v = 0;
        raw(&v, 8);
        return v;
    }
    double f64() {
        double v = 0;
        raw(&v, 8);
        return v;
    }
    std::vector<int> vec() {
        const int c = i32();
        if (bad || c < 0 || (size_t)c > (n - at) / 4) {
            bad = true;
            return {};
        }
        std::vector<int> v((size_t)c);
        if (c) raw(v.data(), (size_t)c * 4);
        return v;
    }
};

bool rowPointers(const std::vector<int>& v, int n) {
    if (v.size() < 2 || v.front() != 0 || v.back() != n) return false;
    for (size_t i = 0; i + 1 < v.size(); ++i)
        if (v[i + 1] < v[i]) return false;
    return true;
}

// the checks flattenSchedule and the tree queries rely on: children come
// after their parent and belong to one (parent, sub-block) slot, so the
// recursion ends and visits every leaf once; row ranges lie in [0, n); the
// leaves reached cover every row exactly once
bool treeOk(const std::vector<ZoneLeaf>& t, int n, int bpt) {
    const int L = (int)t.size();
    if (L == 0) return true;
    if (t[0].effRowZ <= 0) return false;
    std::vector<char> owned((size_t)L, 0);
    std::vector<int> depth((size_t)L, 0);  // recursion depth of flattenSchedule
    for (int p = 0; p < L; ++p) {
        const ZoneLeaf& z = t[p];
        if (z.valueZ.size() < 2 || z.totalSubBlocks < 1 || z.children.size() % 2) return false;
        for (size_t i = 0; i < z.valueZ.size(); ++i) {
            if (z.valueZ[i] < 0 || z.valueZ[i] > n) return false;
            if (i && z.valueZ[i] < z.valueZ[i - 1]) return false;
        }
        // a leaf without children runs its range once: one sub-block
        if (!z.children.empty() && (int)z.children.size() < 2 * z.totalSubBlocks) return false;
        if (z.children.empty() && z.totalSubBlocks != 1) return false;
        for (size_t s = 0; s + 1 < z.children.size(); s += 2) {
            const int c0 = z.children[s], c1 = z.children[s + 1];
            if (c0 <= p || c1 < c0 || c1 > L) return false;
            if ((c1 - c0) / bpt <= 1) continue;  // a leaf call: children unused
            for (int c = c0; c < c0 + (c1 - c0) / bpt * bpt; ++c) {
                if (owned[c]) return false;
                owned[c] = 1;
                depth[c] = depth[p] + 1;
                if (depth[c] > 256) return false;
            }
        }
    }
    return true;
}

bool coversOnce(const Schedule& s, int n) {
    std::vector<char> seen((size_t)n, 0);
    long rows = 0;
    for (size_t l = 0; l < s.leafLo.size(); ++l) {
        for (int r = s.leafLo[l]; r < s.leafHi[l]; ++r) {
            if (r < 0 || r >= n || seen[r]) return false;
            seen[r] = 1;
        }
        rows += s.leafHi[l] - s.leafLo[l];
    }
    return rows == n;
}

}  // namespace

uint64_t patternHash(int n, const int* rowPtr, const int* col) {
    uint64_t h = fnv(reinterpret_cast<const unsigned char*>(&n), 4);
    h = fnv(reinterpret_cast<const unsigned char*>(rowPtr), (size_t)(n + 1) * 4, h);
    return fnv(reinterpret_cast<const unsigned char*>(col), (size_t)rowPtr[n] * 4, h);
}

std::vector<unsigned char> encodeSchedule(const SavedSchedule& s) {
    Writer w;
    w.raw(kMagic, 8);
    w.u32(kVersion);
    for (int v : {s.nrow, s.dist, s.d2, s.lbTarget, s.requested, s.available, s.highestPower, s.highestSubPower})
        w.i32(v);
    w.u64(s.pattern);
    w.vec(s.perm);
    w.vec(s.inv);
    w.i32(s.hasTree ? 1 : 0);
    if (s.hasTree) {
        w.i32((int)s.tree.size());
        for (const ZoneLeaf& z : s.tree) {
            for (int v : {z.nthreadsZ, z.idealNthreadsZ, z.totalSubBlocks, z.parentZ, z.effRowZ, z.stage,
                          z.reachedLimit ? 1 : 0})
                w.i32(v);
            w.vec(z.valueZ);
            w.vec(z.children);
        }
    }
    w.i32(s.hasMpk ? 1 : 0);
    if (s.hasMpk) {
        w.vec(s.mpk.levelPtr);
        w.vec(s.mpk.groupPtr);
        w.vec(s.mpk.groupLevel);
        w.vec(s.mpk.levelNnz);
        w.vec(s.mpk.hopeless);
        w.f64(s.mpk.cacheElem);
        w.f64(s.mpk.bytePerNnz);
        w.i32(s.mpk.highestPower);
    }
    w.u64(fnv(w.b.data(), w.b.size()));
    return w.b;
}

int decodeSchedule(const void* buf, size_t size, SavedSchedule* out) {
    if (!buf || !out || size < 8 + 4 + 8) return ERR_INVALID_ARG;
    const unsigned char* p = static_cast<const unsigned char*>(buf);
    uint64_t sum;
    std::memcpy(&sum, p + size - 8, 8);
    if (fnv(p, size - 8) != sum) return ERR_INVALID_ARG;
    Reader r{p, size - 8};
    char magic[8];
    r.raw(magic, 8);
    if (r.bad || std::memcmp(magic, kMagic, 8) || r.u32() != kVersion) return ERR_INVALID_ARG;
    SavedSchedule s;
    s.nrow = r.i32();
    s.dist = r.i32();
    s.d2 = r.i32();
    s.lbTarget = r.i32();
    s.requested = r.i32();
    s.available = r.i32();
    s.highestPower = r.i32();
    s.highestSubPower = r.i32();
    s.pattern = r.u64();
    s.perm = r.vec();
    s.inv = r.vec();
    s.hasTree = r.i32() != 0;
    if (s.hasTree) {
        const int L = r.i32();
        if (r.bad || L < 0 || (size_t)L > size / 36) return ERR_INVALID_ARG;
        s.tree.resize((size_t)L);
        for (ZoneLeaf& z : s.tree) {
            z.nthreadsZ = r.i32();
            z.idealNthreadsZ = r.i32();
            z.totalSubBlocks = r.i32();
            z.parentZ = r.i32();
            z.effRowZ = r.i32();
            z.stage = r.i32();
            z.reachedLimit = r.i32() != 0;
            z.valueZ = r.vec();
            z.children = r.vec();
            if (r.bad) return ERR_INVALID_ARG;
        }
    }
    s.hasMpk = r.i32() != 0;
    if (s.hasMpk) {
        s.mpk.levelPtr = r.vec();
        s.mpk.groupPtr = r.vec();
        s.mpk.groupLevel = r.vec();
        s.mpk.levelNnz = r.vec();
        s.mpk.hopeless = r.vec();
        s.mpk.cacheElem = r.f64();
        s.mpk.bytePerNnz = r.f64();
        s.mpk.highestPower = r.i32();
    }
    if (r.bad || r.at != r.n) return ERR_INVALID_ARG;

    // contents
    const int n = s.nrow;
    if (n <= 0 || s.dist < DIST_ONE || s.dist > DIST_POWER || s.requested < 1 || s.available < 1)
        return ERR_INVALID_ARG;
    if (s.highestPower < 1 || s.highestSubPower < 1) return ERR_INVALID_ARG;
    if ((int)s.perm.size() != n || (int)s.inv.size() != n) return ERR_INVALID_ARG;
    for (int i = 0; i < n; ++i)
        if (s.perm[i] < 0 || s.perm[i] >= n || s.inv[s.perm[i]] != i) return ERR_INVALID_ARG;
    if ((s.dist == DIST_POWER) != s.hasMpk || s.hasTree == s.hasMpk) return ERR_INVALID_ARG;
    if (s.hasTree) {
        const int bpt = blockPerThread(s.dist, s.d2);
        if (!treeOk(s.tree, n, bpt)) return ERR_INVALID_ARG;
        ZoneTree zt(s.dist, s.d2, s.lbTarget);
        zt.tree = s.tree;
        if (!coversOnce(flattenSchedule(zt, false, n), n)) return ERR_INVALID_ARG;
    }
    if (s.hasMpk) {
        const MPKPartition& m = s.mpk;
        if (!rowPointers(m.levelPtr, n) || !rowPointers(m.groupPtr, n)) return ERR_INVALID_ARG;
        // hopeless: the group of each hopeless region, ascending
        for (size_t i = 0; i < m.hopeless.size(); ++i)
            if (m.hopeless[i] < 0 || m.hopeless[i] + 1 >= (int)m.groupPtr.size() ||
                (i > 0 && m.hopeless[i] <= m.hopeless[i - 1]))
                return ERR_INVALID_ARG;
        // groupLevel: level group -> first BFS level, ending with the level count
        const int L = (int)m.levelNnz.size();
        if (m.levelPtr.size() != (size_t)L + 1 || m.groupLevel.size() != m.groupPtr.size() ||
            !rowPointers(m.groupLevel, L))
            return ERR_INVALID_ARG;
        for (size_t g = 0; g < m.groupPtr.size(); ++g)
            if (m.groupPtr[g] != m.levelPtr[m.groupLevel[g]]) return ERR_INVALID_ARG;
        if (m.highestPower < 1) return ERR_INVALID_ARG;
    }
    *out = std::move(s);
    return OK;
}

}  // namespace race
