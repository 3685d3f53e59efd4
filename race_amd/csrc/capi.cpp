// capi.cpp -- C ABI of the MI355X RACE engine (include/race_amd.h).
//
// A handle mirrors RACE::Interface (include/interface.h:54-205,
// src/interface.cpp): it owns the graph (copied to the device), the GPU
// traverser, the zone tree or MPK partition and the registered functions.
#include <hip/hip_runtime.h>
#include <omp.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/race_amd.h"
#include "bfs.hpp"
#include "simdify.hpp"
#include "engine.hpp"
#include "exec.hpp"
#include "graph_cache.hpp"
#include "schedule_io.hpp"

using namespace race;

namespace {

struct HostFn {
    race_color_fn color = nullptr;
    race_power_fn powerFn = nullptr;
    void* args = nullptr;
    int power = 1, subPower = 1;
    bool serial = false;  // setSerial: one host thread
};

struct Function {
    std::unique_ptr<HostFn> host;
    std::unique_ptr<DeviceFn> dev;
    std::unique_ptr<GraphCache> graph;  // with the "graph" option: captured launch sequences
    // attached halo-exchange callback (one of the two forms); one device has
    // no remote rows, so it is never called
    race_comm_fn comm = nullptr;
    race_comm_simple_fn commSimple = nullptr;
    void* commArgs = nullptr;
};

// Pattern symmetrisation: add (c,r) for every (r,c) without its pair
// (RACE::makeSymmetricGraph, src/graph.cpp:47-142).  Added columns are
// appended after the row's own entries.
void symmetrizePattern(int n, const int* rp, const int* col, std::vector<int>& orp,
                       std::vector<int>& ocol, bool* changed) {
    std::vector<std::vector<int>> extra(n);
    bool sym = true;
    for (int r = 0; r < n; ++r)
        for (int j = rp[r]; j < rp[r + 1]; ++j) {
            const int c = col[j];
            if (c == r || c < 0 || c >= n) continue;
            bool has = false;
            for (int k = rp[c]; k < rp[c + 1]; ++k)
                if (col[k] == r) {
                    has = true;
                    break;
                }
            if (!has) {
                extra[c].push_back(r);
                sym = false;
            }
        }
    *changed = !sym;
    if (sym) return;
    orp.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) orp[r + 1] = orp[r] + (rp[r + 1] - rp[r]) + (int)extra[r].size();
    ocol.resize(orp[n]);
    for (int r = 0; r < n; ++r) {
        int o = orp[r];
        for (int j = rp[r]; j < rp[r + 1]; ++j) ocol[o++] = col[j];
        for (int c : extra[r]) ocol[o++] = c;
    }
}

}  // namespace

struct race_handle_s {
    int nrow = 0, requested = 1, dist = DIST_ONE, d2 = TWO_BLOCK, lbTarget = LB_NNZ;
    // rows the colouring traverses: nrow minus the dense rows moved to the
    // serial tail (RACE_MAX_DENSE); graphRows[q] = original row of graph row
    // q, denseRows = original rows of the tail, in order
    int nGraph = 0;
    std::vector<int> graphRows, denseRows;
    int smt = 1, pin = 1;
    bool symmHint = false;
    std::vector<int> rowPtr, col;  // graph pattern used for traversal (copied)
    long nnzInput = 0;
    std::vector<int> initPerm, initInv;
    hipStream_t stream = nullptr;
    std::unique_ptr<GpuTraverser> trav;
    std::unique_ptr<ZoneTree> tree;
    int availableThreads = -1;
    bool colored = false;
    Schedule fwd, rev;
    MPKPartition mpk;
    int highestPower = 1, highestSubPower = 1;
    std::vector<int> perm, inv;
    std::vector<Function> funcs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int optLayout = 0, optBundleRows = 128, optBundleLocal = 6144, optBundleThreads = 256, optBundleTpr = 0, optXcd = 0;
    int optBfsBlockMax = 1 << 17, optSymmMode = 0, optGraph = 0, optFlow = 0;
    ~race_handle_s() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

static int checkHandle(race_handle h) { return h ? OK : ERR_INVALID_ARG; }

// A square CRS the device kernels can index: rowPtr[0] = 0, non-decreasing,
// every column in [0, nrow), values present.  Checked on the host before
// anything is laid out or uploaded.
static int checkCrs(const race_kernel_desc* d) {
    if (!d || d->nrow <= 0 || !d->rowPtr || !d->col || !d->val) return ERR_INVALID_ARG;
    if (d->rowPtr[0] != 0) return ERR_INVALID_ARG;
    for (int r = 0; r < d->nrow; ++r)
        if (d->rowPtr[r + 1] < d->rowPtr[r]) return ERR_INVALID_ARG;
    const int nnz = d->rowPtr[d->nrow];
    for (int j = 0; j < nnz; ++j)
        if (d->col[j] < 0 || d->col[j] >= d->nrow) return ERR_INVALID_ARG;
    // Gauss-Seidel divides by the first entry of a row and SymmSpMV takes it
    // as the diagonal (GS_KERNEL / SYMM_SPMV_KERNEL_BODY,
    // examples/src/helpers/kernels.cpp:174-199, 403-427): every row must
    // hold its diagonal first
    if (d->kind == RACE_K_GS || d->kind == RACE_K_SYMMSPMV)
        for (int r = 0; r < d->nrow; ++r)
            if (d->rowPtr[r + 1] == d->rowPtr[r] || d->col[d->rowPtr[r]] != r) return ERR_INVALID_ARG;
    return OK;
}

// A batch of functions on many matrices (race_batch_*): host copies of each
// matrix, its handle (colour schedules) and vectors, the layout options and
// the built fused device function.
struct race_batch_s {
    int kind = 0, valPrec = RACE_FP64, vecPrec = RACE_FP64, nrhs = 1;
    struct Mat {
        race_handle h = nullptr;
        std::vector<int> rp, col;
        std::vector<double> val;
        void* x = nullptr;
        void* b = nullptr;
    };
    std::vector<Mat> mats;
    ExecContext ctx;
    std::unique_ptr<DeviceFn> fn;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~race_batch_s() {
        fn.reset();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// the tuner's candidate layouts (12 per kernel family):
//   GS / KACZ bundles: (workgroup threads, rows per bundle, lanes per row;
//                       0 = from the longest row)
//   SpMV / SymmSpMV SELL: (rows per chunk C, sorting window sigma)
static const int kBundleCand[12][3] = {{512, 256, 0}, {256, 256, 0}, {1024, 256, 0}, {256, 64, 0},
                                       {512, 128, 0}, {128, 128, 0}, {512, 256, 1},  {512, 256, 2},
                                       {512, 256, 4}, {256, 128, 2}, {256, 128, 4},  {1024, 512, 2}};
static const int kSellCand[12][2] = {{64, 1},   {32, 1},  {16, 1},  {8, 1},  {64, 32}, {32, 32},
                                     {16, 32},  {8, 32},  {64, 1024}, {32, 1024}, {16, 1024}, {8, 1024}};

static bool bundleKind(int k) { return k == RACE_K_GS || k == RACE_K_KACZ; }

static void applyCandidate(race_batch b, int i) {
    if (bundleKind(b->kind)) {
        b->ctx.bundleThreads = kBundleCand[i][0];
        b->ctx.bundleRows = kBundleCand[i][1];
        b->ctx.bundleTpr = kBundleCand[i][2];
    } else {
        b->ctx.sellC = kSellCand[i][0];
        b->ctx.sellSigma = kSellCand[i][1];
    }
}

static int buildBatch(race_batch b) {
    b->fn.reset();
    std::vector<BatchItem> items(b->mats.size());
    for (size_t m = 0; m < b->mats.size(); ++m) {
        race_batch_s::Mat& M = b->mats[m];
        BatchItem& it = items[m];
        it.fwd = &M.h->fwd;
        it.rev = &M.h->rev;
        it.d.kind = b->kind;
        it.d.nrow = (int)M.rp.size() - 1;
        it.d.rowPtr = M.rp.data();
        it.d.col = M.col.data();
        it.d.val = M.val.data();
        it.d.val_precision = b->valPrec;
        it.d.vec_precision = b->vecPrec;
        it.d.nrhs = b->nrhs;
        it.d.x = M.x;
        it.d.b = M.b;
    }
    return bundleKind(b->kind) ? makeBundleBatchFn(items, b->ctx, &b->fn) : makeSellBatchFn(items, b->ctx, &b->fn);
}

extern "C" {

const char* race_error_string(int s) {
    switch (s) {
        case RACE_SUCCESS: return "RACE_SUCCESS";
        case RACE_ERR_INVALID_ARG: return "RACE_ERR_INVALID_ARG";
        case RACE_ERR_MATRIX_SYMM: return "RACE_ERR_MATRIX_SYMM";
        case RACE_ERR_D2_COLOR: return "RACE_ERR_D2_COLOR";
        case RACE_ERR_D1_COLOR: return "RACE_ERR_D1_COLOR";
        case RACE_ERR_GRAPH_TRAVERSAL: return "RACE_ERR_GRAPH_TRAVERSAL";
        case RACE_ERR_INCOMPATIBILITY: return "RACE_ERR_INCOMPATIBILITY";
        case RACE_ERR_HWLOC: return "RACE_ERR_HWLOC";
        case RACE_ERR_PIN: return "RACE_ERR_PIN";
        case RACE_ERR_INTERNAL: return "RACE_ERR_INTERNAL";
        case RACE_ERR_NOT_IMPLEMENTED: return "RACE_ERR_NOT_IMPLEMENTED";
        default: return "RACE_ERR_UNKNOWN";
    }
}

int race_create(race_handle* out, int nrow, int nthreads, int dist, const int* rowPtr, const int* col,
                int symm_hint, int smt, int pin_method, const int* init_perm, const int* init_inv_perm,
                int d2_method, int lb_target) {
    if (!out || nrow <= 0 || !rowPtr || !col || nthreads <= 0) return ERR_INVALID_ARG;
    if (dist != DIST_ONE && dist != DIST_TWO && dist != DIST_POWER) return ERR_INVALID_ARG;
    if ((init_perm == nullptr) != (init_inv_perm == nullptr)) return ERR_INVALID_ARG;
    auto h = std::make_unique<race_handle_s>();
    h->nrow = nrow;
    h->requested = nthreads;
    h->dist = dist;
    h->d2 = d2_method;
    h->lbTarget = lb_target;
    h->smt = smt;
    h->pin = pin_method;
    h->symmHint = symm_hint != 0;
    h->nnzInput = rowPtr[nrow];
    // RACE_MAX_RECURSION_STAGES=0 lets a POWER graph skip symmetrisation
    // (src/graph_SoA.cpp:441-460)
    int maxRec = 10;
    if (dist == DIST_POWER) {
        std::vector<int> v = envIntList("RACE_MAX_RECURSION_STAGES");
        if (!v.empty()) maxRec = v[0];
    }
    bool changed = false;
    if (maxRec > 0 && !h->symmHint) symmetrizePattern(nrow, rowPtr, col, h->rowPtr, h->col, &changed);
    if (!changed) {
        h->rowPtr.assign(rowPtr, rowPtr + nrow + 1);
        h->col.assign(col, col + rowPtr[nrow]);
    }
    if (init_perm) {
        h->initPerm.assign(init_perm, init_perm + nrow);
        h->initInv.assign(init_inv_perm, init_inv_perm + nrow);
    }
    h->nGraph = nrow;
    // Dense rows (RACE_MAX_DENSE, Graph::getStatistics /
    // permuteAndRemoveSerialPart src/graph_SoA.cpp:321-435, 173-200): rows
    // longer than RACE_MAX_DENSE x (nnz / nrow, integer division) go to a
    // serial tail after the coloured rows, in the row order of the initial
    // permutation, and the graph loses them and every entry pointing to them.
    // This follows the reference built without RACE_PERMUTE_ON_FLY; with it
    // (its default) the reference reads past its arrays (it skips columns by
    // their unpermuted index, src/traverse_GAP_graphSoA.cpp:161, and warns
    // that the removal is not implemented, src/graph_SoA.cpp:389).
    if (const char* md = std::getenv("RACE_MAX_DENSE")) {
        if (dist == DIST_POWER) return ERR_NOT_IMPLEMENTED;  // the power executor has no serial part
        const double maxDense = std::atof(md);
        const int mean = (int)(h->nnzInput / nrow);
        std::vector<int> cur(nrow), newId(nrow, -1);
        for (int r = 0; r < nrow; ++r) cur[r] = init_perm ? init_perm[r] : r;
        for (int r = 0; r < nrow; ++r) {
            const int o = cur[r];
            if (h->rowPtr[o + 1] - h->rowPtr[o] > maxDense * mean) {
                h->denseRows.push_back(o);
            } else {
                newId[o] = (int)h->graphRows.size();
                h->graphRows.push_back(o);
            }
        }
        if (!h->denseRows.empty()) {
            std::vector<int> rp(1, 0), cl;
            for (int o : h->graphRows) {
                for (int j = h->rowPtr[o]; j < h->rowPtr[o + 1]; ++j)
                    if (newId[h->col[j]] >= 0) cl.push_back(newId[h->col[j]]);
                rp.push_back((int)cl.size());
            }
            h->rowPtr.swap(rp);
            h->col.swap(cl);
            h->nGraph = (int)h->graphRows.size();
            h->initPerm.clear();  // applied: graph row q is original row graphRows[q]
            h->initInv.clear();
        } else {
            h->graphRows.clear();
        }
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return ERR_INTERNAL;
    *out = h.release();
    return OK;
}

int race_destroy(race_handle h) {
    if (!h) return ERR_INVALID_ARG;
    hipStream_t s = h->stream;
    if (s) (void)hipStreamSynchronize(s);
    delete h;
    if (s) (void)hipStreamDestroy(s);
    return OK;
}

// Full-length permutation (graph rows, then the dense tail) and the colour
// schedules, with the tail as one last phase in both directions: the
// reference calls the serial part after the coloured part, f(n', n) forward
// and reverse (FuncManager::Run src/functionManager.cpp:1493-1500).
static void appendSerialPhase(race_handle h, Schedule* s) {
    if (h->nGraph >= h->nrow) return;
    s->leafLo.push_back(h->nGraph);
    s->leafHi.push_back(h->nrow);
    s->phasePtr.push_back((int)s->leafLo.size());
}

static void setPermAndSchedules(race_handle h) {
    const std::vector<int>& tp = h->trav->perm();
    if (h->graphRows.empty()) {
        h->perm = tp;
        h->inv = h->trav->invPerm();
    } else {
        h->perm.resize(h->nrow);
        for (int i = 0; i < h->nGraph; ++i) h->perm[i] = h->graphRows[tp[i]];
        for (size_t d = 0; d < h->denseRows.size(); ++d) h->perm[h->nGraph + d] = h->denseRows[d];
        h->inv.resize(h->nrow);
        for (int i = 0; i < h->nrow; ++i) h->inv[h->perm[i]] = i;
    }
    h->fwd = flattenSchedule(*h->tree, false, h->nGraph);
    h->rev = flattenSchedule(*h->tree, true, h->nGraph);
    appendSerialPhase(h, &h->fwd);
    appendSerialPhase(h, &h->rev);
}

static int ensureTraverser(race_handle h) {
    if (h->trav) return OK;
    h->trav = std::make_unique<GpuTraverser>();
    h->trav->rangeBlockMax = h->optBfsBlockMax;
    return h->trav->init(h->nGraph, h->rowPtr.data(), h->col.data(),
                         h->initPerm.empty() ? nullptr : h->initPerm.data(),
                         h->initInv.empty() ? nullptr : h->initInv.data(), h->stream);
}

int race_color(race_handle h) {
    if (checkHandle(h)) return ERR_INVALID_ARG;
    if (h->dist == DIST_POWER) return ERR_INVALID_ARG;  // src/interface.cpp:263-267
    int e = ensureTraverser(h);
    if (e) return e;
    std::vector<int> effI = envIntList("RACE_EFFICIENCY");
    std::vector<double> eff(effI.begin(), effI.end());
    std::vector<int> lvlThreads = envIntList("RACE_THREADS");
    h->tree = std::make_unique<ZoneTree>(h->dist, h->d2, h->lbTarget);
    LevelRecursion lr(h->trav.get(), h->nGraph, h->requested, h->dist, h->d2, h->lbTarget, eff, lvlThreads);
    e = lr.levelBalancing(h->tree.get());
    if (e) return e;
    h->availableThreads = lr.availableThreads;
    setPermAndSchedules(h);
    h->colored = true;
    return OK;
}

int race_color_power(race_handle h, int highest_power, double cache_size, double safety,
                     char mtx_type, int highest_sub_power) {
    if (checkHandle(h)) return ERR_INVALID_ARG;
    if (highest_sub_power < 1 || highest_power < 1) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER) return ERR_INVALID_ARG;  // src/interface.cpp:220-224
    // "N" (BFS levels), "L" / "U" (row order / reversed row order as levels,
    // src/traverse_GAP_graphSoA.cpp:316-335); any other matrix type is an
    // error in the reference (src/matrixPower.cpp:55-66)
    if (mtx_type != 'N' && mtx_type != 'L' && mtx_type != 'U') return ERR_INVALID_ARG;
    h->highestPower = highest_power;
    h->highestSubPower = highest_sub_power;
    int e = ensureTraverser(h);
    if (e) return e;
    // cache-violation cut-off of stage 0 (src/matrixPowerRecursive.cpp:37-74,88-109)
    std::vector<int> def = envIntList("RACE_CACHE_VIOLATION_CUTOFF_DEFAULT");
    std::vector<int> cut{def.empty() ? 50 : def[0]};
    for (int v : envIntList("RACE_CACHE_VIOLATION_CUTOFF")) cut.push_back(v);
    std::vector<int> mr = envIntList("RACE_MAX_RECURSION_STAGES");
    int maxRec = mr.empty() ? std::max(0, (int)cut.size()) : mr[0];
    // mtxPowerRecursive::get_cache_violation_cutoff: stage s uses entry s+1
    // of the list (the default when the list is shorter), no cut-off from
    // stage maxRec on
    auto cutoffOf = [&](int stage) {
        if (stage >= maxRec) return 0x7fffffff;
        return stage < (int)cut.size() - 1 ? cut[stage + 1] : cut[0];
    };
    e = mpkPartition(h->trav.get(), h->nrow, h->nnzInput, highest_power * highest_sub_power,
                     cache_size, safety, cutoffOf(0), cutoffOf(1), mtx_type, &h->mpk);
    if (e == ERR_NOT_IMPLEMENTED)
        std::fprintf(stderr,
                     "race_amd: hopeless levels inside a recursion stage (or with mtxType L/U): a second "
                     "recursion stage is not implemented\n");
    if (e) return e;
    h->perm = h->trav->perm();
    h->inv = h->trav->invPerm();
    h->availableThreads = h->requested;
    h->colored = true;
    return OK;
}

int race_get_perm(race_handle h, int* perm, int len) {
    if (checkHandle(h) || !perm || len != h->nrow || h->perm.empty()) return ERR_INVALID_ARG;
    std::memcpy(perm, h->perm.data(), sizeof(int) * len);
    return OK;
}

int race_get_inv_perm(race_handle h, int* inv, int len) {
    if (checkHandle(h) || !inv || len != h->nrow || h->inv.empty()) return ERR_INVALID_ARG;
    std::memcpy(inv, h->inv.data(), sizeof(int) * len);
    return OK;
}

int race_get_num_threads(race_handle h, int* out) {
    if (checkHandle(h) || !out) return ERR_INVALID_ARG;
    *out = h->availableThreads;
    return OK;
}

int race_get_efficiency(race_handle h, double* out) {
    if (checkHandle(h) || !out || !h->colored) return ERR_INVALID_ARG;
    if (h->dist == DIST_POWER) {
        *out = 1.0;  // single device: every row is a main row
    } else {
        // (nrow / effRow) / requested threads (src/interface.cpp:304-321)
        *out = ((double)h->nrow / (double)h->tree->tree[0].effRowZ) / h->requested;
    }
    return OK;
}

int race_get_max_stage_depth(race_handle h, int* out) {
    if (checkHandle(h) || !out) return ERR_INVALID_ARG;
    *out = h->tree ? h->tree->maxStage() : 1;
    return OK;
}

int race_get_highest_power(race_handle h, int* out) {
    if (checkHandle(h) || !out) return ERR_INVALID_ARG;
    *out = h->highestPower;
    return OK;
}

int race_register_function(race_handle h, race_color_fn f, void* args, int* id) {
    if (checkHandle(h) || !f || !id) return ERR_INVALID_ARG;
    if (h->dist == DIST_POWER || !h->colored) return ERR_INVALID_ARG;
    Function fn;
    fn.host = std::make_unique<HostFn>();
    fn.host->color = f;
    fn.host->args = args;
    h->funcs.push_back(std::move(fn));
    *id = (int)h->funcs.size() - 1;
    return OK;
}

int race_register_power_function(race_handle h, race_power_fn f, void* args, int power,
                                 int sub_power, int numa_split, int* id) {
    (void)numa_split;
    if (checkHandle(h) || !f || !id) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER || !h->colored) return ERR_INVALID_ARG;
    if (power < 1 || sub_power < 1) return ERR_INVALID_ARG;
    Function fn;
    fn.host = std::make_unique<HostFn>();
    fn.host->powerFn = f;
    fn.host->args = args;
    fn.host->power = power;
    fn.host->subPower = sub_power;
    h->funcs.push_back(std::move(fn));
    *id = (int)h->funcs.size() - 1;
    return OK;
}

int race_attach_communication(race_handle h, int id, race_comm_fn f, void* args) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !f) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER) return ERR_NOT_IMPLEMENTED;  // src/interface.cpp:489-494
    Function& fn = h->funcs[id];
    fn.comm = f;
    fn.commSimple = nullptr;
    fn.commArgs = args;
    return OK;
}

int race_attach_communication_simple(race_handle h, int id, race_comm_simple_fn f, void* args) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !f) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER) return ERR_NOT_IMPLEMENTED;
    if (h->highestSubPower != 1) return ERR_INCOMPATIBILITY;  // src/interface.cpp:521-531
    Function& fn = h->funcs[id];
    fn.comm = nullptr;
    fn.commSimple = f;
    fn.commArgs = args;
    return OK;
}

int race_register_device(race_handle h, const race_kernel_desc* d, int* id) {
    if (checkHandle(h) || !d || !id) return ERR_INVALID_ARG;
    if (d->nrow != h->nrow || checkCrs(d)) return ERR_INVALID_ARG;
    const bool power = d->kind == RACE_K_MPK || d->kind == RACE_K_MPK_HIGHEST;
    // SpMV and SpMTV need no schedule: they run on any handle
    const bool free = d->kind == RACE_K_SPMV || d->kind == RACE_K_SPMTV;
    if (power != (h->dist == DIST_POWER) && !free) return ERR_INVALID_ARG;
    if (!free && !h->colored) return ERR_INVALID_ARG;
    // power > highestPower is allowed, as in the reference (which only warns,
    // src/interface.cpp:449-452): the wavefront is correct for any power,
    // only the cache blocking was sized for highestPower
    if (power && d->power < 1) return ERR_INVALID_ARG;
    // A kernel that writes the columns of its rows (Kaczmarz, colour-phase
    // SymmSpMV) relies on distance-2 independence of concurrent rows; rows
    // removed as dense are no graph vertices, so two concurrent rows may
    // share a dense column and write it at once (the reference's OpenMP
    // execution races there the same way).  Refused instead.
    if (h->nGraph < h->nrow &&
        (d->kind == RACE_K_KACZ || (d->kind == RACE_K_SYMMSPMV && h->optSymmMode == 1)))
        return ERR_INCOMPATIBILITY;
    ExecContext ctx;
    ctx.fwd = &h->fwd;
    ctx.rev = &h->rev;
    ctx.mpk = &h->mpk;
    ctx.layout = h->optLayout;
    ctx.bundleRows = h->optBundleRows;
    ctx.bundleLocal = h->optBundleLocal;
    ctx.xcdOrder = h->optXcd;
    ctx.bundleThreads = h->optBundleThreads;
    ctx.bundleTpr = h->optBundleTpr;
    ctx.symmMode = h->optSymmMode;
    ctx.flow = h->optFlow;
    Function fn;
    int e = makeDeviceFn(*d, ctx, &fn.dev);
    if (e) return e;
    h->funcs.push_back(std::move(fn));
    *id = (int)h->funcs.size() - 1;
    return OK;
}

int race_set_option(race_handle h, const char* key, int value) {
    if (checkHandle(h) || !key) return ERR_INVALID_ARG;
    const std::string k(key);
    if (k == "layout") {
        if (value < 0 || value > 2) return ERR_INVALID_ARG;
        h->optLayout = value;
    } else if (k == "bundle_rows") {
        if (value < 1 || value > 8192) return ERR_INVALID_ARG;
        h->optBundleRows = value;
    } else if (k == "bundle_threads") {
        if (value != 64 && value != 128 && value != 256 && value != 512 && value != 1024) return ERR_INVALID_ARG;
        h->optBundleThreads = value;
    } else if (k == "bundle_tpr") {
        if (value != 0 && value != 1 && value != 2 && value != 4) return ERR_INVALID_ARG;
        h->optBundleTpr = value;
    } else if (k == "bundle_local") {
        if (value < 1 || value > 65535) return ERR_INVALID_ARG;
        h->optBundleLocal = value;
    } else if (k == "xcd_order") {
        if (value != 0 && value != 1) return ERR_INVALID_ARG;
        h->optXcd = value;
    } else if (k == "symm_mode") {
        if (value < 0 || value > 3) return ERR_INVALID_ARG;
        h->optSymmMode = value;
    } else if (k == "graph") {
        if (value != 0 && value != 1) return ERR_INVALID_ARG;
        h->optGraph = value;
    } else if (k == "flow") {
        if (value != 0 && value != 1) return ERR_INVALID_ARG;
        h->optFlow = value;
    } else if (k == "bfs_block_max") {
        if (value < 0) return ERR_INVALID_ARG;
        h->optBfsBlockMax = value;
        if (h->trav) h->trav->rangeBlockMax = value;
    } else {
        return ERR_INVALID_ARG;
    }
    return OK;
}

int race_register_refinement(race_handle h, const race_kernel_desc* a64, int inner_id, int inner_sweeps,
                             int* id) {
    if (checkHandle(h) || !a64 || !id || inner_id < 0 || inner_id >= (int)h->funcs.size())
        return ERR_INVALID_ARG;
    if (!h->funcs[inner_id].dev || a64->nrow != h->nrow || checkCrs(a64)) return ERR_INVALID_ARG;
    Function fn;
    int e = makeRefineFn(*a64, h->funcs[inner_id].dev.get(), inner_sweeps, &fn.dev);
    if (e) return e;
    h->funcs.push_back(std::move(fn));
    *id = (int)h->funcs.size() - 1;
    return OK;
}

int race_refine(race_handle h, int id, double tol, int max_outer, int* outer, long* sweeps,
                double* relres) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    return refineSolve(h->funcs[id].dev.get(), tol, max_outer, h->stream, outer, sweeps, relres);
}

int race_refine_noalloc(race_handle h, int id, double tol, int max_outer, int* outer, long* sweeps,
                        double* relres) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    return refineSolveStatic(h->funcs[id].dev.get(), tol, max_outer, h->stream, outer, sweeps, relres);
}

int race_refine_noalloc_max_bytes(size_t* bytes) {
    if (!bytes) return ERR_INVALID_ARG;
    *bytes = refineStaticScratchBytes();
    return OK;
}

int race_residual(race_handle h, int id, double* relres) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !h->funcs[id].dev || !relres)
        return ERR_INVALID_ARG;
    return refineResidual(h->funcs[id].dev.get(), h->stream, relres);
}

int race_set_vectors(race_handle h, int id, void* x, void* b) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    h->funcs[id].dev->x = x;
    h->funcs[id].dev->b = b;
    return OK;
}

// OpenMP threads for host callbacks: RACE's thread count (the reference runs
// one OpenMP thread per RACE thread), at most the host's, 1 when serial.
static int hostThreads(race_handle h, const HostFn& f) {
    if (f.serial) return 1;
    const int avail = h->availableThreads > 0 ? h->availableThreads : h->requested;
    return std::max(1, std::min(avail, omp_get_max_threads()));
}

// Host colouring callbacks run phase by phase, the leaves of one phase in
// parallel on OpenMP threads (they are distance-k independent, the
// concurrency of recursiveCall src/functionManager.cpp:199-265), with the
// implicit barrier of the worksharing loop between phases.
static void runHostColor(race_handle h, HostFn& f, bool rev) {
    const Schedule& s = rev ? h->rev : h->fwd;
    const int nt = hostThreads(h, f);
#pragma omp parallel num_threads(nt) if (nt > 1)
    for (int p = 0; p < s.numPhases(); ++p) {
#pragma omp for schedule(dynamic, 1)
        for (int l = s.phasePtr[p]; l < s.phasePtr[p + 1]; ++l) {
            if (rev)
                f.color(s.leafHi[l] - 1, s.leafLo[l] - 1, f.args);
            else
                f.color(s.leafLo[l], s.leafHi[l], f.args);
        }
    }
    // the serial part is the schedule's last phase when dense rows were
    // split off; otherwise the reference's empty call f(n, n)
    if (h->nGraph >= h->nrow) f.color(h->nrow, h->nrow, f.args);
}

// Host power callbacks over the level groups in a slope-2 wavefront: at step
// L, product t (= (pow-1) subPower + subPow, 1-based) runs on group
// L - 2 (t-1).  Product t of group g reads product t-1 of groups g-1..g+1,
// finished at steps <= L-1, and two units of one step are two or more groups
// apart, so the units of a step are independent; the rows of a unit are cut
// into slices for the threads (the reference's per-thread slices of a level,
// FuncManager::powerCallGeneral src/functionManager.cpp:618-822).
static void runHostPower(race_handle h, HostFn& f) {
    const std::vector<int>& g = h->mpk.groupPtr;
    const int G = (int)g.size() - 1;
    const int k = f.power * f.subPower;
    const int nt = hostThreads(h, f);
    struct Slice {
        int lo, hi, t;
    };
    std::vector<Slice> work;
    for (int L = 0; L < G + 2 * (k - 1); ++L) {
        work.clear();
        for (int t = 1; t <= k; ++t) {
            const int grp = L - 2 * (t - 1);
            if (grp < 0 || grp >= G) continue;
            const int lo = g[grp], hi = g[grp + 1];
            const int parts = std::max(1, std::min(nt, hi - lo));
            for (int q = 0; q < parts; ++q)
                work.push_back({lo + (int)((long)(hi - lo) * q / parts), lo + (int)((long)(hi - lo) * (q + 1) / parts),
                                t});
        }
        const int nw = (int)work.size();
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1 && nw > 1)
        for (int w = 0; w < nw; ++w) {
            const int t = work[w].t - 1;
            f.powerFn(work[w].lo, work[w].hi, t / f.subPower + 1, t % f.subPower + 1, 0, f.args);
        }
    }
}

int race_execute_async(race_handle h, int id, int rev, void* stream) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size()) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    if (fn.host) {
        if (fn.host->color)
            runHostColor(h, *fn.host, rev != 0);
        else
            runHostPower(h, *fn.host);
        return OK;
    }
    hipStream_t st = stream ? (hipStream_t)stream : h->stream;
    if (h->optGraph && fn.dev->graphable()) {
        if (!fn.graph) fn.graph = std::make_unique<GraphCache>();
        return fn.graph->run(fn.dev.get(), rev != 0, st);
    }
    return fn.dev->run(rev != 0, st);
}

int race_execute(race_handle h, int id, int rev) {
    int e = race_execute_async(h, id, rev, nullptr);
    if (e) return e;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return ERR_INTERNAL;
    return OK;
}

int race_set_power(race_handle h, int id, int power) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || power < 1) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    if (fn.host) {
        fn.host->power = power;
    } else {
        // a device MPK holds vector slots for the registered products only
        if ((long)power * fn.dev->subPower > fn.dev->powerMax) return ERR_INVALID_ARG;
        fn.dev->power = power;
    }
    return OK;
}

int race_set_sub_power(race_handle h, int id, int sub_power) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || sub_power < 1) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    if (fn.host) {
        fn.host->subPower = sub_power;
    } else {
        if ((long)fn.dev->power * sub_power > fn.dev->powerMax) return ERR_INVALID_ARG;
        fn.dev->subPower = sub_power;
    }
    return OK;
}

int race_get_sub_power(race_handle h, int id, int* sub_power) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !sub_power) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    *sub_power = fn.host ? fn.host->subPower : fn.dev->subPower;
    return OK;
}

int race_set_serial(race_handle h, int id, int serial) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size()) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    if (fn.host) fn.host->serial = serial != 0;  // device functions have no host threads
    return OK;
}

int race_reset_time(race_handle h) { return checkHandle(h) ? ERR_INVALID_ARG : OK; }

int race_sleep(race_handle h) { return checkHandle(h) ? ERR_INVALID_ARG : OK; }

int race_pin_thread(race_handle h, int thread_id) {
    if (checkHandle(h) || thread_id < 0) return ERR_INVALID_ARG;
    cpu_set_t allowed;
    CPU_ZERO(&allowed);
    if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return ERR_PIN;
    const int ncpu = CPU_COUNT(&allowed);
    if (ncpu < 1) return ERR_PIN;
    // thread i goes to the (i * smt mod ncpu)-th CPU the process may use
    // (RACE's FILL order over its allowed cores, one hardware thread per core
    // with smt > 1)
    int target = (int)(((long)thread_id * std::max(1, h->smt)) % ncpu), seen = 0, cpu = -1;
    for (int c = 0; c < CPU_SETSIZE; ++c)
        if (CPU_ISSET(c, &allowed) && seen++ == target) {
            cpu = c;
            break;
        }
    if (cpu < 0) return ERR_PIN;
    cpu_set_t one;
    CPU_ZERO(&one);
    CPU_SET(cpu, &one);
    return sched_setaffinity(0, sizeof(one), &one) == 0 ? OK : ERR_PIN;
}

// The reference's first-touch initialisers (src/interface.cpp:852-1040): a
// power callback registered with power 1 and executed once, so every row
// range is written by the thread that will later compute it.
namespace {
struct NumaArgs {
    int nrow;
    int* rowPtr;
    int* col;
    double* val;
    double* x;
    int** rowPtrL;
    int** colL;
    double** valL;
};
void numaRowPtrFn(int start, int end, int pow, int subPow, int, void* a) {
    NumaArgs* n = (NumaArgs*)a;
    if (pow == 1 && subPow == 1)
        for (int r = start; r < end; ++r) n->rowPtr[r] = n->rowPtr[r + 1] = 0;
}
void numaRowPtrLocalFn(int start, int end, int pow, int subPow, int numa, void* a) {
    NumaArgs* n = (NumaArgs*)a;
    if (pow == 1 && subPow == 1)
        for (int r = start; r < end; ++r) n->rowPtrL[numa][r] = 0;
}
void numaMtxVecFn(int start, int end, int pow, int subPow, int, void* a) {
    NumaArgs* n = (NumaArgs*)a;
    for (int r = start; r < end; ++r) {
        if (pow == 1 && subPow == 1)
            for (int j = n->rowPtr[r]; j < n->rowPtr[r + 1]; ++j) {
                n->val[j] = 0;
                n->col[j] = 0;
            }
        if (n->x) {
            n->x[(long)pow * n->nrow + r] = 0;
            n->x[(long)(pow - 1) * n->nrow + r] = 0;
        }
    }
}
void numaMtxVecLocalFn(int start, int end, int pow, int subPow, int numa, void* a) {
    NumaArgs* n = (NumaArgs*)a;
    if (pow == 1 && subPow == 1)
        for (int r = start; r < end; ++r)
            for (int j = n->rowPtrL[numa][r]; j < n->rowPtrL[numa][r + 1]; ++j) {
                n->valL[numa][j] = 0;
                n->colL[numa][j] = 0;
            }
}
}  // namespace

static int runNumaInit(race_handle h, race_power_fn fn, NumaArgs* a) {
    if (checkHandle(h)) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER) return ERR_NOT_IMPLEMENTED;  // "NUMA not implemented for coloring"
    int id = -1;
    int e = race_register_power_function(h, fn, a, 1, 1, 0, &id);
    if (e) return e;
    return race_execute(h, id, 0);
}

int race_numa_init_row_ptr(race_handle h, int* rowPtr) {
    if (!rowPtr) return ERR_INVALID_ARG;
    NumaArgs a{h ? h->nrow : 0, rowPtr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    return runNumaInit(h, numaRowPtrFn, &a);
}

int race_numa_init_row_ptr_local(race_handle h, int** rowPtr) {
    if (!rowPtr) return ERR_INVALID_ARG;
    NumaArgs a{h ? h->nrow : 0, nullptr, nullptr, nullptr, nullptr, rowPtr, nullptr, nullptr};
    return runNumaInit(h, numaRowPtrLocalFn, &a);
}

int race_numa_init_mtx_vec(race_handle h, int* rowPtr, int* col, double* val, double* x, int power) {
    (void)power;  // unused by the reference as well
    if (!rowPtr || !col || !val) return ERR_INVALID_ARG;
    NumaArgs a{h ? h->nrow : 0, rowPtr, col, val, x, nullptr, nullptr, nullptr};
    return runNumaInit(h, numaMtxVecFn, &a);
}

int race_numa_init_mtx_vec_local(race_handle h, int** rowPtr, int** col, double** val, int power) {
    (void)power;
    if (!rowPtr || !col || !val) return ERR_INVALID_ARG;
    NumaArgs a{h ? h->nrow : 0, nullptr, nullptr, nullptr, nullptr, rowPtr, col, val};
    return runNumaInit(h, numaMtxVecLocalFn, &a);
}

int race_get_numa_splitting(race_handle h, int* split, int cap, int* len) {
    if (checkHandle(h) || !len) return ERR_INVALID_ARG;
    if (h->dist != DIST_POWER || !h->colored) return ERR_INVALID_ARG;
    // one device, one domain: the main rows [0, nrow) (src/interface.cpp:1042-1052)
    *len = 2;
    if (split) {
        if (cap < 2) return ERR_INVALID_ARG;
        split[0] = h->mpk.levelPtr.empty() ? 0 : h->mpk.levelPtr.front();
        split[1] = h->mpk.levelPtr.empty() ? h->nrow : h->mpk.levelPtr.back();
    }
    return OK;
}

int race_get_power(race_handle h, int id, int* power) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !power) return ERR_INVALID_ARG;
    Function& fn = h->funcs[id];
    *power = fn.host ? fn.host->power : fn.dev->power;
    return OK;
}

// time per power for p = 1..highestPower, keep the fastest
// (Interface::tuneFunction src/interface.cpp:605-647)
int race_tune_function(race_handle h, int id, int rev, int* best_power) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !best_power) return ERR_INVALID_ARG;
    double bestT = 1e300;
    int best = 1;
    const Function& fn = h->funcs[id];
    const int pmax = fn.dev ? std::min(h->highestPower, fn.dev->powerMax / fn.dev->subPower) : h->highestPower;
    for (int p = 1; p <= pmax; ++p) {
        int e0 = race_set_power(h, id, p);
        if (e0) return e0;
        auto t0 = std::chrono::steady_clock::now();
        double t = 0;
        int it = 0;
        while (t < 0.2) {
            int e = race_execute(h, id, rev);
            if (e) return e;
            t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            ++it;
        }
        double per = t / (it * (double)p);
        if (per < bestT) {
            bestT = per;
            best = p;
        }
    }
    race_set_power(h, id, best);
    *best_power = best;
    return OK;
}

static void coverageFn(int start, int end, int pow, int, int, void* args) {
    std::vector<int>* cnt = (std::vector<int>*)args;
    for (int r = start; r < end; ++r) (*cnt)[r] += 1;
    (void)pow;
}

int race_check_power_coverage(race_handle h) {
    if (checkHandle(h) || h->dist != DIST_POWER || !h->colored) return ERR_INVALID_ARG;
    std::vector<int> cnt(h->nrow, 0);
    HostFn f;
    f.powerFn = coverageFn;
    f.args = &cnt;
    f.power = h->highestPower;
    runHostPower(h, f);
    for (int r = 0; r < h->nrow; ++r)
        if (cnt[r] != h->highestPower) return ERR_INTERNAL;
    return OK;
}

int race_get_schedule(race_handle h, int rev, int* nphases, int* nleaves, int* phase_ptr,
                      int* leaf_lo, int* leaf_hi, int cap) {
    if (checkHandle(h) || !h->colored || h->dist == DIST_POWER) return ERR_INVALID_ARG;
    const Schedule& s = rev ? h->rev : h->fwd;
    if (nphases) *nphases = s.numPhases();
    if (nleaves) *nleaves = (int)s.leafLo.size();
    if (phase_ptr && (int)s.phasePtr.size() <= cap)
        std::memcpy(phase_ptr, s.phasePtr.data(), sizeof(int) * s.phasePtr.size());
    if (leaf_lo && leaf_hi && (int)s.leafLo.size() <= cap) {
        std::memcpy(leaf_lo, s.leafLo.data(), sizeof(int) * s.leafLo.size());
        std::memcpy(leaf_hi, s.leafHi.data(), sizeof(int) * s.leafHi.size());
    }
    return OK;
}

int race_get_tree(race_handle h, int* nleaves, int* data6, int cap) {
    if (checkHandle(h) || !h->tree || !nleaves) return ERR_INVALID_ARG;
    const auto& T = h->tree->tree;
    *nleaves = (int)T.size();
    if (data6 && (int)T.size() <= cap)
        for (size_t i = 0; i < T.size(); ++i) {
            data6[6 * i + 0] = T[i].valueZ.front();
            data6[6 * i + 1] = T[i].valueZ.back();
            data6[6 * i + 2] = T[i].parentZ;
            data6[6 * i + 3] = T[i].nthreadsZ;
            data6[6 * i + 4] = T[i].stage;
            data6[6 * i + 5] = T[i].effRowZ;
        }
    return OK;
}

int race_schedule_save(race_handle h, void* buf, size_t cap, size_t* size) {
    if (checkHandle(h) || !size || !h->colored) return ERR_INVALID_ARG;
    if (h->nGraph < h->nrow) return ERR_NOT_IMPLEMENTED;  // a dense-row tail is not stored
    SavedSchedule s;
    s.nrow = h->nrow;
    s.dist = h->dist;
    s.d2 = h->d2;
    s.lbTarget = h->lbTarget;
    s.requested = h->requested;
    s.available = h->availableThreads;
    s.highestPower = h->highestPower;
    s.highestSubPower = h->highestSubPower;
    s.pattern = patternHash(h->nrow, h->rowPtr.data(), h->col.data());
    s.perm = h->perm;
    s.inv = h->inv;
    if (h->dist == DIST_POWER) {
        s.hasMpk = true;
        s.mpk = h->mpk;
    } else {
        if (!h->tree) return ERR_INTERNAL;
        s.hasTree = true;
        s.tree = h->tree->tree;
    }
    const std::vector<unsigned char> b = encodeSchedule(s);
    *size = b.size();
    if (!buf) return OK;
    if (cap < b.size()) return ERR_INVALID_ARG;
    std::memcpy(buf, b.data(), b.size());
    return OK;
}

int race_schedule_load(race_handle h, const void* buf, size_t size) {
    if (checkHandle(h) || !buf) return ERR_INVALID_ARG;
    // registered functions were laid out for the schedule they saw
    if (!h->funcs.empty()) return ERR_INVALID_ARG;
    SavedSchedule s;
    int e = decodeSchedule(buf, size, &s);
    if (e) return e;
    if (s.nrow != h->nrow || s.dist != h->dist || s.d2 != h->d2 || s.lbTarget != h->lbTarget ||
        s.requested != h->requested)
        return ERR_INVALID_ARG;
    if (s.pattern != patternHash(h->nrow, h->rowPtr.data(), h->col.data())) return ERR_INVALID_ARG;
    if (s.hasTree) {
        h->tree = std::make_unique<ZoneTree>(h->dist, h->d2, h->lbTarget);
        h->tree->tree = std::move(s.tree);
        h->fwd = flattenSchedule(*h->tree, false, h->nGraph);
        h->rev = flattenSchedule(*h->tree, true, h->nGraph);
        appendSerialPhase(h, &h->fwd);
        appendSerialPhase(h, &h->rev);
    } else {
        h->tree.reset();
        h->mpk = std::move(s.mpk);
        h->highestPower = s.highestPower;
        h->highestSubPower = s.highestSubPower;
    }
    h->perm = std::move(s.perm);
    h->inv = std::move(s.inv);
    h->availableThreads = s.available;
    h->colored = true;
    return OK;
}

int race_schedule_info(const void* buf, size_t size, int* nrow, int* dist, int* nthreads) {
    SavedSchedule s;
    int e = decodeSchedule(buf, size, &s);
    if (e) return e;
    if (nrow) *nrow = s.nrow;
    if (dist) *dist = s.dist;
    if (nthreads) *nthreads = s.available;
    return OK;
}

int race_get_mpk_groups(race_handle h, int* ngroups, int* group_ptr, int cap) {
    if (checkHandle(h) || h->dist != DIST_POWER || !h->colored || !ngroups) return ERR_INVALID_ARG;
    *ngroups = (int)h->mpk.groupPtr.size() - 1;
    if (group_ptr && (int)h->mpk.groupPtr.size() <= cap)
        std::memcpy(group_ptr, h->mpk.groupPtr.data(), sizeof(int) * h->mpk.groupPtr.size());
    return OK;
}

int race_get_launch_count(race_handle h, int id, int rev, int* launches) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !launches || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    *launches = h->funcs[id].dev->launches(rev != 0);
    return OK;
}

int race_get_flow_aborts(race_handle h, int id, int* aborts) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !aborts || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    RACE_HIP(hipStreamSynchronize(h->stream));
    return h->funcs[id].dev->flowAborts(aborts);
}

int race_get_wave_count(race_handle h, int id, int rev, int* waves) {
    if (checkHandle(h) || id < 0 || id >= (int)h->funcs.size() || !waves || !h->funcs[id].dev)
        return ERR_INVALID_ARG;
    *waves = h->funcs[id].dev->waves(rev != 0);
    return OK;
}

int race_timer_begin(race_handle h) {
    if (checkHandle(h)) return ERR_INVALID_ARG;
    if (!h->ev0 && hipEventCreate(&h->ev0) != hipSuccess) return ERR_INTERNAL;
    if (!h->ev1 && hipEventCreate(&h->ev1) != hipSuccess) return ERR_INTERNAL;
    return hipEventRecord(h->ev0, h->stream) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_timer_end(race_handle h, float* ms) {
    if (checkHandle(h) || !ms || !h->ev0) return ERR_INVALID_ARG;
    if (hipEventRecord(h->ev1, h->stream) != hipSuccess) return ERR_INTERNAL;
    if (hipEventSynchronize(h->ev1) != hipSuccess) return ERR_INTERNAL;
    return hipEventElapsedTime(ms, h->ev0, h->ev1) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_get_preprocess_stats(race_handle h, double* bfs_ms, int* traversals) {
    if (checkHandle(h) || !h->trav) return ERR_INVALID_ARG;
    if (bfs_ms) *bfs_ms = h->trav->lastTraverseMs;
    if (traversals) *traversals = (int)h->trav->traversals;
    return OK;
}

int race_sell_c_sigmize(race_handle h, int simd_width, int C, int* col, const int* chunk_start,
                        const int* row_len, const int* chunk_len) {
    if (checkHandle(h) || !col || !chunk_start || !row_len || !chunk_len || simd_width <= 0 || C <= 0)
        return ERR_INVALID_ARG;
    if (h->dist == DIST_POWER || !h->colored) return ERR_INVALID_ARG;
    sellCSigmize(h->fwd, simd_width, C, col, chunk_start, row_len, chunk_len);
    return OK;
}

int race_simdify_resolve(int simd_width, int C, int nrows, int* col, const int* chunk_start,
                         const int* chunk_len, void* val, int val_precision, int diag_first) {
    if (!col || !chunk_start || !chunk_len || !val || nrows < 0) return ERR_INVALID_ARG;
    if (val_precision == RACE_FP64)
        return simdifyResolve<double>(simd_width, C, nrows, col, chunk_start, chunk_len, (double*)val,
                                      diag_first != 0);
    if (val_precision == RACE_FP32)
        return simdifyResolve<float>(simd_width, C, nrows, col, chunk_start, chunk_len, (float*)val,
                                     diag_first != 0);
    return ERR_INVALID_ARG;
}

int race_simdify(race_handle h, int simd_width, int C, int nrows, int* col, const int* chunk_start,
                 const int* row_len, const int* chunk_len, void* val, int val_precision, int diag_first) {
    if (val_precision != RACE_FP64 && val_precision != RACE_FP32) return ERR_INVALID_ARG;
    int e = race_sell_c_sigmize(h, simd_width, C, col, chunk_start, row_len, chunk_len);
    if (e) return e;
    return race_simdify_resolve(simd_width, C, nrows, col, chunk_start, chunk_len, val, val_precision, diag_first);
}

int race_simdify_d1(int simd_width, int C, int nrows, int* col, const int* chunk_start, const int* chunk_len,
                    void* val, int val_precision, int dist2_compatible) {
    if (!col || !chunk_start || !chunk_len || !val || nrows < 0) return ERR_INVALID_ARG;
    if (val_precision == RACE_FP64)
        return simdifyD1<double>(simd_width, C, nrows, col, chunk_start, chunk_len, (double*)val,
                                 dist2_compatible != 0);
    if (val_precision == RACE_FP32)
        return simdifyD1<float>(simd_width, C, nrows, col, chunk_start, chunk_len, (float*)val,
                                dist2_compatible != 0);
    return ERR_INVALID_ARG;
}

int race_batch_create(race_batch* out, int kind, int val_precision, int vec_precision, int nrhs) {
    if (!out) return ERR_INVALID_ARG;
    *out = nullptr;
    if (kind != RACE_K_GS && kind != RACE_K_KACZ && kind != RACE_K_SYMMSPMV && kind != RACE_K_SPMV)
        return ERR_INVALID_ARG;
    if (nrhs < 1 || val_precision < RACE_FP64 || val_precision > RACE_BF16 || vec_precision < RACE_FP64 ||
        vec_precision > RACE_FP32)
        return ERR_INVALID_ARG;
    auto b = std::make_unique<race_batch_s>();
    b->kind = kind;
    b->valPrec = val_precision;
    b->vecPrec = vec_precision;
    b->nrhs = nrhs;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return ERR_INTERNAL;
    *out = b.release();
    return OK;
}

int race_batch_add(race_batch b, race_handle h, const race_kernel_desc* d) {
    if (!b || checkHandle(h) || !d || !d->rowPtr || !d->col || !d->val || !d->x) return ERR_INVALID_ARG;
    if (checkCrs(d)) return ERR_INVALID_ARG;
    if (d->kind != b->kind || d->val_precision != b->valPrec || d->vec_precision != b->vecPrec ||
        (d->nrhs > 0 ? d->nrhs : 1) != b->nrhs || d->nrow != h->nrow)
        return ERR_INVALID_ARG;
    if (bundleKind(b->kind) && (h->dist == DIST_POWER || !h->colored)) return ERR_INVALID_ARG;
    if (!d->b) return ERR_INVALID_ARG;
    race_batch_s::Mat m;
    m.h = h;
    m.rp.assign(d->rowPtr, d->rowPtr + d->nrow + 1);
    m.col.assign(d->col, d->col + m.rp.back());
    m.val.assign(d->val, d->val + m.rp.back());
    m.x = d->x;
    m.b = d->b;
    b->mats.push_back(std::move(m));
    b->fn.reset();
    return OK;
}

int race_batch_set_option(race_batch b, const char* key, int value) {
    if (!b || !key) return ERR_INVALID_ARG;
    const std::string k(key);
    if (k == "bundle_threads") {
        if (value != 64 && value != 128 && value != 256 && value != 512 && value != 1024) return ERR_INVALID_ARG;
        b->ctx.bundleThreads = value;
    } else if (k == "bundle_rows") {
        if (value < 1 || value > 8192) return ERR_INVALID_ARG;
        b->ctx.bundleRows = value;
    } else if (k == "bundle_tpr") {
        if (value != 0 && value != 1 && value != 2 && value != 4) return ERR_INVALID_ARG;
        b->ctx.bundleTpr = value;
    } else if (k == "bundle_local") {
        if (value < 1 || value > 65535) return ERR_INVALID_ARG;
        b->ctx.bundleLocal = value;
    } else if (k == "sell_c") {
        if (value != 8 && value != 16 && value != 32 && value != 64) return ERR_INVALID_ARG;
        b->ctx.sellC = value;
    } else if (k == "sell_sigma") {
        if (value < 1) return ERR_INVALID_ARG;
        b->ctx.sellSigma = value;
    } else if (k == "flow") {
        if (value != 0 && value != 1) return ERR_INVALID_ARG;
        b->ctx.flow = value;
    } else {
        return ERR_INVALID_ARG;
    }
    b->fn.reset();
    return OK;
}

int race_batch_get_option(race_batch b, const char* key, int* value) {
    if (!b || !key || !value) return ERR_INVALID_ARG;
    const std::string k(key);
    if (k == "bundle_threads") *value = b->ctx.bundleThreads;
    else if (k == "bundle_rows") *value = b->ctx.bundleRows;
    else if (k == "bundle_tpr") *value = b->ctx.bundleTpr;
    else if (k == "bundle_local") *value = b->ctx.bundleLocal;
    else if (k == "sell_c") *value = b->ctx.sellC;
    else if (k == "sell_sigma") *value = b->ctx.sellSigma;
    else if (k == "flow") *value = b->ctx.flow;
    else return ERR_INVALID_ARG;
    return OK;
}

int race_batch_get_flow_aborts(race_batch b, int* aborts) {
    if (!b || !aborts || !b->fn) return ERR_INVALID_ARG;
    RACE_HIP(hipDeviceSynchronize());
    return b->fn->flowAborts(aborts);
}

int race_batch_get_launch_count(race_batch b, int rev, int* launches) {
    if (!b || !launches || !b->fn) return ERR_INVALID_ARG;
    *launches = b->fn->launches(rev != 0);
    return OK;
}

int race_batch_build(race_batch b) {
    if (!b || b->mats.empty()) return ERR_INVALID_ARG;
    return buildBatch(b);
}

int race_batch_execute_async(race_batch b, int rev, void* stream) {
    if (!b) return ERR_INVALID_ARG;
    if (!b->fn) {
        int e = race_batch_build(b);
        if (e) return e;
    }
    return b->fn->run(rev != 0, stream ? (hipStream_t)stream : b->stream);
}

int race_batch_execute(race_batch b, int rev) {
    int e = race_batch_execute_async(b, rev, nullptr);
    if (e) return e;
    return hipStreamSynchronize(b->stream) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_batch_time(race_batch b, int rev, int reps, float* ms) {
    if (!b || !ms || reps < 1) return ERR_INVALID_ARG;
    if (!b->ev0 && hipEventCreate(&b->ev0) != hipSuccess) return ERR_INTERNAL;
    if (!b->ev1 && hipEventCreate(&b->ev1) != hipSuccess) return ERR_INTERNAL;
    int e = race_batch_execute_async(b, rev, nullptr);  // warm-up (and build)
    if (e) return e;
    if (hipEventRecord(b->ev0, b->stream) != hipSuccess) return ERR_INTERNAL;
    for (int r = 0; r < reps; ++r)
        if ((e = b->fn->run(rev != 0, b->stream))) return e;
    if (hipEventRecord(b->ev1, b->stream) != hipSuccess) return ERR_INTERNAL;
    if (hipEventSynchronize(b->ev1) != hipSuccess) return ERR_INTERNAL;
    float t = 0;
    if (hipEventElapsedTime(&t, b->ev0, b->ev1) != hipSuccess) return ERR_INTERNAL;
    *ms = t / reps;
    return OK;
}

int race_batch_candidates(int kind, int* count) {
    if (!count) return ERR_INVALID_ARG;
    *count = 12;
    (void)kind;
    return OK;
}

int race_batch_candidate(int kind, int index, int* p0, int* p1, int* p2) {
    if (index < 0 || index >= 12 || !p0 || !p1 || !p2) return ERR_INVALID_ARG;
    if (bundleKind(kind)) {
        *p0 = kBundleCand[index][0];
        *p1 = kBundleCand[index][1];
        *p2 = kBundleCand[index][2];
    } else {
        *p0 = kSellCand[index][0];
        *p1 = kSellCand[index][1];
        *p2 = 0;
    }
    return OK;
}

int race_batch_use_candidate(race_batch b, int index) {
    if (!b || index < 0 || index >= 12) return ERR_INVALID_ARG;
    applyCandidate(b, index);
    b->fn.reset();
    return OK;
}

int race_batch_autotune(race_batch b, int rev, int reps, int* best, float* times, int cap) {
    if (!b || b->mats.empty() || reps < 1) return ERR_INVALID_ARG;
    int bi = -1;
    float bt = 0;
    for (int i = 0; i < 12; ++i) {
        applyCandidate(b, i);
        float t = -1;
        int e = buildBatch(b);
        if (e == OK) e = race_batch_time(b, rev, reps, &t);
        if (e == ERR_INTERNAL) return e;
        if (e != OK) t = -1;  // layout not possible for these matrices
        if (times && i < cap) times[i] = t;
        if (t >= 0 && (bi < 0 || t < bt)) {
            bi = i;
            bt = t;
        }
    }
    if (bi < 0) return ERR_INCOMPATIBILITY;
    applyCandidate(b, bi);
    if (best) *best = bi;
    return buildBatch(b);
}

int race_batch_destroy(race_batch b) {
    delete b;
    return OK;
}

int race_lb_split(const int* level_row, const int* level_nnz, int nlevels, int nthreads,
                  double efficiency, int dist, int d2_method, int lb_target, int cap, int* nblocks,
                  int* zone_rows, int* level_ptr, int* nthreads_out, int* scale) {
    if (!level_row || !level_nnz || nlevels <= 0 || nthreads <= 0) return ERR_INVALID_ARG;
    LevelData ld;
    ld.levelRow.assign(level_row, level_row + nlevels);
    ld.levelNnz.assign(level_nnz, level_nnz + nlevels);
    LBResult r = splitZones(ld, nthreads, efficiency, dist, d2_method, lb_target);
    if (nblocks) *nblocks = (int)r.zoneRows.size();
    if (nthreads_out) *nthreads_out = r.threads();
    if ((int)r.levelPtr.size() > cap) return ERR_INVALID_ARG;
    if (zone_rows) std::memcpy(zone_rows, r.zoneRows.data(), sizeof(int) * r.zoneRows.size());
    if (level_ptr) std::memcpy(level_ptr, r.levelPtr.data(), sizeof(int) * r.levelPtr.size());
    if (scale) std::memcpy(scale, r.scale.data(), sizeof(int) * r.scale.size());
    return OK;
}

int race_split_threads(const int* weights, int ngroups, int total, int parent_threads, int* out) {
    if (!weights || !out || ngroups <= 0 || total <= 0 || parent_threads < 0) return ERR_INVALID_ARG;
    const std::vector<int> r = splitThreads(std::vector<int>(weights, weights + ngroups), total, parent_threads);
    std::memcpy(out, r.data(), sizeof(int) * ngroups);
    return OK;
}

int race_mpk_hopeless_regions(const int* violated, int nviolated, int k, int total_level, int* out, int cap,
                              int* nout) {
    if ((nviolated > 0 && !violated) || nviolated < 0 || k < 1 || total_level < 1 || !nout) return ERR_INVALID_ARG;
    const std::vector<int> r =
        identifyHopelessRegions(std::vector<int>(violated, violated + nviolated), k, total_level);
    *nout = (int)r.size();
    if ((int)r.size() > cap || (!out && !r.empty())) return ERR_INVALID_ARG;
    if (!r.empty()) std::memcpy(out, r.data(), sizeof(int) * r.size());
    return OK;
}

int race_mpk_consolidate(const int* level_ptr, const double* elem, int nlevels, int k, double cache_elem,
                         const int* regions, int nregions, int* groups, int* glevel, int* hid, int cap,
                         int* ngroups, int* nhid) {
    if (!level_ptr || !elem || nlevels < 1 || k < 1 || nregions < 0 || (nregions && !regions) || !ngroups ||
        !nhid || !groups || !glevel || !hid)
        return ERR_INVALID_ARG;
    std::vector<int> g, gl, h;
    consolidateLevels(std::vector<int>(level_ptr, level_ptr + nlevels + 1), std::vector<double>(elem, elem + nlevels),
                      k, cache_elem, std::vector<int>(regions, regions + nregions), &g, &gl, &h);
    *ngroups = (int)g.size() - 1;
    *nhid = (int)h.size();
    if ((int)g.size() > cap || (int)h.size() > cap) return ERR_INVALID_ARG;
    std::memcpy(groups, g.data(), sizeof(int) * g.size());
    std::memcpy(glevel, gl.data(), sizeof(int) * gl.size());
    if (!h.empty()) std::memcpy(hid, h.data(), sizeof(int) * h.size());
    return OK;
}

int race_get_mpk_children(race_handle h, int* nchildren, int* data4, int cap) {
    if (checkHandle(h) || h->dist != DIST_POWER || !h->colored || !nchildren) return ERR_INVALID_ARG;
    const auto& c = h->mpk.children;
    *nchildren = (int)c.size();
    if (data4 && (int)c.size() <= cap)
        for (size_t i = 0; i < c.size(); ++i) {
            data4[4 * i + 0] = c[i].lo;
            data4[4 * i + 1] = c[i].hi;
            data4[4 * i + 2] = c[i].group;
            data4[4 * i + 3] = (int)c[i].groupPtr.size() - 1;
        }
    return OK;
}

int race_device_malloc(void** p, size_t bytes) {
    if (!p) return ERR_INVALID_ARG;
    return hipMalloc(p, bytes ? bytes : 1) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_device_free(void* p) { return hipFree(p) == hipSuccess ? OK : ERR_INTERNAL; }

int race_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_memset_device(void* dst, int value, size_t bytes) {
    return hipMemset(dst, value, bytes) == hipSuccess ? OK : ERR_INTERNAL;
}

int race_device_copy(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return ERR_INVALID_ARG;
    int e = deviceCopy(dst, src, bytes, nullptr);
    if (e) return e;
    return hipDeviceSynchronize() == hipSuccess ? OK : ERR_INTERNAL;
}

int race_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? OK : ERR_INTERNAL; }

int race_device_copy_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return ERR_INVALID_ARG;
    return deviceCopy(dst, src, bytes, (hipStream_t)stream);
}

int race_get_stream(race_handle h, void** stream) {
    if (checkHandle(h) || !stream) return ERR_INVALID_ARG;
    *stream = (void*)h->stream;
    return OK;
}

int race_device_count(int* count) {
    if (!count) return ERR_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return OK;
}

}  // extern "C"
