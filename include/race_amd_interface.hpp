// race_amd_interface.hpp -- header-only RACE::Interface over the C ABI.
//
// Drop-in for RRZE-HPC/RACE's C++ interface (include/interface.h:54-205):
// same class name, constructor, RACEColor overloads, getPerm/getInvPerm
// ownership (new[] arrays the caller delete[]s), registerFunction overloads
// and executeFunction.  Host callbacks keep working; GPU kernels are added
// through registerDevice (the typed device registration of race_amd.h).
#ifndef RACE_AMD_INTERFACE_HPP
#define RACE_AMD_INTERFACE_HPP

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "race_amd.h"

namespace RACE {

// include/type.h:31-55
enum dist { ONE = RACE_ONE, TWO = RACE_TWO, POWER = RACE_POWER };
enum d2Method { TWO_BLOCK = RACE_TWO_BLOCK, THREE_BLOCK = RACE_THREE_BLOCK };
enum LBTarget { ROW = RACE_LB_ROW, NNZ = RACE_LB_NNZ };
enum PinMethod { FILL = RACE_FILL, SCATTER = RACE_SCATTER };

class Interface {
  public:
    // Interface::Interface include/interface.h:127
    Interface(int nrow, int nthreads, RACE::dist dist, int* rowPtr, int* col, bool symm_hint = false,
              int SMT = 1, RACE::PinMethod method = RACE::SCATTER, int* initPerm = NULL,
              int* initInvPerm = NULL, RACE::d2Method d2Type = RACE::TWO_BLOCK,
              RACE::LBTarget lbTarget = RACE::NNZ)
        : nrow_(nrow) {
        int s = race_create(&h_, nrow, nthreads, (int)dist, rowPtr, col, symm_hint ? 1 : 0, SMT,
                            (int)method, initPerm, initInvPerm, (int)d2Type, (int)lbTarget);
        if (s != RACE_SUCCESS) {
            std::fprintf(stderr, "RACE::Interface: %s\n", race_error_string(s));
            h_ = nullptr;
        }
    }
    ~Interface() {
        if (h_) race_destroy(h_);
    }
    Interface(const Interface&) = delete;
    Interface& operator=(const Interface&) = delete;

    // preprocessing (include/interface.h:129-130)
    RACE_error RACEColor(int highestPower, double cacheSize, double safetyFactor = 2,
                         std::string mtxType = "N", int highestSubPower = 1) {
        if (!h_) return RACE_ERR_INVALID_ARG;
        if (mtxType.size() != 1) return RACE_ERR_INVALID_ARG;
        RACE_error e = (RACE_error)race_color_power(h_, highestPower, cacheSize, safetyFactor, mtxType[0],
                                                    highestSubPower);
        if (e == RACE_SUCCESS) highestSubPower_ = highestSubPower;
        return e;
    }
    RACE_error RACEColor() { return h_ ? (RACE_error)race_color(h_) : RACE_ERR_INVALID_ARG; }

    // schedule save / load (no counterpart in the reference): what RACEColor
    // computed, as bytes; loadSchedule replaces RACEColor on a handle built
    // with the same matrix pattern and constructor arguments
    std::vector<unsigned char> saveSchedule() {
        size_t n = 0;
        if (race_schedule_save(h_, nullptr, 0, &n) != RACE_SUCCESS) return {};
        std::vector<unsigned char> b(n);
        if (race_schedule_save(h_, b.data(), b.size(), &n) != RACE_SUCCESS) return {};
        return b;
    }
    RACE_error loadSchedule(const std::vector<unsigned char>& b) {
        return h_ ? (RACE_error)race_schedule_load(h_, b.data(), b.size()) : RACE_ERR_INVALID_ARG;
    }

    double getEfficiency() {
        double e = 0;
        race_get_efficiency(h_, &e);
        return e;
    }
    int getMaxStageDepth() {
        int d = 0;
        race_get_max_stage_depth(h_, &d);
        return d;
    }
    int getNumThreads() {
        int t = 0;
        race_get_num_threads(h_, &t);
        return t;
    }
    int getHighestPower() {
        int p = 0;
        race_get_highest_power(h_, &p);
        return p;
    }
    // perm[new] = old; new[] arrays owned by the caller (perf_test's harness
    // frees them with delete[])
    void getPerm(int** perm, int* len) {
        *perm = new int[nrow_];
        *len = nrow_;
        race_get_perm(h_, *perm, nrow_);
    }
    void getInvPerm(int** invPerm, int* len) {
        *invPerm = new int[nrow_];
        *len = nrow_;
        race_get_inv_perm(h_, *invPerm, nrow_);
    }
    void printZoneTree() {
        int n = 0;
        race_get_tree(h_, &n, nullptr, 0);
        std::vector<int> d(6 * (size_t)n);
        race_get_tree(h_, &n, d.data(), n);
        for (int i = 0; i < n; ++i)
            std::printf("%d. Range:[%d, %d] Parent: %d, nthreads: %d, stage:%d, effRow: %d\n", i,
                        d[6 * i], d[6 * i + 1], d[6 * i + 2], d[6 * i + 3], d[6 * i + 4], d[6 * i + 5]);
    }

    // registration (include/interface.h:158-171).  As in the reference, a
    // failure returns the RACE_error value instead of an id.
    int registerFunction(void (*f)(int, int, void*), void* args) {
        int id = -1;
        int s = race_register_function(h_, f, args, &id);
        return s == RACE_SUCCESS ? id : s;
    }
    int registerFunction(std::function<void(int, int, void*)> f, void* args) {
        colorFns_.push_back(std::make_unique<ColorThunk>(ColorThunk{std::move(f), args}));
        int id = -1;
        int s = race_register_function(h_, &ColorThunk::call, colorFns_.back().get(), &id);
        return s == RACE_SUCCESS ? id : s;
    }
    int registerFunction(void (*f)(int, int, int, int, int, void*), void* args, int power = 1,
                         int subPower = 1, int numaSplit = false) {
        int id = -1;
        int s = race_register_power_function(h_, f, args, power, subPower, numaSplit, &id);
        return s == RACE_SUCCESS ? id : s;
    }
    int registerFunction(std::function<void(int, int, int, int, int, void*)> f, void* args,
                         int power = 1, int subPower = 1, int numaSplit = false) {
        powerFns_.push_back(std::make_unique<PowerThunk>(PowerThunk{std::move(f), args}));
        int id = -1;
        int s = race_register_power_function(h_, &PowerThunk::call, powerFns_.back().get(), power,
                                             subPower, numaSplit, &id);
        return s == RACE_SUCCESS ? id : s;
    }
    // simplified power callback f(start, end, pow, numa, args), only when
    // highestSubPower == 1 (src/interface.cpp:468-485, else
    // RACE_ERR_INCOMPATIBILITY)
    int registerFunction(std::function<void(int, int, int, int, void*)> f, void* args, int power = 1,
                         int numaSplit = false) {
        if (highestSubPower_ != 1) return RACE_ERR_INCOMPATIBILITY;
        std::function<void(int, int, int, int, int, void*)> g =
            [f](int s, int e, int p, int, int numa, void* a) { f(s, e, p, numa, a); };
        return registerFunction(g, args, power, 1, numaSplit);
    }
    int registerFunction(void (*f)(int, int, int, int, void*), void* args, int power = 1,
                         int numaSplit = false) {
        return registerFunction(std::function<void(int, int, int, int, void*)>(f), args, power, numaSplit);
    }
    // Interface::attachCommunicationToFunction (include/interface.h:161-165):
    // stored, never called on one device (no remote rows), as in the
    // reference's non-MPI build
    int attachCommunicationToFunction(int funcId, void (*f)(int, int, void*), void* args) {
        return race_attach_communication(h_, funcId, f, args);
    }
    int attachCommunicationToFunction(int funcId, std::function<void(int, int, void*)> f, void* args) {
        commFns_.push_back(std::make_unique<CommThunk>(CommThunk{std::move(f), args}));
        return race_attach_communication(h_, funcId, &CommThunk::call, commFns_.back().get());
    }
    int attachCommunicationToFunction(int funcId, void (*f)(int, void*), void* args) {
        return race_attach_communication_simple(h_, funcId, f, args);
    }
    int attachCommunicationToFunction(int funcId, std::function<void(int, void*)> f, void* args) {
        commSimpleFns_.push_back(std::make_unique<CommSimpleThunk>(CommSimpleThunk{std::move(f), args}));
        return race_attach_communication_simple(h_, funcId, &CommSimpleThunk::call, commSimpleFns_.back().get());
    }
    // engine options of race_set_option ("layout", bundle shape, "graph",
    // "flow", ...), set before registering (no reference counterpart)
    int setOption(const char* key, int value) { return race_set_option(h_, key, value); }
    // GPU kernels: the typed device registration (no reference counterpart)
    int registerDevice(const race_kernel_desc& d) {
        int id = -1;
        int s = race_register_device(h_, &d, &id);
        return s == RACE_SUCCESS ? id : s;
    }

    // Interface::executeFunction (src/interface.cpp:541-572), synchronous
    int executeFunction(int funcId, bool rev = false) { return race_execute(h_, funcId, rev ? 1 : 0); }
    // asynchronous on a HIP stream passed as void* (NULL: the handle's stream)
    int executeFunctionAsync(int funcId, bool rev, void* stream) {
        return race_execute_async(h_, funcId, rev ? 1 : 0, stream);
    }
    void setPower(int funcId, int pow) { race_set_power(h_, funcId, pow); }
    int getPower(int funcId) {
        int p = 0;
        race_get_power(h_, funcId, &p);
        return p;
    }
    // include/interface.h:168-172
    void setSubPower(int funcId, int subPow) { race_set_sub_power(h_, funcId, subPow); }
    int getSubPower(int funcId) {
        int p = 0;
        race_get_sub_power(h_, funcId, &p);
        return p;
    }
    void setSerial(int funcId) { race_set_serial(h_, funcId, 1); }
    void unsetSerial(int funcId) { race_set_serial(h_, funcId, 0); }
    // include/interface.h:176: per-leaf timers (none are kept here)
    void resetTime() { race_reset_time(h_); }
    // include/interface.h:141: parks RACE's own pool (OpenMP threads here)
    void sleep() { race_sleep(h_); }
    // include/interface.h:141-142: MPI helpers, empty in the reference too
    // (src/interface.cpp:63-67, 70-202)
    void passGlobalRows(int globalStartRow, int globalEndRow) {
        (void)globalStartRow;
        (void)globalEndRow;
    }
    void compressColIdx() {}
    // include/interface.h:189-201
    void pinThread(int threadId) { race_pin_thread(h_, threadId); }
    void numaInitRowPtr(int* rowPtr) { report(race_numa_init_row_ptr(h_, rowPtr), "numaInitRowPtr"); }
    void numaInitMtxVec(int* rowPtr, int* col, double* val, double* x, int power = 1) {
        report(race_numa_init_mtx_vec(h_, rowPtr, col, val, x, power), "numaInitMtxVec");
    }
    void numaInitRowPtr(int** rowPtr) { report(race_numa_init_row_ptr_local(h_, rowPtr), "numaInitRowPtr"); }
    void numaInitMtxVec(int** rowPtr, int** col, double** val, int power = 1) {
        report(race_numa_init_mtx_vec_local(h_, rowPtr, col, val, power), "numaInitMtxVec");
    }
    // split is malloc'ed, as in the reference (src/interface.cpp:1042-1052)
    void getNumaSplitting(int** split, int* splitLen) {
        int n = 0;
        *split = nullptr;
        *splitLen = 0;
        if (race_get_numa_splitting(h_, nullptr, 0, &n) != RACE_SUCCESS) return;
        *split = (int*)std::malloc(sizeof(int) * n);
        if (*split && race_get_numa_splitting(h_, *split, n, &n) == RACE_SUCCESS) *splitLen = n;
    }
    int tuneFunction(int funcId, bool rev = false) {
        int p = 1;
        race_tune_function(h_, funcId, rev ? 1 : 0, &p);
        return p;
    }
    void checkPowerCoverage() {
        if (race_check_power_coverage(h_) != RACE_SUCCESS)
            std::fprintf(stderr, "RACE: power coverage check failed\n");
    }
    // Interface::simdify / simdifyD1 (include/interface.h:179-185,
    // src/interface.cpp:824-842): SELL-C-sigma rearrangement in place
    bool simdify(int simdWidth, int C, int nrows, int* col, int* chunkStart, int* rl, int* clp, double* val,
                 bool diagFirst = false) {
        return race_simdify(h_, simdWidth, C, nrows, col, chunkStart, rl, clp, val, RACE_FP64, diagFirst ? 1 : 0) ==
               RACE_SUCCESS;
    }
    bool simdify(int simdWidth, int C, int nrows, int* col, int* chunkStart, int* rl, int* clp, float* val,
                 bool diagFirst = false) {
        return race_simdify(h_, simdWidth, C, nrows, col, chunkStart, rl, clp, val, RACE_FP32, diagFirst ? 1 : 0) ==
               RACE_SUCCESS;
    }
    bool simdifyD1(int simdWidth, int C, int nrows, int* col, int* chunkStart, int* rl, int* clp, double* val,
                   bool dist2Compatible = false) {
        (void)rl;
        return race_simdify_d1(simdWidth, C, nrows, col, chunkStart, clp, val, RACE_FP64, dist2Compatible ? 1 : 0) ==
               RACE_SUCCESS;
    }
    bool simdifyD1(int simdWidth, int C, int nrows, int* col, int* chunkStart, int* rl, int* clp, float* val,
                   bool dist2Compatible = false) {
        (void)rl;
        return race_simdify_d1(simdWidth, C, nrows, col, chunkStart, clp, val, RACE_FP32, dist2Compatible ? 1 : 0) ==
               RACE_SUCCESS;
    }
    race_handle handle() const { return h_; }

  private:
    struct ColorThunk {
        std::function<void(int, int, void*)> f;
        void* args;
        static void call(int s, int e, void* self) {
            ColorThunk* t = static_cast<ColorThunk*>(self);
            t->f(s, e, t->args);
        }
    };
    struct PowerThunk {
        std::function<void(int, int, int, int, int, void*)> f;
        void* args;
        static void call(int s, int e, int p, int sp, int numa, void* self) {
            PowerThunk* t = static_cast<PowerThunk*>(self);
            t->f(s, e, p, sp, numa, t->args);
        }
    };
    struct CommThunk {
        std::function<void(int, int, void*)> f;
        void* args;
        static void call(int p, int sp, void* self) {
            CommThunk* t = static_cast<CommThunk*>(self);
            t->f(p, sp, t->args);
        }
    };
    struct CommSimpleThunk {
        std::function<void(int, void*)> f;
        void* args;
        static void call(int p, void* self) {
            CommSimpleThunk* t = static_cast<CommSimpleThunk*>(self);
            t->f(p, t->args);
        }
    };
    // the reference prints its errors from these void methods
    static void report(int s, const char* what) {
        if (s != RACE_SUCCESS) std::fprintf(stderr, "RACE::Interface::%s: %s\n", what, race_error_string(s));
    }
    race_handle h_ = nullptr;
    int nrow_ = 0;
    int highestSubPower_ = 1;
    std::vector<std::unique_ptr<ColorThunk>> colorFns_;
    std::vector<std::unique_ptr<PowerThunk>> powerFns_;
    std::vector<std::unique_ptr<CommThunk>> commFns_;
    std::vector<std::unique_ptr<CommSimpleThunk>> commSimpleFns_;
};

}  // namespace RACE

// sell_c_sigmize (include/sell_c_sigmize.h:41, src/sell_c_sigmize.cpp:48-62)
inline void sell_c_sigmize(int simdWidth, int C, int* col, int* chunkStart, int* rl, int* clp,
                           RACE::Interface* ce) {
    race_sell_c_sigmize(ce->handle(), simdWidth, C, col, chunkStart, rl, clp);
}

#endif  // RACE_AMD_INTERFACE_HPP
