// A RACE user program written against RACE::Interface exactly as with the
// reference (colour, permute, register a CPU Gauss-Seidel callback, execute),
// plus the GPU kernel through registerDevice.  Prints the relative difference
// between the CPU-callback result and the GPU result.  A second part runs the
// POWER interface the reference's harness uses (NUMA first touch, power
// callbacks with sub-powers, serial mode, communication hooks, the tuner)
// and compares a CPU matrix-power callback with the GPU MPK.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#include "race_amd_interface.hpp"

struct Mat {
    int n;
    std::vector<int> rp, col;
    std::vector<double> val;
};

struct GsArgs {
    const Mat* A;
    double* x;
    const double* b;
};

// GS_KERNEL of examples/src/helpers/kernels.cpp, diagonal first
static void gs_kernel(int start, int end, void* a) {
    GsArgs* g = (GsArgs*)a;
    const Mat& A = *g->A;
    int inc = start > end ? -1 : 1;
    for (int row = start; row != end; row += inc) {
        double t = g->b[row];
        for (int j = A.rp[row] + 1; j < A.rp[row + 1]; ++j) t += -A.val[j] * g->x[A.col[j]];
        g->x[row] = t / A.val[A.rp[row]];
    }
}

#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            std::printf("check failed: %s (line %d)\n", #c, __LINE__); \
            return 4;                                                \
        }                                                            \
    } while (0)

struct PowArgs {
    const Mat* A;
    double* x;  // (k+1) n: power p of row r at x[p n + r]
};

// MAT_SPMV_KERNEL of examples/src/helpers/kernels.cpp:749-777 (one product
// per (pow, subPow) pair, slot (pow-1) subPower + subPow)
static int gSubPower = 1;
static void mpk_kernel(int start, int end, int pow, int subPow, int, void* a) {
    PowArgs* p = (PowArgs*)a;
    const Mat& A = *p->A;
    const int t = (pow - 1) * gSubPower + subPow;
    for (int row = start; row < end; ++row) {
        double s = 0;
        for (int j = A.rp[row]; j < A.rp[row + 1]; ++j) s += A.val[j] * p->x[(long)(t - 1) * A.n + A.col[j]];
        p->x[(long)t * A.n + row] = s;
    }
}

// MAT_NUM_VEC_ACCESSES_w_subPower (examples/src/helpers/sparsemat.cpp:1399-1412):
// every row sees (pow, subPow) once and in order
struct CountArgs {
    std::vector<int>* cnt;
    int subPower;
    std::atomic<int>* bad;
};
static void count_kernel(int start, int end, int pow, int subPow, int, void* a) {
    CountArgs* c = (CountArgs*)a;
    for (int row = start; row < end; ++row) {
        int& v = (*c->cnt)[row];
        if (v != (pow - 1) * c->subPower + subPow - 1) c->bad->fetch_add(1);
        v = (pow - 1) * c->subPower + subPow;
    }
}

static int powerPart() {
    const int g = 20, n = g * g * g;
    Mat A;
    A.n = n;
    A.rp.push_back(0);
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j)
            for (int k = 0; k < g; ++k) {
                const int r = (i * g + j) * g + k;
                const int nb[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                for (auto& o : nb) {
                    int ii = i + o[0], jj = j + o[1], kk = k + o[2];
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= g || jj >= g || kk >= g) continue;
                    A.col.push_back((ii * g + jj) * g + kk);
                    A.val.push_back(r == A.col.back() ? 0.5 : -0.08);
                }
                A.rp.push_back((int)A.col.size());
            }
    const int power = 3, subPower = 2, k = power * subPower;
    RACE::Interface ce(n, 8, RACE::POWER, A.rp.data(), A.col.data(), true);
    CHECK(ce.RACEColor(power, 64.0 * 1024, 2, "N", subPower) == RACE_SUCCESS);
    CHECK(ce.getHighestPower() == power);
    int *split = nullptr, splitLen = 0;
    ce.getNumaSplitting(&split, &splitLen);
    CHECK(splitLen == 2 && split[0] == 0 && split[1] == n);
    std::free(split);
    ce.pinThread(0);
    ce.sleep();
    ce.resetTime();
    ce.passGlobalRows(0, n);
    ce.compressColIdx();
    int *perm, *inv, len;
    ce.getPerm(&perm, &len);
    ce.getInvPerm(&inv, &len);
    // first touch (the harness's order, examples/src/helpers/sparsemat.cpp:1019,1050)
    Mat P;
    P.n = n;
    P.rp.assign(n + 1, 7);
    ce.numaInitRowPtr(P.rp.data());
    for (int r = 0; r <= n; ++r) CHECK(P.rp[r] == 0);
    P.rp[0] = 0;
    for (int r = 0; r < n; ++r) P.rp[r + 1] = P.rp[r] + (A.rp[perm[r] + 1] - A.rp[perm[r]]);
    P.col.assign(P.rp[n], 5);
    P.val.assign(P.rp[n], 5.0);
    std::vector<double> xs((size_t)(k + 1) * n, 9.0);
    ce.numaInitMtxVec(P.rp.data(), P.col.data(), P.val.data(), xs.data(), power);
    for (int j = 0; j < P.rp[n]; ++j) CHECK(P.col[j] == 0 && P.val[j] == 0.0);
    for (int r = 0; r < 2 * n; ++r) CHECK(xs[r] == 0.0);
    int* rpL[1] = {P.rp.data()};
    int* colL[1] = {P.col.data()};
    double* valL[1] = {P.val.data()};
    ce.numaInitMtxVec(rpL, colL, valL, power);
    // the permuted matrix
    for (int r = 0; r < n; ++r) {
        int d = P.rp[r];
        for (int j = A.rp[perm[r]]; j < A.rp[perm[r] + 1]; ++j, ++d) {
            P.col[d] = inv[A.col[j]];
            P.val[d] = A.val[j];
        }
    }
    // sub-power coverage through a std::function callback, parallel and serial
    std::vector<int> cnt(n, 0);
    std::atomic<int> bad{0};
    CountArgs ca{&cnt, subPower, &bad};
    std::function<void(int, int, int, int, int, void*)> cf = count_kernel;
    int cid = ce.registerFunction(cf, &ca, power, subPower);
    CHECK(cid >= 0);
    CHECK(ce.getSubPower(cid) == subPower && ce.getPower(cid) == power);
    ce.executeFunction(cid);
    for (int r = 0; r < n; ++r) CHECK(cnt[r] == k);
    CHECK(bad.load() == 0);
    ce.setSerial(cid);
    cnt.assign(n, 0);
    ce.executeFunction(cid);
    ce.unsetSerial(cid);
    for (int r = 0; r < n; ++r) CHECK(cnt[r] == k);
    CHECK(bad.load() == 0);
    ce.setSubPower(cid, 1);
    CHECK(ce.getSubPower(cid) == 1);
    // the 4-argument (simplified) prototypes need highestSubPower 1
    std::function<void(int, int, int, int, void*)> simple = [](int, int, int, int, void*) {};
    CHECK(ce.registerFunction(simple, nullptr, power) == RACE_ERR_INCOMPATIBILITY);
    // communication hooks (stored; one device has no remote rows)
    int comm = 0;
    std::function<void(int, int, void*)> cm = [&comm](int, int, void*) { ++comm; };
    CHECK(ce.attachCommunicationToFunction(cid, cm, nullptr) == RACE_SUCCESS);
    CHECK(ce.attachCommunicationToFunction(cid, std::function<void(int, void*)>([](int, void*) {}), nullptr) ==
          RACE_ERR_INCOMPATIBILITY);
    ce.checkPowerCoverage();

    // CPU matrix powers through the callback against the GPU MPK with sub-powers
    gSubPower = subPower;
    for (int r = 0; r < n; ++r) xs[r] = std::cos(0.37 * r);
    PowArgs pa{&P, xs.data()};
    int pid = ce.registerFunction(&mpk_kernel, &pa, power, subPower);
    CHECK(pid >= 0);
    ce.executeFunction(pid);
    double* dx;
    const size_t bytes = (size_t)(k + 1) * n * sizeof(double);
    if (hipMalloc(&dx, bytes) != hipSuccess) return 3;
    std::vector<double> xg((size_t)(k + 1) * n, 0.0);
    for (int r = 0; r < n; ++r) xg[r] = xs[r];
    if (hipMemcpy(dx, xg.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return 3;
    race_kernel_desc d{};
    d.kind = RACE_K_MPK;
    d.nrow = n;
    d.rowPtr = P.rp.data();
    d.col = P.col.data();
    d.val = P.val.data();
    d.val_precision = RACE_FP64;
    d.vec_precision = RACE_FP64;
    d.nrhs = 1;
    d.x = dx;
    d.power = power;
    d.sub_power = subPower;
    int did = ce.registerDevice(d);
    CHECK(did >= 0);
    CHECK(ce.getSubPower(did) == subPower);
    ce.executeFunction(did);
    if (hipMemcpy(xg.data(), dx, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 3;
    double worst = 0;
    for (int t = 1; t <= k; ++t) {
        double num = 0, den = 0;
        for (int r = 0; r < n; ++r) {
            const double a = xg[(size_t)t * n + r], b = xs[(size_t)t * n + r];
            num += (a - b) * (a - b);
            den += b * b;
        }
        worst = std::max(worst, std::sqrt(num / den));
    }
    // a device MPK runs at most its registered products
    ce.setSubPower(did, subPower + 1);
    CHECK(ce.getSubPower(did) == subPower);
    ce.setSubPower(did, 1);
    ce.setPower(did, k);
    CHECK(ce.getPower(did) == k);
    CHECK(ce.tuneFunction(did) >= 1);
    std::printf("power: products=%d relerr=%.3e comm_calls=%d\n", k, worst, comm);
    delete[] perm;
    delete[] inv;
    (void)hipFree(dx);
    return worst <= 1e-12 && comm == 0 ? 0 : 1;
}

int main() {
    const int g = 10, n = g * g * g;
    Mat A;
    A.n = n;
    A.rp.push_back(0);
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j)
            for (int k = 0; k < g; ++k) {
                for (int a = -1; a <= 1; ++a)
                    for (int b = -1; b <= 1; ++b)
                        for (int c = -1; c <= 1; ++c) {
                            int ii = i + a, jj = j + b, kk = k + c;
                            if (ii < 0 || jj < 0 || kk < 0 || ii >= g || jj >= g || kk >= g) continue;
                            A.col.push_back((ii * g + jj) * g + kk);
                        }
                A.rp.push_back((int)A.col.size());
            }
    RACE::Interface ce(n, 16, RACE::ONE, A.rp.data(), A.col.data(), true);
    if (ce.RACEColor() != RACE_SUCCESS) return 2;
    int *perm, *inv, len;
    ce.getPerm(&perm, &len);
    ce.getInvPerm(&inv, &len);
    // permuted matrix, diagonal first; off-diagonals -1, diagonal 27
    Mat P;
    P.n = n;
    P.rp.push_back(0);
    for (int r = 0; r < n; ++r) {
        int o = perm[r];
        P.col.push_back(r);
        P.val.push_back(27.0);
        for (int j = A.rp[o]; j < A.rp[o + 1]; ++j)
            if (A.col[j] != o) {
                P.col.push_back(inv[A.col[j]]);
                P.val.push_back(-1.0);
            }
        P.rp.push_back((int)P.col.size());
    }
    std::vector<double> b(n), xc(n, 0.0), xg(n, 0.0);
    for (int r = 0; r < n; ++r) b[r] = std::sin(0.1 * r);
    GsArgs ga{&P, xc.data(), b.data()};
    int hid = ce.registerFunction(&gs_kernel, &ga);
    ce.executeFunction(hid);
    ce.executeFunction(hid, true);

    double *dx, *db;
    const std::vector<double> xstart(xg.begin(), xg.begin() + n);
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess) return 3;
    if (hipMalloc(&db, n * sizeof(double)) != hipSuccess) return 3;
    if (hipMemcpy(dx, xg.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 3;
    if (hipMemcpy(db, b.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 3;
    race_kernel_desc d{};
    d.kind = RACE_K_GS;
    d.nrow = n;
    d.rowPtr = P.rp.data();
    d.col = P.col.data();
    d.val = P.val.data();
    d.val_precision = RACE_FP64;
    d.vec_precision = RACE_FP64;
    d.nrhs = 1;
    d.x = dx;
    d.b = db;
    int did = ce.registerDevice(d);
    ce.executeFunction(did);
    ce.executeFunction(did, true);
    if (hipMemcpy(xg.data(), dx, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    double num = 0, den = 0;
    for (int r = 0; r < n; ++r) {
        num += (xg[r] - xc[r]) * (xg[r] - xc[r]);
        den += xc[r] * xc[r];
    }
    std::printf("threads=%d eff=%.3f relerr=%.3e\n", ce.getNumThreads(), ce.getEfficiency(),
                std::sqrt(num / den));
    const bool colorOk = std::sqrt(num / den) <= 1e-12;
    // the same device sweeps, one launch per sweep (engine option "flow"):
    // bit for bit the phase launches
    {
        CHECK(ce.setOption("flow", 1) == RACE_SUCCESS);
        double* dx2;
        if (hipMalloc(&dx2, n * sizeof(double)) != hipSuccess) return 3;
        if (hipMemcpy(dx2, xstart.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 3;
        race_kernel_desc d2 = d;
        d2.x = dx2;
        const int fid = ce.registerDevice(d2);
        CHECK(fid >= 0);
        ce.executeFunction(fid);
        ce.executeFunction(fid, true);
        std::vector<double> xf(n);
        if (hipMemcpy(xf.data(), dx2, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 3;
        for (int r = 0; r < n; ++r) CHECK(xf[r] == xg[r]);
        CHECK(ce.setOption("flow", 0) == RACE_SUCCESS);
        (void)hipFree(dx2);
    }
    // the same sweeps on one thread (setSerial) give the same x
    std::vector<double> xs(n, 0.0);
    GsArgs gs1{&P, xs.data(), b.data()};
    int sid = ce.registerFunction(std::function<void(int, int, void*)>(gs_kernel), &gs1);
    ce.setSerial(sid);
    ce.executeFunction(sid);
    ce.executeFunction(sid, true);
    ce.unsetSerial(sid);
    for (int r = 0; r < n; ++r) CHECK(xs[r] == xc[r]);
    // NUMA helpers are POWER-only, as in the reference
    CHECK(race_numa_init_row_ptr(ce.handle(), P.rp.data()) == RACE_ERR_NOT_IMPLEMENTED);
    delete[] perm;
    delete[] inv;
    (void)hipFree(dx);
    (void)hipFree(db);
    if (!colorOk) return 1;
    return powerPart();
}
