// A RACE user program written against RACE::Interface exactly as with the
// reference (colour, permute, register a CPU Gauss-Seidel callback, execute),
// plus the GPU kernel through registerDevice.  Prints the relative difference
// between the CPU-callback result and the GPU result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
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
    delete[] perm;
    delete[] inv;
    (void)hipFree(dx);
    (void)hipFree(db);
    return std::sqrt(num / den) <= 1e-12 ? 0 : 1;
}
