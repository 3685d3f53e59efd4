/*
 * race_oracle.c -- CPU oracle for the RACE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (race_amd/csrc) never links, loads or calls it.
 *
 * It restates, in plain sequential C, the parts of RRZE-HPC/RACE that work
 * row by row (citations are into the reference tree):
 *
 *   or_symmetrize      RACE::makeSymmetricGraph        src/graph.cpp:47-142
 *   or_traverse        RACE::Traverse ctor/TDStep/calculateDistance/
 *                      findLevelData/permuteGraph      src/traverse_GAP_graphSoA.cpp:31-94,
 *                                                      135-221,236-383,385-504,506-641
 *                      updatePerm                      include/utility.h:114-149
 *   or_symmspmv_range  SYMM_SPMV_KERNEL_BODY           examples/src/helpers/kernels.cpp:403-427
 *   or_gs_range        GS_KERNEL / GS_KERNEL_BODY      examples/src/helpers/kernels.cpp:174-199
 *   or_kacz_range      KACZ_KERNEL / KACZ_KERNEL_BODY  examples/src/helpers/kernels.cpp:285-316
 *   or_spmv_range      SPMV_KERNEL_BODY                examples/src/helpers/kernels.cpp:5-25
 *   or_mpk_serial      MAT_SPMV_KERNEL applied power by power
 *                                                      examples/src/helpers/kernels.cpp:749-777
 *   or_run_calls       the leaf calls of recursiveCall  src/functionManager.cpp:199-265
 *                      (serialised; concurrent leaves are independent)
 *   or_run_phases      the same phase by phase, the leaves of a phase on
 *                      OpenMP threads (RACE's own execution on the host, the
 *                      CPU baseline)
 *
 * How each function is pinned (the reference cannot be built here; see the
 * test files for the checks):
 *   or_traverse     -- closed-form level counts of 7-/27-point grids, shortest
 *                      path invariants on random graphs, stage-0 permutation ==
 *                      stable sort by level, island rule on disconnected graphs.
 *   or_symmspmv_range -- equals the full SpMV A*x (textbook reduction).
 *   or_gs_range     -- one leaf over all rows equals textbook forward/backward
 *                      Gauss-Seidel computed independently with scipy.
 *   or_kacz_range   -- one leaf equals textbook Kaczmarz projection (the
 *                      updated x satisfies row r exactly after its step).
 *   or_mpk_serial   -- equals repeated scipy SpMV.
 *   or_symmetrize   -- result pattern is symmetric and a superset; brute force.
 */
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <string.h>
#include <math.h>

enum { OR_ONE = 1, OR_TWO = 2, OR_POWER = 3 };

/* ------------------------------------------------------------------------ */
/* Graph symmetrisation: add (c,r) for every (r,c) whose pair is missing.    */
/* Returns new nnz; out_rowPtr[n+1], out_col[cap] are caller buffers.  If    */
/* cap is too small, returns -needed.                                        */
/* ------------------------------------------------------------------------ */
long or_symmetrize(int n, const int *rowPtr, const int *col, int *out_rowPtr,
                   int *out_col, long cap)
{
    int *extraCnt = (int *)calloc((size_t)n, sizeof(int));
    long extra = 0;
    for (int r = 0; r < n; ++r)
        for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) {
            int c = col[j];
            if (c != r && c >= 0 && c < n) {
                int has = 0;
                for (int k = rowPtr[c]; k < rowPtr[c + 1]; ++k)
                    if (col[k] == r) { has = 1; break; }
                if (!has) { extraCnt[c]++; extra++; }
            }
        }
    long nnz = (long)rowPtr[n] + extra;
    if (nnz > cap) { free(extraCnt); return -nnz; }
    out_rowPtr[0] = 0;
    for (int r = 0; r < n; ++r)
        out_rowPtr[r + 1] = out_rowPtr[r] + (rowPtr[r + 1] - rowPtr[r]) + extraCnt[r];
    int *fill = (int *)malloc(sizeof(int) * (size_t)n);
    for (int r = 0; r < n; ++r) {
        int o = out_rowPtr[r];
        for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) out_col[o++] = col[j];
        fill[r] = o;
    }
    for (int r = 0; r < n; ++r)
        for (int j = rowPtr[r]; j < rowPtr[r + 1]; ++j) {
            int c = col[j];
            if (c != r && c >= 0 && c < n) {
                int has = 0;
                for (int k = rowPtr[c]; k < rowPtr[c + 1]; ++k)
                    if (col[k] == r) { has = 1; break; }
                if (!has) out_col[fill[c]++] = r;
            }
        }
    free(fill);
    free(extraCnt);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Traversal workspace: dist/parent are kept at -1 between calls; only the   */
/* entries touched by a traversal are reset afterwards.                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int *dist, *parent, *queue, *touched, *perm, *cnt;
    long ntouched;
} or_ws;

void *or_ws_create(int n)
{
    or_ws *w = (or_ws *)calloc(1, sizeof(or_ws));
    w->n = n;
    w->dist = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    w->parent = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    w->queue = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    w->touched = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    w->perm = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    w->cnt = (int *)malloc(sizeof(int) * (size_t)(3 * (size_t)n + 4));
    for (int i = 0; i < n; ++i) { w->dist[i] = -1; w->parent[i] = -1; }
    return w;
}

void or_ws_destroy(void *p)
{
    or_ws *w = (or_ws *)p;
    if (!w) return;
    free(w->dist); free(w->parent); free(w->queue); free(w->touched);
    free(w->perm); free(w->cnt); free(w);
}

static inline void ws_visit(or_ws *w, int c, int par)
{
    w->parent[c] = par;
    w->touched[w->ntouched++] = c;
}

/*
 * Level-synchronous BFS inside [lo,hi) with RACE's distance semantics,
 * followed by the level histograms and the stable by-level permutation
 * composed into totalPerm/totalInvPerm (RACE_PERMUTE_ON_FLY semantics:
 * the graph stays in original numbering; u -> totalPerm[u] on reads,
 * child -> totalInvPerm[child]).
 *
 * levelRow/levelNnz: caller buffers of at least 3*(hi-lo)+2 ints (with
 * distance-2 bridges an isolated row of the range can cost three levels:
 * the restart's empty level, its own, and the one that expands its
 * bridges).
 * distOut (optional, may be NULL): receives dist for rows [lo,hi) in the
 * index space BEFORE the permutation is applied.
 * Returns totalLevel, or -1 when rows go missing (the reference exits).
 */
int or_traverse(void *wsp, int n, const int *rowPtr, const int *col,
                int *totalPerm, int *totalInvPerm, int dist, int lo, int hi,
                int *levelRow, int *levelNnz, int *distOut)
{
    or_ws *w = (or_ws *)wsp;
    if (lo >= hi) return 0;
    int *q = w->queue;
    long qhead = 0, qtail = 0;      /* current frontier = q[qhead, qtail) */
    long counter = 0;
    int currLvl = 0;
    w->ntouched = 0;

    /* root = {lo}  (src/traverse_GAP_graphSoA.cpp:33-39, 258-265) */
    q[qtail++] = lo;
    w->dist[lo] = currLvl;
    ws_visit(w, lo, 0);
    counter++;
    int prevIslandStop = lo;

    while (qtail > qhead) {
        currLvl += 1;
        long fEnd = qtail;
        /* TDStep(currLvl) :135-221 */
        for (long qi = qhead; qi < fEnd; ++qi) {
            int u = q[qi];
            int pu = totalPerm[u];
            for (int j = rowPtr[pu]; j < rowPtr[pu + 1]; ++j) {
                int c = col[j];
                if (c < n) {
                    c = totalInvPerm[c];
                    if (w->parent[c] == -1) {
                        if (c >= lo && c < hi) {
                            ws_visit(w, c, u);
                            q[qtail++] = c;
                            w->dist[c] = currLvl;
                            counter++;
                        } else if (dist == OR_TWO) {
                            if (u >= lo && u < hi) {
                                ws_visit(w, c, u);
                                q[qtail++] = c;
                            }
                        }
                    }
                }
            }
        }
        qhead = fEnd;
        /* island detection :277-306 */
        if (qtail == qhead && counter != (long)(hi - lo)) {
            for (int i = prevIslandStop; i < hi; ++i) {
                if (w->dist[i] == -1) {
                    if (dist != OR_POWER) currLvl += 1;
                    q[qtail++] = i;
                    counter++;
                    w->dist[i] = currLvl;
                    ws_visit(w, i, i);
                    prevIslandStop = i;
                    break;
                }
            }
            if (qtail == qhead) return -1;
        }
    }
    int totalLevel = currLvl;

    /* findLevelData :385-447 */
    for (int L = 0; L < totalLevel; ++L) { levelRow[L] = 0; levelNnz[L] = 0; }
    for (int i = lo; i < hi; ++i) {
        int d = w->dist[i];
        if (d < 0) return -1;
        levelRow[d] += 1;
        int pr = totalPerm[i];
        levelNnz[d] += rowPtr[pr + 1] - rowPtr[pr];
    }
    if (distOut)
        for (int i = lo; i < hi; ++i) distOut[i - lo] = w->dist[i];

    /* permuteGraph :506-641 -- stable sort of [lo,hi) by level (counting
     * sort visiting rows in index order is the unique stable order). */
    int *off = w->cnt;
    int acc = 0;
    for (int L = 0; L < totalLevel; ++L) { off[L] = acc; acc += levelRow[L]; }
    int *perm = w->perm;
    for (int i = lo; i < hi; ++i) perm[lo + off[w->dist[i]]++] = i;
    /* updatePerm: totalPerm_new[i] = totalPerm_old[perm[i]] on the range */
    int *tmp = w->queue; /* queue no longer needed */
    for (int i = lo; i < hi; ++i) tmp[i - lo] = totalPerm[perm[i]];
    for (int i = lo; i < hi; ++i) {
        totalPerm[i] = tmp[i - lo];
        totalInvPerm[totalPerm[i]] = i;
    }

    /* reset the workspace */
    for (long t = 0; t < w->ntouched; ++t) {
        int c = w->touched[t];
        w->dist[c] = -1;
        w->parent[c] = -1;
    }
    w->ntouched = 0;
    return totalLevel;
}

/*
 * The POWER traversal of a recursion stage (mtxPowerRecursive, a hopeless
 * region [lo,hi) with its boundary regions): Traverse ctor with a boundary
 * bitmap (src/traverse_GAP_graphSoA.cpp:76-88), TDStep's boundary branch
 * (:184-197: a boundary row reached from the frontier gets the level and is
 * expanded, but does not count towards the range), island restarts over the
 * range only (:277-306, adjacent levels for POWER), createLevelData
 * (:449-495: untouched boundary rows go to one extra last level, at least
 * one level), findLevelData for the range and every boundary region, and
 * permuteGraph (:506-575: each region stably sorted by level, composed into
 * totalPerm).
 * nb boundary regions [blo[b], bhi[b]); levelRow/levelNnz hold cap entries,
 * bLevelRow/bLevelNnz nb * cap (region b at b * cap).  Returns totalLevel,
 * or -1 when rows go missing or cap is too small.
 */
int or_traverse_boundary(void *wsp, int n, const int *rowPtr, const int *col, int *totalPerm,
                         int *totalInvPerm, int lo, int hi, int nb, const int *blo, const int *bhi,
                         int *levelRow, int *levelNnz, int *bLevelRow, int *bLevelNnz, int cap)
{
    or_ws *w = (or_ws *)wsp;
    if (lo >= hi) return 0;
    char *inB = (char *)calloc((size_t)n, 1);
    for (int b = 0; b < nb; ++b)
        for (int i = blo[b]; i < bhi[b]; ++i) inB[i] = 1;
    int *q = w->queue;
    long qhead = 0, qtail = 0;
    long counter = 0;
    int currLvl = 0;
    w->ntouched = 0;
    q[qtail++] = lo;
    w->dist[lo] = currLvl;
    ws_visit(w, lo, 0);
    counter++;
    int prevIslandStop = lo;
    int status = 0;
    while (qtail > qhead) {
        currLvl += 1;
        long fEnd = qtail;
        for (long qi = qhead; qi < fEnd; ++qi) {
            int u = q[qi];
            int pu = totalPerm[u];
            for (int j = rowPtr[pu]; j < rowPtr[pu + 1]; ++j) {
                int c = col[j];
                if (c >= n) continue;
                c = totalInvPerm[c];
                if (w->parent[c] != -1) continue;
                if (c >= lo && c < hi) {
                    ws_visit(w, c, u);
                    q[qtail++] = c;
                    w->dist[c] = currLvl;
                    counter++;
                } else if (inB[c]) {
                    ws_visit(w, c, u);
                    q[qtail++] = c;
                    w->dist[c] = currLvl;
                }
            }
        }
        qhead = fEnd;
        if (qtail == qhead && counter != (long)(hi - lo)) {
            for (int i = prevIslandStop; i < hi; ++i) {
                if (w->dist[i] == -1) {
                    q[qtail++] = i;
                    counter++;
                    w->dist[i] = currLvl;
                    ws_visit(w, i, i);
                    prevIslandStop = i;
                    break;
                }
            }
            if (qtail == qhead) { status = -1; break; }
        }
    }
    int totalLevel = currLvl;
    if (status == 0) {
        int untouched = 0;
        for (int b = 0; b < nb; ++b)
            for (int i = blo[b]; i < bhi[b]; ++i)
                if (w->dist[i] == -1) {
                    w->dist[i] = totalLevel;
                    w->touched[w->ntouched++] = i;
                    untouched = 1;
                }
        if (untouched) totalLevel += 1;
        if (totalLevel < 1) totalLevel = 1;
        if (totalLevel > cap) status = -1;
    }
    if (status == 0) {
        for (int L = 0; L < totalLevel; ++L) { levelRow[L] = 0; levelNnz[L] = 0; }
        for (int i = lo; i < hi; ++i) {
            int d = w->dist[i], pr = totalPerm[i];
            levelRow[d] += 1;
            levelNnz[d] += rowPtr[pr + 1] - rowPtr[pr];
        }
        for (int b = 0; b < nb; ++b) {
            int *lr = bLevelRow + (size_t)b * cap, *ln = bLevelNnz + (size_t)b * cap;
            for (int L = 0; L < totalLevel; ++L) { lr[L] = 0; ln[L] = 0; }
            for (int i = blo[b]; i < bhi[b]; ++i) {
                int d = w->dist[i], pr = totalPerm[i];
                lr[d] += 1;
                ln[d] += rowPtr[pr + 1] - rowPtr[pr];
            }
        }
        /* stable sort of every region by level, composed into totalPerm */
        int *off = w->cnt, *perm = w->perm, *tmp = (int *)malloc(sizeof(int) * (size_t)n);
        for (int region = -1; region < nb; ++region) {
            const int a = region < 0 ? lo : blo[region], e = region < 0 ? hi : bhi[region];
            if (a >= e) continue;
            for (int L = 0; L <= totalLevel; ++L) off[L] = 0;
            for (int i = a; i < e; ++i) off[w->dist[i] + 1]++;
            for (int L = 0; L < totalLevel; ++L) off[L + 1] += off[L];
            for (int i = a; i < e; ++i) perm[a + off[w->dist[i]]++] = i;
            for (int i = a; i < e; ++i) tmp[i - a] = totalPerm[perm[i]];
            for (int i = a; i < e; ++i) {
                totalPerm[i] = tmp[i - a];
                totalInvPerm[totalPerm[i]] = i;
            }
        }
        free(tmp);
    }
    for (long t = 0; t < w->ntouched; ++t) {
        int c = w->touched[t];
        w->dist[c] = -1;
        w->parent[c] = -1;
    }
    w->ntouched = 0;
    free(inB);
    return status == 0 ? totalLevel : -1;
}

/* ------------------------------------------------------------------------ */
/* Kernel callback bodies (sequential over the given row range).             */
/* CRS arrays are the *permuted* matrix with the diagonal stored first.      */
/* Vectors with nrhs>1 are column-major: v[r + k*n].                         */
/* ------------------------------------------------------------------------ */
void or_symmspmv_range(int n, const int *rp, const int *col, const double *val,
                       const double *x, double *b, int start, int end, int nrhs)
{
    for (int k = 0; k < nrhs; ++k) {
        const double *xk = x + (size_t)k * n;
        double *bk = b + (size_t)k * n;
        for (int row = start; row < end; ++row) {
            double xr = xk[row];
            bk[row] += val[rp[row]] * xr;
            double t = 0;
            for (int idx = rp[row] + 1; idx < rp[row + 1]; ++idx) {
                double m = val[idx];
                int c = col[idx];
                t += m * xk[c];
                bk[c] += m * xr;
            }
            bk[row] += t;
        }
    }
}

void or_gs_range(int n, const int *rp, const int *col, const double *val,
                 double *x, const double *b, int start, int end, int nrhs)
{
    int inc = (start > end) ? -1 : 1;
    for (int k = 0; k < nrhs; ++k) {
        double *xk = x + (size_t)k * n;
        const double *bk = b + (size_t)k * n;
        for (int row = start; row != end; row += inc) {
            double t = bk[row];
            for (int idx = rp[row] + 1; idx < rp[row + 1]; ++idx)
                t += (-val[idx]) * xk[col[idx]];
            xk[row] = t / val[rp[row]];
        }
    }
}

void or_kacz_range(int n, const int *rp, const int *col, const double *val,
                   double *x, const double *b, int start, int end, int nrhs)
{
    int inc = (start > end) ? -1 : 1;
    for (int k = 0; k < nrhs; ++k) {
        double *xk = x + (size_t)k * n;
        const double *bk = b + (size_t)k * n;
        for (int row = start; row != end; row += inc) {
            double rowNorm = 0.0, scale = -bk[row];
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx) {
                double m = val[idx];
                scale += m * xk[col[idx]];
                rowNorm += m * m;
            }
            scale /= rowNorm;
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx)
                xk[col[idx]] = xk[col[idx]] - scale * val[idx];
        }
    }
}

void or_spmv_range(int n, const int *rp, const int *col, const double *val,
                   const double *x, double *b, int start, int end, int nrhs)
{
    for (int k = 0; k < nrhs; ++k) {
        const double *xk = x + (size_t)k * n;
        double *bk = b + (size_t)k * n;
        for (int row = start; row < end; ++row) {
            double t = 0;
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx) t += val[idx] * xk[col[idx]];
            bk[row] = t;
        }
    }
}

/* SPMTV_KERNEL_BODY examples/src/helpers/kernels.cpp:106-115: b += A^T x,
 * scattering row by row in call order */
void or_spmtv_range(int n, const int *rp, const int *col, const double *val,
                    const double *x, double *b, int start, int end, int nrhs)
{
    for (int k = 0; k < nrhs; ++k) {
        const double *xk = x + (size_t)k * n;
        double *bk = b + (size_t)k * n;
        for (int row = start; row < end; ++row) {
            double xr = xk[row];
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx) bk[col[idx]] += val[idx] * xr;
        }
    }
}

/* MAT_SPMV_KERNEL_only_highest examples/src/helpers/kernels.cpp:706-734, one
 * power at a time over all rows: x has 2*n entries, power p goes to slot
 * p%2 and reads slot (p-1)%2 */
void or_mpk_highest_serial(int n, const int *rp, const int *col, const double *val,
                           double *x, int k)
{
    for (int p = 1; p <= k; ++p) {
        const double *src = x + (size_t)((p - 1) % 2) * n;
        double *dst = x + (size_t)(p % 2) * n;
        for (int row = 0; row < n; ++row) {
            double t = 0;
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx) t += val[idx] * src[col[idx]];
            dst[row] = t;
        }
    }
}

/* x has (k+1)*n entries: x[p*n + row] = (A^p x0)[row]  (1-based powers) */
void or_mpk_serial(int n, const int *rp, const int *col, const double *val,
                   double *x, int k)
{
    for (int p = 1; p <= k; ++p) {
        const double *src = x + (size_t)(p - 1) * n;
        double *dst = x + (size_t)p * n;
        for (int row = 0; row < n; ++row) {
            double t = 0;
            for (int idx = rp[row]; idx < rp[row + 1]; ++idx) t += val[idx] * src[col[idx]];
            dst[row] = t;
        }
    }
}

/* Run `ncalls` leaf calls (start,end pairs) of a kernel in the given order. */
enum { OR_K_SYMMSPMV = 0, OR_K_GS = 1, OR_K_KACZ = 2, OR_K_SPMV = 3, OR_K_SPMTV = 5 };

void or_run_calls(int kind, int n, const int *rp, const int *col,
                  const double *val, double *x, double *b, int ncalls,
                  const int *calls, int nrhs)
{
    for (int c = 0; c < ncalls; ++c) {
        int s = calls[2 * c], e = calls[2 * c + 1];
        switch (kind) {
        case OR_K_SYMMSPMV: or_symmspmv_range(n, rp, col, val, x, b, s, e, nrhs); break;
        case OR_K_GS: or_gs_range(n, rp, col, val, x, b, s, e, nrhs); break;
        case OR_K_KACZ: or_kacz_range(n, rp, col, val, x, b, s, e, nrhs); break;
        case OR_K_SPMV: or_spmv_range(n, rp, col, val, x, b, s, e, nrhs); break;
        case OR_K_SPMTV: or_spmtv_range(n, rp, col, val, x, b, s, e, nrhs); break;
        default: break;
        }
    }
}

/* The leaves of each phase (phase_ptr[p] .. phase_ptr[p+1] into calls) run
 * concurrently on OpenMP threads, each thread a contiguous run of leaves
 * (RACE's threads own neighbouring zones), phases one after the other: how
 * RACE runs them on the host (recursiveCall, src/functionManager.cpp:199-265).  Leaves
 * of one phase are distance-k independent, so the result equals
 * or_run_calls over the same calls in order. */
void or_run_phases(int kind, int n, const int *rp, const int *col, const double *val, double *x,
                   double *b, int nphases, const int *phase_ptr, const int *calls, int nrhs,
                   int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    for (int p = 0; p < nphases; ++p) {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int c = phase_ptr[p]; c < phase_ptr[p + 1]; ++c)
            or_run_calls(kind, n, rp, col, val, x, b, 1, calls + 2 * c, nrhs);
    }
    (void)nthreads;
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
