/*
 * race_amd.h -- C ABI of the MI355X RACE engine.
 *
 * Drop-in boundary for RRZE-HPC/RACE's RACE::Interface (include/interface.h).
 * Every entry point returns an int status whose values equal the
 * reference's RACE_error enum (include/error.h:29-41); none of them exits the
 * process.  A handle is not thread-safe: use one host thread per handle.
 * All pointers are plain host pointers unless documented as device pointers.
 * The C++ facade include/race_amd_interface.hpp rebuilds RACE::Interface on
 * top of these functions.
 */
#ifndef RACE_AMD_H
#define RACE_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the reference's RACE_error, same names and values
 * (include/error.h:29-41) */
enum RACE_error {
    RACE_SUCCESS = 0,
    RACE_ERR_INVALID_ARG = 1,
    RACE_ERR_MATRIX_SYMM = 2,
    RACE_ERR_D2_COLOR = 3,
    RACE_ERR_D1_COLOR = 4,
    RACE_ERR_GRAPH_TRAVERSAL = 5,
    RACE_ERR_INCOMPATIBILITY = 6,
    RACE_ERR_HWLOC = 7,
    RACE_ERR_PIN = 8,
    RACE_ERR_INTERNAL = 9,
    RACE_ERR_NOT_IMPLEMENTED = 10
};
#ifndef __cplusplus
typedef enum RACE_error RACE_error;
#endif

/* RACE::dist, d2Method, LBTarget, PinMethod (include/type.h:31-55) */
enum race_dist { RACE_ONE = 1, RACE_TWO = 2, RACE_POWER = 3 };
enum race_d2_method { RACE_TWO_BLOCK = 0, RACE_THREE_BLOCK = 1 };
enum race_lb_target { RACE_LB_ROW = 0, RACE_LB_NNZ = 1 };
enum race_pin_method { RACE_FILL = 0, RACE_SCATTER = 1 };

typedef struct race_handle_s *race_handle;

/* Interface::Interface (include/interface.h:127, src/interface.cpp:33-36).
 * rowPtr/col are copied (the reference borrows them).  RACE_MAX_DENSE (read
 * here, as the reference's Graph does, src/graph_SoA.cpp:321-435): rows
 * longer than that many times nnz/nrow (integer division) leave the graph
 * and run last, one after the other, in every sweep (a last phase of one
 * leaf, [n', n) of the permutation); colouring only (RACE_ERR_NOT_IMPLEMENTED
 * for POWER, whose executor has no serial part).  Device kernels that write
 * the columns of their rows (Kaczmarz, colour-phase SymmSpMV) are then
 * refused with RACE_ERR_INCOMPATIBILITY: concurrent rows may share a dense
 * column. */
int race_create(race_handle *out, int nrow, int nthreads, int dist, const int *rowPtr,
                const int *col, int symm_hint, int smt, int pin_method, const int *init_perm,
                const int *init_inv_perm, int d2_method, int lb_target);
/* Interface::~Interface (src/interface.cpp:38-60) */
int race_destroy(race_handle h);

/* Interface::RACEColor() -- colouring preprocessing (src/interface.cpp:260-301).
 * RACE_EFFICIENCY and RACE_THREADS are read from the environment like the
 * reference (src/level_recursion.cpp:62-70). */
int race_color(race_handle h);
/* Interface::RACEColor(highestPower, cacheSize, safetyFactor, mtxType,
 * highestSubPower) -- MPK preprocessing (src/interface.cpp:205-257).
 * cache_size is in bytes; mtx_type is 'N', 'L' or 'U'. */
int race_color_power(race_handle h, int highest_power, double cache_size, double safety,
                     char mtx_type, int highest_sub_power);

/* Interface::getPerm / getInvPerm (src/interface.cpp:323-375): perm[new]=old,
 * copied into caller buffers of length len (== nrow). */
int race_get_perm(race_handle h, int *perm, int len);
int race_get_inv_perm(race_handle h, int *inv_perm, int len);
/* Interface::getNumThreads / getEfficiency / getMaxStageDepth /
 * getHighestPower (src/interface.cpp:304-321, 377-380, 1054-1057) */
int race_get_num_threads(race_handle h, int *out);
int race_get_efficiency(race_handle h, double *out);
int race_get_max_stage_depth(race_handle h, int *out);
int race_get_highest_power(race_handle h, int *out);

/* ---- host callbacks (drop-in for registerFunction with CPU kernels) ---- */
typedef void (*race_color_fn)(int start, int end, void *args);
typedef void (*race_power_fn)(int start, int end, int pow, int sub_pow, int numa, void *args);
/* Interface::registerFunction(void(*)(int,int,void*), void*)
 * (include/interface.h:158, src/interface.cpp:397-420).  start>end asks for
 * reverse order, as in the reference.  Host callbacks run on OpenMP threads:
 * the leaves of one colour phase in parallel (min(RACE threads, OpenMP
 * threads) of them), phases in order; power callbacks on row slices of the
 * independent (level group, power) units of a wavefront step. */
int race_register_function(race_handle h, race_color_fn f, void *args, int *id);
/* Interface::registerFunction(void(*)(int,int,int,int,int,void*), void*, power,
 * subPower, numaSplit) (include/interface.h:163, src/interface.cpp:435-458) */
int race_register_power_function(race_handle h, race_power_fn f, void *args, int power,
                                 int sub_power, int numa_split, int *id);
/* Interface::attachCommunicationToFunction (include/interface.h:161-165,
 * src/interface.cpp:487-533): a halo-exchange callback comm(pow, subPow,
 * args) for a POWER function (RACE_ERR_NOT_IMPLEMENTED for colouring
 * handles).  The reference calls it only in its MPI build
 * (RACE_ENABLE_MPI_MPK, src/functionManager.cpp:963-1067); on one device
 * there are no remote rows, so it is stored and never called, as in the
 * reference's non-MPI build.  The _simple form comm(pow, args) needs
 * highestSubPower == 1 (else RACE_ERR_INCOMPATIBILITY). */
typedef void (*race_comm_fn)(int pow, int sub_pow, void *args);
typedef void (*race_comm_simple_fn)(int pow, void *args);
int race_attach_communication(race_handle h, int id, race_comm_fn f, void *args);
int race_attach_communication_simple(race_handle h, int id, race_comm_simple_fn f, void *args);

/* ---- device kernels (the GPU path) ---- */
enum race_kernel_kind {
    RACE_K_SYMMSPMV = 0, /* b += A x, upper triangle stored, diagonal first
                            (SYMM_SPMV_KERNEL examples/src/helpers/kernels.cpp:403-427) */
    RACE_K_GS = 1,       /* Gauss-Seidel sweep (GS_KERNEL kernels.cpp:174-199) */
    RACE_K_KACZ = 2,     /* Kaczmarz sweep (KACZ_KERNEL kernels.cpp:285-316) */
    RACE_K_SPMV = 3,     /* b = A x (SPMV_KERNEL kernels.cpp:5-25) */
    RACE_K_MPK = 4,      /* x_p = A x_{p-1}, p=1..power (MAT_SPMV_KERNEL kernels.cpp:749-777) */
    RACE_K_SPMTV = 5,    /* b += A^T x, full matrix stored (SPMTV_KERNEL kernels.cpp:106-124);
                            needs no colouring: the transpose is formed at registration */
    RACE_K_MPK_HIGHEST = 6 /* MPK keeping two vector slots: power p in x + (p%2)*nrow, x holds
                              2*nrow entries, x_0 in slot 0 (MAT_SPMV_KERNEL_only_highest
                              kernels.cpp:706-734) */
};
enum race_precision { RACE_FP64 = 0, RACE_FP32 = 1, RACE_BF16 = 2 };

typedef struct {
    int kind;            /* race_kernel_kind */
    int nrow;
    const int *rowPtr;   /* host CRS of the permuted matrix, diagonal first in each row */
    const int *col;
    const double *val;
    int val_precision;   /* storage precision of the matrix values on the device */
    int vec_precision;   /* RACE_FP64 or RACE_FP32 for x and b; pairs: FP64/FP64,
                            FP32/FP32, BF16/FP32 (FP32 accumulation) */
    int nrhs;            /* right-hand sides, vectors are column-major v[r + k*nrow] */
    void *x;             /* device pointer */
    void *b;             /* device pointer (unused for MPK) */
    int power;           /* MPK: number of powers; x holds (power*sub_power+1)*nrow
                            entries (RACE_K_MPK_HIGHEST: 2*nrow) */
    int sub_power;       /* MPK: sub-powers per power (0 or 1: none).  The reference's
                            power executor calls the kernel for every (pow, subPow) pair
                            in order (registerFunction(..., power, subPower),
                            include/interface.h:153-157, src/interface.cpp:435-458); on
                            the device each pair is one product with A, so product t =
                            (pow-1)*sub_power + subPow lands in slot t of x */
} race_kernel_desc;

/* Typed device registration: the GPU counterpart of registerFunction; the
 * matrix is uploaded and laid out once, the vectors stay caller-owned. */
int race_register_device(race_handle h, const race_kernel_desc *desc, int *id);
/* Layout options, applied to functions registered afterwards:
 *   "layout"       0 = leaf bundles (execution-order SELL, LDS-resident waves),
 *                  1 = one workgroup per leaf (CRS),
 *                  2 = leaf bundles only (registration fails with
 *                      RACE_ERR_INCOMPATIBILITY if a leaf does not fit)
 *   "bundle_rows"  rows per bundle (default 128)
 *   "bundle_local" LDS columns (rows + outside columns) per bundle (default 6144)
 *   "xcd_order"    1 = consecutive bundles of a phase run on the same XCD (default 0)
 *   "bundle_threads" workgroup size of the bundle kernels: 64 (one wavefront), 128,
 *                  256 (default), 512 or 1024
 *   "bundle_tpr"   lanes per matrix row in the bundle kernels: 1, 2, 4 (0 = auto:
 *                  from the longest row; rows longer than 64 entries select
 *                  variable lanes per row, 16 entries per lane)
 *   "symm_mode"    SymmSpMV: 1 = colour phases (CRS, one thread per row),
 *                  2 = one launch over a SELL-C upper triangle, FP64 atomics,
 *                  3 = one launch over the expanded full matrix (default, 0)
 *   "graph"        1 = executeFunction replays the function's launch sequence
 *                  (one kernel per colour phase or wavefront step) as a HIP
 *                  graph, captured once per sweep direction, vectors and
 *                  power (default 0)
 *   "bfs_block_max" ranges of at most this many rows are traversed one
 *                  workgroup each, batched (default 131072; 0 = every range
 *                  level by level over the whole device); set before colouring
 *   "flow"         1 = one launch per sweep (Gauss-Seidel with one
 *                  right-hand side and 256-thread workgroups; Kaczmarz with
 *                  256-1024-thread workgroups): the chains of all colour
 *                  phases are handed out in phase order and each waits only
 *                  for the earlier chains it conflicts with (per-chain
 *                  completion flags) instead of a launch boundary per phase
 *                  (default 0); set before registering; results identical
 * No counterpart in the reference (its executor has no layout choice). */
int race_set_option(race_handle h, const char *key, int value);
/* re-point a registered device function at other vectors */
int race_set_vectors(race_handle h, int id, void *x, void *b);

/* FP64 iterative refinement around a registered Gauss-Seidel or Kaczmarz
 * function (no counterpart in the reference).  a64: the FP64 matrix (full
 * CRS, diagonal first, permuted) and the FP64 device vectors x and b.  The
 * inner function's vectors are replaced by internal workspace; one step is
 * r = b - A x (FP64), inner_sweeps forward+backward sweeps on A d = r in the
 * inner function's precision from d = 0, x += d (FP64).  race_execute(id)
 * runs one step. */
int race_register_refinement(race_handle h, const race_kernel_desc *a64, int inner_id,
                             int inner_sweeps, int *id);
/* steps until ||b - A x|| / ||b|| <= tol or max_outer steps; reports the
 * steps taken, the inner sweeps run and the final relative residual */
int race_refine(race_handle h, int id, double tol, int max_outer, int *outer, long *sweeps,
                double *relres);
/* relative residual ||b - A x|| / ||b|| of a refinement function's vectors */
int race_residual(race_handle h, int id, double *relres);
/* race_refine without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The residual, the
 * correction and the norms live in one static device array of 64 MiB per
 * GPU (race_refine_noalloc_max_bytes), held until the process ends: the
 * workspace needs 256 + 2 * nrow * nrhs * (8 for FP64 inner sweeps, 4 for
 * FP32 / BF16) bytes (each vector rounded up to 256), and larger problems
 * give RACE_ERR_INVALID_ARG.  Calls on one GPU share that array: calls that
 * overlap on one GPU (other threads, other streams) race.  Results equal
 * race_refine's. */
int race_refine_noalloc(race_handle h, int id, double tol, int max_outer, int *outer, long *sweeps,
                        double *relres);
int race_refine_noalloc_max_bytes(size_t *bytes);

/* Interface::executeFunction (src/interface.cpp:541-572): synchronous. */
int race_execute(race_handle h, int id, int rev);
/* Asynchronous variant on a caller HIP stream (hipStream_t, may be NULL for
 * the handle's own stream). */
int race_execute_async(race_handle h, int id, int rev, void *stream);
/* Interface::setPower/getPower (src/interface.cpp:574-592) */
/* Interface::setPower (src/interface.cpp:574-577).  A device MPK runs powers
 * 1..power; power may not exceed the power it was registered with. */
int race_set_power(race_handle h, int id, int power);
int race_get_power(race_handle h, int id, int *power);
/* Interface::setSubPower / getSubPower (include/interface.h:168-170,
 * src/interface.cpp:579-592): sub-powers per power of a POWER function; a
 * device MPK accepts power * sub_power up to its registered products. */
int race_set_sub_power(race_handle h, int id, int sub_power);
int race_get_sub_power(race_handle h, int id, int *sub_power);
/* Interface::setSerial / unsetSerial (include/interface.h:171-172,
 * src/interface.cpp:594-602, FuncManager::setSerial
 * src/functionManager.cpp:1561-1580): serial != 0 runs a host callback
 * function on one thread, 0 on the handle's threads again.  Device
 * functions have no host threads (accepted, no effect). */
int race_set_serial(race_handle h, int id, int serial);
/* Interface::resetTime (src/interface.cpp:649-652) clears the zone tree's
 * per-leaf timers; no timers are kept here: accepted, no effect. */
int race_reset_time(race_handle h);
/* Interface::sleep (src/interface.cpp:817-822) parks RACE's own thread pool,
 * which exists only without RACE_KERNEL_THREAD_OMP; host callbacks run on
 * OpenMP threads here, as in the reference's default build: no effect. */
int race_sleep(race_handle h);
/* Interface::pinThread (src/interface.cpp:845-848): pins the calling thread
 * to the (thread_id * smt mod n)-th CPU of the process's affinity mask
 * (RACE_ERR_PIN when the OS refuses). */
int race_pin_thread(race_handle h, int thread_id);
/* Interface::numaInitRowPtr / numaInitMtxVec, both overloads
 * (src/interface.cpp:852-1040): first-touch initialisation through a power
 * callback registered with power 1 and executed once (it takes a function id,
 * as in the reference): rowPtr[r] = rowPtr[r+1] = 0; val, col of each row and
 * x[r], x[nrow + r] = 0; the _local forms index domain 0 of the per-domain
 * arrays (one device = one domain).  POWER handles only
 * (RACE_ERR_NOT_IMPLEMENTED otherwise, "NUMA not implemented for
 * coloring"). */
int race_numa_init_row_ptr(race_handle h, int *rowPtr);
int race_numa_init_row_ptr_local(race_handle h, int **rowPtr);
int race_numa_init_mtx_vec(race_handle h, int *rowPtr, int *col, double *val, double *x, int power);
int race_numa_init_mtx_vec_local(race_handle h, int **rowPtr, int **col, double **val, int power);
/* Interface::getNumaSplitting (src/interface.cpp:1042-1052): row boundaries
 * of the NUMA domains; one device is one domain, [0, nrow] (len 2). */
int race_get_numa_splitting(race_handle h, int *split, int cap, int *len);
/* Interface::tuneFunction (src/interface.cpp:605-647): best power by timing */
int race_tune_function(race_handle h, int id, int rev, int *best_power);
/* Interface::checkPowerCoverage (src/interface.cpp:1060-1095): returns
 * RACE_SUCCESS when every (row, power) is computed exactly once. */
int race_check_power_coverage(race_handle h);

/* ---- introspection ---- */
/* flattened colour schedule: phase_ptr[nphases+1], leaf_lo/hi[nleaves] */
int race_get_schedule(race_handle h, int rev, int *nphases, int *nleaves, int *phase_ptr,
                      int *leaf_lo, int *leaf_hi, int cap);
/* zone tree as flat arrays (leaf i: valueZ[0], valueZ[1], parent, nthreads, stage, effRow) */
int race_get_tree(race_handle h, int *nleaves, int *data6, int cap);
/* MPK level groups: group_ptr[ngroups+1] (row offsets) */
int race_get_mpk_groups(race_handle h, int *ngroups, int *group_ptr, int cap);
/* the recursion stage of every hopeless group (mtxPowerRecursive's stage-1
 * leaves, src/matrixPowerRecursive.cpp:156-338): per child lo, hi (rows),
 * its stage-0 group, and the number of its own level groups */
int race_get_mpk_children(race_handle h, int *nchildren, int *data4, int cap);

/* ---- initial permutation ----
 * Reverse Cuthill-McKee order of a symmetric pattern, perm[new] = old, the
 * initial permutation the reference's harness passes to the Interface
 * constructor (sparsemat::doRCM examples/src/helpers/sparsemat.cpp:715-760,
 * via SpMP's CSR::getRCMPermutation, which is not in the reference tree):
 * per component a George-Liu pseudo-peripheral root, neighbours visited in
 * increasing degree.  RACE_ERR_MATRIX_SYMM for an unsymmetric pattern.
 * Host only. */
int race_rcm(int n, const int *rowPtr, const int *col, int *perm, int *inv_perm);

/* ---- schedule save / load ----
 * No counterpart in the reference, which rebuilds its zone tree and MPK tree
 * on every RACEColor call and can only print them (ZoneTree::printTree
 * src/zone_tree.cpp:307-339, mtxPowerRecursive::printTree
 * src/matrixPowerRecursive.cpp:447-490).  The byte string holds what
 * RACEColor computed: permutation, zone tree (colouring) or level groups
 * (POWER), thread count, and a hash of the matrix pattern it was built on.
 * race_schedule_save: buf NULL returns the size only; needs a coloured handle
 * (RACE_ERR_NOT_IMPLEMENTED when RACE_MAX_DENSE moved rows to a serial tail).
 * race_schedule_load: instead of RACEColor, on a handle created with the
 * same matrix pattern, nrow, nthreads, dist, d2 method and LB target, and
 * before any function is registered; anything else, a damaged byte string
 * (checksum) or one whose contents are out of range, gives
 * RACE_ERR_INVALID_ARG.
 * race_schedule_info: checks a byte string without a handle. */
int race_schedule_save(race_handle h, void *buf, size_t cap, size_t *size);
int race_schedule_load(race_handle h, const void *buf, size_t size);
int race_schedule_info(const void *buf, size_t size, int *nrow, int *dist, int *nthreads);
/* dependency waves of a registered GS/KACZ function (critical path length) */
int race_get_wave_count(race_handle h, int id, int rev, int *waves);
/* kernel launches one execution of a registered device function issues
 * (colour phases, MPK wavefront steps) */
int race_get_launch_count(race_handle h, int id, int rev, int *launches);
/* one-launch ("flow") sweeps of a device function whose dependency waits gave
 * up after their poll budget (always 0 for a consistent schedule; a nonzero
 * count means those sweeps' results are invalid).  Synchronises with the
 * device.  No counterpart in the reference. */
int race_get_flow_aborts(race_handle h, int id, int *aborts);
/* HIP-event timer on the handle's stream (brackets work launched there) */
int race_timer_begin(race_handle h);
int race_timer_end(race_handle h, float *ms);
/* seconds spent in GPU traversals during preprocessing */
int race_get_preprocess_stats(race_handle h, double *bfs_ms, int *traversals);

/* ---- host-only scheduling logic (no GPU needed) ---- */
/* LB::splitZones on a level histogram (src/lb.cpp:188-667).  Outputs rows per
 * zone (blocks), nested thread count per thread, level boundaries. */
int race_lb_split(const int *level_row, const int *level_nnz, int nlevels, int nthreads,
                  double efficiency, int dist, int d2_method, int lb_target, int cap,
                  int *nblocks, int *zone_rows, int *level_ptr, int *nthreads_out, int *scale);

/* The thread split of LevelRecursion::calculateIdealNthreads
 * (src/level_recursion.cpp:149-263) for one partitioned range: weights[i]
 * is the nnz (or row count) of child group i, total the range's; out[i] =
 * the ideal thread count of group i. */
int race_split_threads(const int *weights, int ngroups, int total, int parent_threads, int *out);

/* Hopeless regions of the MPK level grouping
 * (mtxPower::identifyHopelessRegions src/matrixPower.cpp:147-502, one node):
 * violated = ascending levels over the cache by more than the cut-off; out
 * receives [start, end] level pairs (inclusive), *nout entries. */
int race_mpk_hopeless_regions(const int *violated, int nviolated, int k, int total_level, int *out, int cap,
                              int *nout);
/* mtxPower::consolidatePartition (src/matrixPower.cpp:685-880, one node):
 * level_ptr[nlevels+1], elem[nlevels] (safety * (k+1) * nnz per level),
 * regions from race_mpk_hopeless_regions; groups[ngroups+1] row pointers,
 * glevel[ngroups+1] first levels, hid[nhid] the group of each region. */
int race_mpk_consolidate(const int *level_ptr, const double *elem, int nlevels, int k, double cache_elem,
                         const int *regions, int nregions, int *groups, int *glevel, int *hid, int cap,
                         int *ngroups, int *nhid);

/* ---- batches: one function on many matrices (no counterpart in the
 * reference) ----
 * Every matrix keeps its own handle (colouring), CRS and device vectors; a
 * batch runs them together: Gauss-Seidel / Kaczmarz as one fused launch per
 * colour-phase index over the bundles of all matrices, SpMV / SymmSpMV as
 * one launch over the SELL chunks of all matrices.  All matrices of a batch
 * share kind, precisions and nrhs.  Handles must stay alive while the batch
 * is built (race_batch_build, race_batch_autotune, first execution). */
typedef struct race_batch_s *race_batch;
int race_batch_create(race_batch *out, int kind, int val_precision, int vec_precision, int nrhs);
/* d: like race_register_device (host arrays are copied) */
int race_batch_add(race_batch b, race_handle h, const race_kernel_desc *d);
/* "bundle_threads", "bundle_rows", "bundle_tpr" (0, 1, 2, 4),
 * "bundle_local" (GS / KACZ); "sell_c" (8, 16, 32, 64), "sell_sigma"
 * (row-sorting window, 1 = none) (SpMV / SymmSpMV); "flow" (Kaczmarz,
 * 256-1024-thread workgroups): 1 = one launch per sweep and matrix-shape
 * group, every matrix advancing through its own colour phases with per-chain
 * dependency flags instead of one fused launch per phase index; the groups'
 * launches run concurrently (the group with the longest dependency path on
 * the given stream, the others on streams of the batch, joined back into it
 * by events before the call's work ends) (default 0; results identical) */
int race_batch_set_option(race_batch b, const char *key, int value);
int race_batch_get_option(race_batch b, const char *key, int *value);
/* kernel launches of one execution of the built batch; one-launch sweeps
 * whose dependency waits gave up (always 0 for consistent schedules;
 * synchronises with the device) */
int race_batch_get_launch_count(race_batch b, int rev, int *launches);
int race_batch_get_flow_aborts(race_batch b, int *aborts);
int race_batch_build(race_batch b);
int race_batch_execute(race_batch b, int rev);
int race_batch_execute_async(race_batch b, int rev, void *stream);
/* average milliseconds of one execution over reps (after one warm-up) */
int race_batch_time(race_batch b, int rev, int reps, float *ms);
/* the layout tuner: 12 candidate layouts per kernel family; race_batch_candidate
 * describes candidate index (bundles: threads, rows, lanes per row; SELL: C,
 * sigma, 0); race_batch_autotune times every candidate (times[i] = -1 where a
 * layout does not fit), keeps the fastest and reports its index */
int race_batch_candidates(int kind, int *count);
int race_batch_candidate(int kind, int index, int *p0, int *p1, int *p2);
int race_batch_use_candidate(race_batch b, int index);
int race_batch_autotune(race_batch b, int rev, int reps, int *best, float *times, int cap);
int race_batch_destroy(race_batch b);

/* ---- the layout tuner's model (no counterpart in the reference) ----
 * A 16-64-64-12 multilayer perceptron (ReLU hidden layers) mapping 16 matrix
 * features to the predicted relative log-time of the 12 candidate layouts
 * of race_batch_candidate.  Parameters: W1 [16][64], c1 [64], W2 [64][64],
 * c2 [64], W3 [64][12], c3 [12], row-major, concatenated.  x: batch x 16,
 * y/t/mask: batch x 12 (host arrays). */
typedef struct race_mlp_s *race_mlp;
int race_mlp_dims(int *n_in, int *n_hidden, int *n_out, int *n_params);
/* The model's 16 input features of a square pattern (host only, no GPU):
 * log10 n, log10 nnz, mean row length, its coefficient of variation, log10
 * longest row, shortest row, share of rows longer than twice the mean and
 * than 32, mean and largest |col - row| / n, share of entries within 64 of
 * the diagonal, log10 of the number of levels of a breadth-first search
 * from row 0 (parts restarted at their smallest row), largest level / n,
 * mean / largest level, Gini coefficient of the row lengths, log10(nnz /
 * levels). */
int race_matrix_features(int n, const int *rowPtr, const int *col, double *out16);
int race_mlp_create(race_mlp *out, unsigned seed);
int race_mlp_set_params(race_mlp m, const float *params, int count);
int race_mlp_get_params(race_mlp m, float *params, int count);
/* precision RACE_FP32 or RACE_BF16 (BF16 weights and activations on the
 * matrix cores, FP32 accumulation) */
int race_mlp_forward(race_mlp m, const float *x, int batch, float *y, int precision);
/* gradients of sum(mask (y - t)^2) / sum(mask) with respect to the
 * parameters (FP32), and that loss */
int race_mlp_gradients(race_mlp m, const float *x, const float *t, const float *mask, int batch,
                       float *grads, float *loss);
/* epochs full-batch Adam steps (beta 0.9 / 0.999); loss of the last step */
int race_mlp_train(race_mlp m, const float *x, const float *t, const float *mask, int batch, int epochs,
                   float lr, float *loss);
int race_mlp_destroy(race_mlp m);

/* ---- SELL-C-sigma helpers ----
 * Entry j of row r of a SELL-C-sigma matrix sits at
 * chunk_start[r / C] + r % C + j * C; row_len[r] is the row length,
 * chunk_len[c] the (padded) length of chunk c.  val_precision is RACE_FP64
 * (val is double*) or RACE_FP32 (float*). */
/* sell_c_sigmize (src/sell_c_sigmize.cpp:48-62): padding slot j >= row_len[r]
 * gets column start + (r - start) % simd_width, start = first row of the
 * colouring leaf that holds r (the reference runs it as a RACE callback).
 * Needs a coloured handle (RACEColor()). */
int race_sell_c_sigmize(race_handle h, int simd_width, int C, int *col, const int *chunk_start,
                        const int *row_len, const int *chunk_len);
/* Interface::simdify (src/interface.cpp:824-832, include/simdify.h:70-357):
 * sell_c_sigmize, then entries are swapped inside rows so that no two rows of
 * a SIMD group (simd_width consecutive rows of a chunk) hold the same column
 * at one position.  RACE_ERR_INCOMPATIBILITY where that is impossible (the
 * reference returns false). */
int race_simdify(race_handle h, int simd_width, int C, int nrows, int *col, const int *chunk_start,
                 const int *row_len, const int *chunk_len, void *val, int val_precision, int diag_first);
/* the rearrangement of race_simdify without the padding step (host only) */
int race_simdify_resolve(int simd_width, int C, int nrows, int *col, const int *chunk_start,
                         const int *chunk_len, void *val, int val_precision, int diag_first);
/* Interface::simdifyD1 (src/interface.cpp:834-842, include/simdify.h:389-447):
 * the entries of each row on the rows of preceding lanes of its SIMD group
 * are swapped to the end of the row (host only). */
int race_simdify_d1(int simd_width, int C, int nrows, int *col, const int *chunk_start,
                    const int *chunk_len, void *val, int val_precision, int dist2_compatible);

/* ---- device memory helpers (for callers without a HIP toolchain) ---- */
int race_device_malloc(void **p, size_t bytes);
int race_device_free(void *p);
int race_memcpy_h2d(void *dst, const void *src, size_t bytes);
int race_memcpy_d2h(void *dst, const void *src, size_t bytes);
int race_memset_device(void *dst, int value, size_t bytes);
int race_device_synchronize(void);
/* streaming device-to-device copy by a plain HIP kernel (bytes % 16 == 0) */
int race_device_copy(void *dst, const void *src, size_t bytes);
/* the same copy enqueued on a HIP stream (NULL: the null stream), no wait */
int race_device_copy_async(void *dst, const void *src, size_t bytes, void *stream);
/* the handle's HIP stream (where race_execute and the timer run) */
int race_get_stream(race_handle h, void **stream);
int race_device_count(int *count);

/* error text, RACE_error_string (src/error.cpp:26-60) */
const char *race_error_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* RACE_AMD_H */
