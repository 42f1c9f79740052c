/*
 * hipfoam.h -- C-ABI of the MI355X (gfx950) LDU Krylov library behind the
 * hipPCG / hipPBiCGStab OpenFOAM solvers.
 *
 * Plain C: opaque context handle, plain pointers and sizes, int status codes
 * (0 = ok).  Host pointers are caller-owned and copied; device state is
 * library-owned.  A context is not thread-safe; every export restores the
 * caller's floating-point environment on return (FP traps are masked while
 * the HIP runtime runs).  Entry points whose names end in _device take HIP
 * device pointers and leave data resident in HBM.
 *
 * Each declaration cites the reference interface it replaces
 * (scotCW/metal4Foam; B = src/metal4Foam/backend/metalFoamBackend,
 *  S = src/metal4Foam/solvers/metalFoamSolver).
 */
#ifndef HIPFOAM_H
#define HIPFOAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPFOAM_ABI_VERSION 1

/* status codes */
#define HIPFOAM_OK 0
#define HIPFOAM_ERR_ARG 1      /* bad argument (NULL, size, order)      */
#define HIPFOAM_ERR_DEVICE 2   /* HIP runtime error / no usable device  */
#define HIPFOAM_ERR_STATE 3    /* call out of order (no topology, ...)  */
#define HIPFOAM_ERR_SIZE 4     /* size beyond a documented maximum      */

/* solver selectors */
#define HIPFOAM_PCG 0          /* metalPCG -> hipPCG                    */
#define HIPFOAM_PBICGSTAB 1    /* metalPBiCGStab -> hipPBiCGStab        */

/* precision selectors */
#define HIPFOAM_FP32 0
#define HIPFOAM_FP64 1

/* topology checks of hipfoam_set_topology (hipfoam_set_topology_check) */
#define HIPFOAM_TOPO_POINTERS 0 /* sizes + lowAddr/uppAddr pointers      */
#define HIPFOAM_TOPO_CONTENT 1  /* sizes + hash of every entry (default) */

typedef struct hipfoam_context hipfoam_context;

/* Inner-solve statistics: metalFoamBackend::Stats (B.H:41-50). */
typedef struct
{
    int32_t nIterations;   /* inner FP32 iterations performed          */
    double finalResSum;    /* FP32 sum|r| at exit (unnormalised)       */
    int32_t converged;     /* reached targetResSum                     */
    int32_t singular;      /* NaN/Inf in the scalars (breakdown)       */
    int32_t stagnated;     /* 50 iterations without improvement        */
    double tGpuSeconds;    /* GPU time of the iteration loop           */
    double tWallSeconds;   /* wall time inside the iteration loop      */
} hipfoam_stats;

/* fvSolution controls of one solve: the lduMatrix::solver controls plus the
 * "hip" sub-dictionary (mirror of the reference's "metal" sub-dictionary,
 * S.C:32-52, 170-171, 224-227, 259-263). */
typedef struct
{
    double tolerance;          /* default 1e-6                            */
    double relTol;             /* default 0                               */
    int32_t maxIter;           /* default 1000 (caps total inner iters)   */
    int32_t minIter;           /* default 0                               */
    int32_t deflate;           /* 0 off, 1 on, 2 auto (default), 3 pinned:
                                  auto, with the inner right-hand side
                                  projected along u = A 1 instead of the
                                  constant (no reference counterpart; lets
                                  right-hand sides with sum != 0 on pinned
                                  Neumann matrices converge)               */
    int32_t chebyshevOrder;    /* default 8 (PCG); 1 = Jacobi             */
    double chebyshevRatio;     /* default 30                              */
    int32_t maxRestarts;       /* default 8                               */
    int32_t iterationsPerSync; /* default 1                               */
    int32_t verbose;           /* per-restart trace on stdout             */
} hipfoam_controls;

/* Result of one solve: OpenFOAM solverPerformance + profile (S.C:127-129,
 * 363-365, 417-426). */
typedef struct
{
    double initialResidual;
    double finalResidual;
    double normFactor;
    int32_t nIterations;  /* total inner iterations                     */
    int32_t restarts;     /* index of the last outer restart            */
    int32_t converged;
    int32_t singular;
    int32_t deflated;     /* constant-mode deflation was used           */
    int32_t polished;     /* switched to the exact operator             */
    double chebBeta;      /* Gershgorin bound used by Chebyshev         */
    double tConvert;      /* seconds: coefficient staging               */
    double tInner;        /* seconds: inner Krylov loops (wall)         */
    double tGpu;          /* seconds: inner Krylov loops (GPU events)   */
    double tTotal;        /* seconds: whole call                        */
} hipfoam_performance;

/* Fills the reference's defaults (S.C:170-263). */
void hipfoam_default_controls(hipfoam_controls* c);

/* ---- context -------------------------------------------------------- */

/* metalFoamBackend::metalFoamBackend (B.C:422-430).  device = HIP ordinal.
 * Always returns a context (check hipfoam_good); NULL only on OOM. */
int32_t hipfoam_create(int32_t device, hipfoam_context** out);
/* metalFoamBackend::~metalFoamBackend (B.C:433). */
void hipfoam_destroy(hipfoam_context* ctx);
/* metalFoamBackend::good (B.H:105-106). */
int32_t hipfoam_good(const hipfoam_context* ctx);
/* metalFoamBackend::description (B.H:108-109). */
const char* hipfoam_description(const hipfoam_context* ctx);
/* Message of the last failing call on this context ("" if none). */
const char* hipfoam_last_error(const hipfoam_context* ctx);

/* ---- backend (inner) interface: metalFoamBackend -------------------- */

/* metalFoamBackend::setTopology (B.C:454-519).  Uploads int32 addressing
 * once per mesh.  With HIPFOAM_TOPO_CONTENT (the default) it is a no-op when
 * nCells, nFaces and a 64-bit hash of every entry of the five arrays are
 * unchanged: addressing rewritten in place at the same addresses is uploaded
 * again, the same mesh at new addresses is not.  The hash reads the whole
 * addressing on every call (about 22 bytes per cell of a hexahedral mesh,
 * split over up to 8 host threads).  With HIPFOAM_TOPO_POINTERS it is a
 * no-op when (nCells, nFaces, lowAddr, uppAddr) are unchanged, exactly as
 * the reference (B.C:469-477), which reuses stale addressing after an
 * in-place rewrite. */
int32_t hipfoam_set_topology(hipfoam_context* ctx, int32_t nCells,
                             int32_t nFaces, const int32_t* lowAddr,
                             const int32_t* uppAddr, const int32_t* ownStart,
                             const int32_t* losort,
                             const int32_t* losortStart);

/* Selects how hipfoam_set_topology decides that the mesh is unchanged:
 * HIPFOAM_TOPO_CONTENT (default) or HIPFOAM_TOPO_POINTERS (the reference's
 * key, B.C:469-477, for callers that replace the addressing arrays whenever
 * the mesh changes and want no per-call hash).  No reference counterpart. */
int32_t hipfoam_set_topology_check(hipfoam_context* ctx, int32_t mode);

/* metalFoamBackend::setCoeffs (B.C:522-565): FP64 -> FP32 operator of the
 * inner solve, rD = float(1/diag).  lower == NULL selects the symmetric
 * path.  *seconds (may be NULL) receives the staging time. */
int32_t hipfoam_set_coeffs(hipfoam_context* ctx, const double* diag,
                           const double* upper, const double* lower,
                           double* seconds);

/* metalFoamBackend::pcg (B.C:568-732): FP32 PCG on A delta = r, delta0 = 0,
 * Chebyshev-Jacobi (chebOrder > 1) or Jacobi. */
int32_t hipfoam_pcg(hipfoam_context* ctx, const double* r,
                    double targetResSum, int32_t maxIter,
                    int32_t iterationsPerSync, int32_t chebOrder,
                    double chebBeta, double chebRatio, hipfoam_stats* stats);

/* metalFoamBackend::bicgstab (B.C:735-959): FP32 Jacobi BiCGStab. */
int32_t hipfoam_bicgstab(hipfoam_context* ctx, const double* r,
                         double targetResSum, int32_t maxIter,
                         hipfoam_stats* stats);

/* metalFoamBackend::delta (B.C:448-451): copies the FP32 correction of the
 * last pcg/bicgstab call into out[nCells]. */
int32_t hipfoam_delta(hipfoam_context* ctx, float* out);

/* ---- solver (outer) interface: metalFoamSolver::solveImpl ------------ */

/* metalFoamSolver::solveImpl (S.C:77-437) with every FP64 step on the GPU:
 * initial residual and normFactor, sumA, deflation, Gershgorin bound, FP32
 * inner solves, psi += delta, true FP64 residual, constant-mode correction,
 * hybrid polish.  solver = HIPFOAM_PCG | HIPFOAM_PBICGSTAB.  lower == NULL
 * for symmetric matrices.  psi is updated in place.  residual (may be NULL)
 * receives the final FP64 residual b - A psi (setResidualField, S.C:429). */
int32_t hipfoam_solve(hipfoam_context* ctx, int32_t solver,
                      const double* diag, const double* upper,
                      const double* lower, double* psi, const double* source,
                      const hipfoam_controls* controls,
                      hipfoam_performance* perf, double* residual);

/* Same, with every array a device pointer (data stays in HBM). */
int32_t hipfoam_solve_device(hipfoam_context* ctx, int32_t solver,
                             const double* diag, const double* upper,
                             const double* lower, double* psi,
                             const double* source,
                             const hipfoam_controls* controls,
                             hipfoam_performance* perf, double* residual);

/* Allocation-free variant of hipfoam_solve_device.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results as hipfoam_solve_device (launched eagerly, never through graphs),
 * but it calls no hipMalloc/hipFree: every per-solve buffer (work vectors,
 * staged FP64/FP32 operator, row-slot coefficients, reduction partials,
 * Krylov and outer states) lives in a fixed 64 MiB scratch array that each
 * GPU holds for the life of the process.  That holds about 324 bytes per
 * cell of a hexahedral mesh (3 faces per cell; 360 with `lower`), i.e. up to
 * roughly 200,000 cells; larger meshes return HIPFOAM_ERR_SIZE.  The mesh's
 * addressing must have been uploaded by hipfoam_set_topology beforehand (it
 * allocates, once per mesh).  The context's own staged operator is left as
 * it was.  All calls on one GPU share its array, which keeps the previous
 * call's contents (the call clears what it needs zeroed): overlapping calls
 * on one GPU would race, so the library runs them one at a time per device
 * within a process (a lock held for the duration of the call). */
int32_t hipfoam_solve_device_noalloc(hipfoam_context* ctx, int32_t solver,
                                     const double* diag, const double* upper,
                                     const double* lower, double* psi,
                                     const double* source,
                                     const hipfoam_controls* controls,
                                     hipfoam_performance* perf,
                                     double* residual);

/* Plain-FP64 GPU comparator (no reference counterpart; the "plain FP64"
 * side of the metric): OpenFOAM's PCG / PBiCGStab loop in FP64 on the exact
 * operator, with the same preconditioner (Chebyshev-Jacobi of
 * controls->chebyshevOrder for PCG, Jacobi for PBiCGStab), the same
 * normFactor and checkConvergence, no refinement. */
int32_t hipfoam_solve_fp64(hipfoam_context* ctx, int32_t solver,
                           const double* diag, const double* upper,
                           const double* lower, double* psi,
                           const double* source,
                           const hipfoam_controls* controls,
                           hipfoam_performance* perf, double* residual);
int32_t hipfoam_solve_fp64_device(hipfoam_context* ctx, int32_t solver,
                                  const double* diag, const double* upper,
                                  const double* lower, double* psi,
                                  const double* source,
                                  const hipfoam_controls* controls,
                                  hipfoam_performance* perf,
                                  double* residual);

/* ---- adjoint (no reference counterpart) ------------------------------- */

/* Gradient of a loss L(x) through the solve A x = b: given the forward
 * solution x and g = dL/dx, solves the transpose system A^T lambda = g with
 * the same refined solver and controls (upper and lower swap roles on the
 * same LDU addressing, lduMatrix::Tmul semantics; lambda0 = 0), then forms
 *   dL/dupper_f = -lambda[l] x[u],  dL/dlower_f = -lambda[u] x[l],
 *   dL/ddiag    = -lambda o x,      dL/db       = lambda.
 * For a symmetric matrix (lower == NULL) the single face coefficient's
 * gradient -(lambda[l] x[u] + lambda[u] x[l]) goes to dUpper and dLower is
 * not written.  Any gradient pointer may be NULL.  perf describes the
 * transpose solve.  precision selects the transpose solver: HIPFOAM_FP32 =
 * the refined FP32-inner solve of hipfoam_solve, HIPFOAM_FP64 = the plain
 * FP64 solve of hipfoam_solve_fp64 (for tolerances below what FP32 inner
 * iterations reach on the right-hand side g). */
int32_t hipfoam_adjoint(hipfoam_context* ctx, int32_t solver, int32_t precision,
                        const double* diag, const double* upper,
                        const double* lower, const double* x, const double* g,
                        double* lambda, double* dUpper, double* dLower,
                        double* dDiag, double* dSource,
                        const hipfoam_controls* controls,
                        hipfoam_performance* perf);
int32_t hipfoam_adjoint_device(hipfoam_context* ctx, int32_t solver,
                               int32_t precision, const double* diag,
                               const double* upper,
                               const double* lower, const double* x,
                               const double* g, double* lambda, double* dUpper,
                               double* dLower, double* dDiag, double* dSource,
                               const hipfoam_controls* controls,
                               hipfoam_performance* perf);

/* ---- batched solves (no reference counterpart) ------------------------ */

/* Many independent LDU systems, each with nRhs right-hand sides sharing the
 * system's matrix (segregated Ux/Uy/Uz/... or independent meshes), solved
 * together: every step of the refined solve is one launch over all systems
 * and right-hand sides, each (system, rhs) segment keeping its own scalars,
 * convergence test, deflation, restarts and polish exactly as hipfoam_solve
 * does for one system.  A batch holds either symmetric systems (lower NULL,
 * HIPFOAM_PCG or HIPFOAM_PBICGSTAB) or asymmetric ones (HIPFOAM_PBICGSTAB).
 * Batched hipPCG uses the Chebyshev-Jacobi preconditioner (order >= 2). */
typedef struct hipfoam_batch hipfoam_batch;

typedef struct
{
    int32_t nCells, nFaces;
    const int32_t* lowAddr;      /* [nFaces] */
    const int32_t* uppAddr;      /* [nFaces] */
    const int32_t* ownStart;     /* [nCells + 1] */
    const int32_t* losort;       /* [nFaces] */
    const int32_t* losortStart;  /* [nCells + 1] */
    const double* diag;          /* [nCells] */
    const double* upper;         /* [nFaces] */
    const double* lower;         /* [nFaces] or NULL (symmetric) */
} hipfoam_system;

/* Uploads addressing and coefficients of every system (host pointers, copied). */
int32_t hipfoam_batch_create(hipfoam_context* ctx, int32_t solver,
                             int32_t nSystems, const hipfoam_system* systems,
                             int32_t nRhs, hipfoam_batch** out);
/* source, psi: [nRhs][nCells] (rhs-major); psi NULL starts from zero. */
int32_t hipfoam_batch_set_rhs(hipfoam_batch* batch, int32_t system,
                              const double* source, const double* psi);
/* Sets every psi of the batch to zero on the device. */
int32_t hipfoam_batch_reset_psi(hipfoam_batch* batch);
/* Refined solve of every segment; perf (may be NULL) receives
 * nSystems * nRhs records, segment = system * nRhs + rhs. */
int32_t hipfoam_batch_solve(hipfoam_batch* batch,
                            const hipfoam_controls* controls,
                            hipfoam_performance* perf);
/* psi: [nRhs][nCells] of one system. */
int32_t hipfoam_batch_get_psi(hipfoam_batch* batch, int32_t system, double* psi);
/* Right-hand sides per workgroup of the batch's Krylov iteration kernels,
 * 1..8 (default 8): each row's matrix slots are loaded once per group.
 * Results do not depend on it.  No reference counterpart (the reference
 * has no batched solver). */
int32_t hipfoam_batch_set_rhs_group(hipfoam_batch* batch, int32_t group);
void hipfoam_batch_destroy(hipfoam_batch* batch);

/* ---- operator primitives --------------------------------------------- */

/* y = A x with the operator last staged: precision HIPFOAM_FP32 uses the
 * inner FP32 operator (set_coeffs / solve), HIPFOAM_FP64 the FP64 matrix of
 * the last solve or hipfoam_set_matrix.  The GPU form of the row-gather LDU
 * SpMV (metalFoamKernels.metal:62-150) and of lduMatrix::Amul (S.C:109).
 * x, y: float or double [nCells], host pointers. */
int32_t hipfoam_amul(hipfoam_context* ctx, int32_t precision, const void* x,
                     void* y);
int32_t hipfoam_amul_device(hipfoam_context* ctx, int32_t precision,
                            const void* x, void* y);

/* Stage an FP64 matrix (and its FP32 inner operator, undeflated) without
 * solving: the per-solve coefficient upload of S.C:211-220 on its own. */
int32_t hipfoam_set_matrix(hipfoam_context* ctx, const double* diag,
                           const double* upper, const double* lower);
int32_t hipfoam_set_matrix_device(hipfoam_context* ctx, const double* diag,
                                  const double* upper, const double* lower);

/* ---- device memory helpers for callers that keep fields in HBM ------- */
int32_t hipfoam_device_alloc(hipfoam_context* ctx, uint64_t bytes,
                             void** ptr);
int32_t hipfoam_device_free(hipfoam_context* ctx, void* ptr);
int32_t hipfoam_copy_to_device(hipfoam_context* ctx, void* dst,
                               const void* src, uint64_t bytes);
int32_t hipfoam_copy_to_host(hipfoam_context* ctx, void* dst,
                             const void* src, uint64_t bytes);
int32_t hipfoam_copy_on_device(hipfoam_context* ctx, void* dst,
                               const void* src, uint64_t bytes);
int32_t hipfoam_synchronize(hipfoam_context* ctx);

/* ---- schedule / measurement ------------------------------------------ */

/* SpMV schedules of the inner operator:
 *   0 = losort-ordered lower coefficients (one indirection),
 *   1 = direct lower[losort[k]] / lowAddr[losort[k]] (two indirections),
 *   2 = row slots (the default): every row's <= 3 owner and <= 3 neighbour
 *       coefficients in slot-major arrays (a symmetric matrix stores each
 *       face once, on its owner side), columns from per-256-row-tile bases
 *       (a 1-byte presence mask per row where every face sits at its base,
 *       int16 deltas elsewhere), no row-start indirection.  A tile with a
 *       row of more than 3 owner or 3 neighbour faces, or column offsets
 *       beyond the deltas' range, is a face tile: its rows run the face
 *       loops of schedule 0 (hipfoam_slot_tiles counts them); when every
 *       tile would be one, schedule 0 runs.  Same terms in the same order as
 *       0 and 1, so all three give identical results,
 * and the iteration driver:
 *   graphChunk = iterations captured per hipGraph launch (even), 0 = eager
 *                launches, -1 = auto (the default): eager from 524,288 rows
 *                on, where the kernels outlast their launches, graphs of 4
 *                iterations below.  Results do not depend on it.
 * blockSize 0 (auto, the default) picks the launch shape from the mesh.
 * Kernels that reduce: one 256-thread workgroup per 256-row tile up to 512
 * tiles, 512 workgroups of 512 threads up to 49,152 tiles (every
 * reduction's second stage then runs inside the next row kernel), and 16
 * workgroups of 512 threads per CU beyond.  Kernels that reduce nothing (the
 * interior Chebyshev steps, the plain SpMV): 16 workgroups of 256 threads
 * per CU up to 32,768 tiles, 8 of 512 beyond.  An explicit blockSize
 * launches min(ceil(n / blockSize),
 * gridLimit) workgroups; each walks a contiguous run of blockSize-row tiles,
 * and second stages run inside the next kernel when there are at most 512
 * workgroups of at least 256 threads.  Results are deterministic for a given
 * mesh and schedule; grids of different sizes group the reductions
 * differently, so their FP64 sums may differ in the last bits. */
typedef struct
{
    int32_t spmvSchedule;
    int32_t blockSize;    /* 0 = auto (default), or 64 .. 1024, multiple of 64 */
    int32_t xcdRemap;     /* 0 = hardware order (workgroup b on XCD b mod 8),
                             1 = contiguous row range per XCD,
                             2 = auto (default): 0 when the mesh's largest
                             stride is a multiple of 8 workgroups and its
                             second stride at most 1/8 of a workgroup's
                             rows, else 1; batches use 1.  Results do not
                             depend on it */
    int32_t graphChunk;
    int32_t gridLimit;    /* max workgroups per row kernel; 0 = 8 per CU
                             (auto blockSize: its own grid) */
    int32_t rowForm;      /* 0 = unrolled predicated rows when every row has
                             <= 3 owner and <= 3 neighbour faces, else loops;
                             1 = always loops */
    int32_t persistent;   /* Chebyshev PCG inner loop as one persistent
                             launch on one XCD, on the per-tile launch shape
                             (one 256-thread workgroup per 256-row tile):
                             -1 = auto (default): up to 160 tiles (40,960
                             rows) for the FP32 inner loop, 96 for the
                             plain-FP64 loop; 1 = up to 512 tiles; 0 =
                             never.  The
                             same rows, partials and fold order as the
                             launched loop: identical results */
    int32_t localLoop;    /* inner loop in one workgroup of 512 threads
                             with every vector in LDS, on a mesh whose
                             tiles are all row-slot tiles, on the per-tile
                             launch shape: Chebyshev PCG up to 2,816 rows
                             (FP32 inner loop) or 1,536 (plain FP64),
                             BiCGStab up to 1,536 (FP64 1,280) rows:
                             -1 = auto (default), 0 = never.  Taken before
                             persistent where both apply; the same rows,
                             partials and fold order as the launched loop:
                             identical results */
} hipfoam_schedule;

int32_t hipfoam_get_schedule(const hipfoam_context* ctx, hipfoam_schedule* s);
/* The SpMV schedule the inner operator of the current mesh runs: the
 * requested one, or 0 when row slots (2) were requested and the mesh does
 * not allow them.  No reference counterpart (the reference has one kernel). */
int32_t hipfoam_active_schedule(const hipfoam_context* ctx,
                                int32_t* spmvSchedule);
/* Row-slot layout of the current mesh (schedule 2): *patternedRows = rows
 * in tiles whose faces all sit at the tile's per-slot column offsets, which
 * need 1 byte of addressing per row instead of 13; *symmetricPlanes = the
 * number of coefficient planes a symmetric matrix stores (3: each face's
 * coefficient once, in its owner row; an asymmetric matrix stores 6).  Both
 * 0 when the mesh does not allow row slots.  No reference counterpart. */
int32_t hipfoam_slot_layout(const hipfoam_context* ctx, int64_t* patternedRows,
                            int32_t* symmetricPlanes);
/* How the current mesh's rows split over the tile kinds of the row-slot
 * layout: patterned tiles (1 byte of addressing per row), explicit tiles
 * (int16 deltas) and face tiles, whose rows run the LDU face loops because
 * they do not fit the slots (a row with more than 3 owner or 3 neighbour
 * faces, or column offsets scattered beyond the deltas' range).  Without a
 * row-slot layout every row is counted in *faceRows.  No reference
 * counterpart. */
int32_t hipfoam_slot_tiles(const hipfoam_context* ctx, int64_t* patternedRows,
                           int64_t* explicitRows, int64_t* faceRows);
int32_t hipfoam_set_schedule(hipfoam_context* ctx, const hipfoam_schedule* s);
/* Inner loops this context ran as one persistent launch (schedule field
 * persistent), how many of those were abandoned part-way and run again on
 * the launched loop, and why the last one was: 0x10000 | workgroup << 4 |
 * XCD id when that workgroup found a participant on another XCD (the id),
 * 0x20000 | barrier index for a barrier wait that timed out, 0x40000 for
 * one reported by hipfoam_simulate_abandon (0: none abandoned).  Any
 * pointer may be NULL.  No reference counterpart. */
int32_t hipfoam_persistent_stats(const hipfoam_context* ctx, int64_t* launches,
                                 int64_t* abandoned, int64_t* lastAbort);
/* Inner loops this context ran in one workgroup (schedule field localLoop).
 * No reference counterpart. */
int32_t hipfoam_local_stats(const hipfoam_context* ctx, int64_t* launches);
/* For tests of the recovery from an abandoned persistent launch: the next
 * count persistent launches of this context run to their end and are then
 * reported as abandoned -- their Krylov state is set back to running and
 * their abort word to 0x40000, as a launch that gave up leaves them -- so the
 * driver runs the inner solve again (and whatever it had queued behind the
 * launch does not run).  count 0 cancels.  No reference counterpart. */
int32_t hipfoam_simulate_abandon(hipfoam_context* ctx, int32_t count);
/* Times the candidate SpMV schedules on the staged operator and keeps the
 * fastest; *best (may be NULL) receives it. */
int32_t hipfoam_autotune(hipfoam_context* ctx, hipfoam_schedule* best);
/* Times the batched SpMV for every SpMV schedule, row form, XCD mapping and
 * RHS group, and keeps the fastest: the schedule in the context (as
 * hipfoam_autotune does, so it also applies to the context's
 * single-system solves), the group in the batch.  The default (row slots,
 * XCD mapping 1, one group of all RHS up to 8) is kept unless another
 * candidate is faster by more than 2 %.  best and group may be NULL.
 * Results do not depend on the choice.  No reference counterpart. */
int32_t hipfoam_batch_autotune(hipfoam_batch* batch, hipfoam_schedule* best,
                               int32_t* group);

/* Kernel timing: when enabled, the inner loops launch without graphs and
 * bracket every launch of kernel class `kclass` with HIP events on the
 * context's stream.  hipfoam_kernel_timing returns the launch count and the
 * summed seconds since the last reset.  kclass: 0 = interior Chebyshev
 * SpMV step, 1 = p SpMV (A p), 2 = BiCGStab SpMV, 3 = first Chebyshev step,
 * 4 = last Chebyshev step, 5 = plain SpMV of hipfoam_amul*. */
int32_t hipfoam_set_kernel_timing(hipfoam_context* ctx, int32_t enable,
                                  int32_t kclass);
int32_t hipfoam_kernel_timing(hipfoam_context* ctx, int64_t* launches,
                              double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* HIPFOAM_H */
