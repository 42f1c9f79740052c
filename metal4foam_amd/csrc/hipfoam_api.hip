// hipfoam_api.hip -- host side of the MI355X LDU Krylov library: context,
// device-resident inner loops (hipGraph chunks), the FP64-on-GPU refined
// solve, and the extern "C" exports declared in include/hipfoam.h.
//
// Reference interfaces replaced (scotCW/metal4Foam):
//   metalFoamBackend   src/metal4Foam/backend/metalFoamBackend.{H,C}
//   metalFoamSolver    src/metal4Foam/solvers/metalFoamSolver.C:77-437
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <fenv.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hipfoam.h"
#include "hipfoam_batched.hpp"
#include "hipfoam_device.hpp"
#include "hipfoam_local.hpp"
#include "hipfoam_persist.hpp"
#include "hipfoam_slots.hpp"
#include "hipfoam_stage.hpp"

using namespace hipfoam;

namespace {

double now_seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// OpenFOAM traps SIGFPE; run library code with the default FP environment
// and restore the caller's on exit (metalFoamBackend.C:48-57).
struct FpTrapGuard {
  fenv_t saved;
  FpTrapGuard() {
    fegetenv(&saved);
    fesetenv(FE_DFL_ENV);
  }
  ~FpTrapGuard() { fesetenv(&saved); }
};

// Set while hipfoam_solve_device_noalloc runs on this thread: a buffer that
// would have to be allocated fails instead (every buffer that call uses is
// carved from the scratch array beforehand).
thread_local bool t_noAlloc = false;

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t cap = 0;
  bool borrowed = false;  // points into the scratch array: never freed
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  hipError_t ensure(size_t n) {
    if (n == 0) n = 1;
    if (p && n <= cap) return hipSuccess;
    if (t_noAlloc || borrowed) return hipErrorOutOfMemory;
    release();
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      cap = 0;
      return e;
    }
    cap = n;
    return hipSuccess;
  }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    borrowed = false;
  }
};

// Krylov workspace of one precision.
template <typename T>
struct Work {
  // r1: the residual's second buffer (Chebyshev PCG: r ping-pongs, k_cheb_first)
  Buf<T> r, r1, p0, p1, wA, Ap, delta, deltaBest, z0, z1, w, rA0, yA, AyA, AyA1, sA, zA, tA;
  size_t n = 0;
  hipError_t ensure(size_t nn) {
    if (n == nn && r.p) return hipSuccess;
    Buf<T>* all[] = {&r, &r1, &p0, &p1, &wA, &Ap, &delta, &deltaBest, &z0,
                     &z1, &w, &rA0, &yA, &AyA, &AyA1, &sA, &zA, &tA};
    for (auto* b : all) {
      b->release();
      hipError_t e = b->ensure(nn);
      if (e != hipSuccess) return e;
    }
    n = nn;
    return hipSuccess;
  }
  void release() {
    Buf<T>* all[] = {&r, &r1, &p0, &p1, &wA, &Ap, &delta, &deltaBest, &z0,
                     &z1, &w, &rA0, &yA, &AyA, &AyA1, &sA, &zA, &tA};
    for (auto* b : all) b->release();
    n = 0;
  }
};

enum Kind { KIND_PCG = 0, KIND_BICG = 1 };

struct GraphKey {
  int kind, prec, order, chunk, direct, remap, block, grid, form, gen;
  bool operator<(const GraphKey& o) const {
    return std::tie(kind, prec, order, chunk, direct, remap, block, grid, form, gen) <
           std::tie(o.kind, o.prec, o.order, o.chunk, o.direct, o.remap, o.block, o.grid,
                    o.form, o.gen);
  }
};

// Per-launch HIP-event timing of one kernel class on the context stream.
struct KernelTimer {
  bool on = false;
  int kclass = 0;
  std::vector<hipEvent_t> ev;
  std::vector<int> count;  // launches inside each event pair
  size_t used = 0;
  int64_t launches = 0;
  double seconds = 0;
};

}  // namespace

struct hipfoam_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool good = false;
  std::string desc = "hipfoam: no device";
  std::string err;
  int numCUs = 0;

  // topology
  int n = 0, nf = 0;
  int maxOwn = 0, maxNbr = 0;  // longest owner / neighbour face run of a row
  bool slotOK = false;  // row-slot layout possible (see build_slots)
  int64_t slotPatterned = 0;  // rows in patterned tiles
  int64_t slotFaceRows = 0;   // rows in face tiles (LDU face loops)
  int64_t strideMajor = 0, strideMinor = 0;  // dominant offsets (SlotLayout)
  Buf<uint32_t> slotOff;
  Buf<int32_t> slotTile;
  Buf<uint8_t> slotRow;
  Buf<float> slotCoef32;
  Buf<double> slotCoef64;
  const int32_t* lastLow = nullptr;
  const int32_t* lastUpp = nullptr;
  uint64_t topoHash = 0;  // addressing_hash of the uploaded mesh
  bool topoHashValid = false;
  int gen = 0;  // bumps whenever device buffers are reallocated
  Buf<int32_t> ownStart, losortStart, uppAddr, lowAddr, losort, losortLow;

  // FP64 matrix (staged per solve) and its FP32 inner operator
  Buf<double> diag64, upper64, lower64, lowerL64, rD64;
  bool sym64 = true;
  bool have64 = false;
  Buf<float> diag32, upper32, lowerL32, lower32, rD32;
  bool sym32 = true;
  bool have32 = false;

  // FP64 outer vectors
  Buf<double> psi0, psi1, b, r64, u64, scratch64;
  // adjoint staging (host-pointer entry point)
  Buf<double> adjX, adjG, adjLam, adjDU, adjDL, adjDD, adjDB;
  Work<float> w32;
  Work<double> w64;

  Buf<double> partials;
  Buf<double> partials2;  // partials of the kernels that merge a second stage
  int pstride = 0;
  Buf<KState> kst;
  Buf<KState> kst2;  // the Krylov state between the two merging kernels
  Buf<OState> ost;
  KState* hk = nullptr;  // pinned
  OState* ho = nullptr;  // pinned
  int32_t* hstatus = nullptr;  // pinned: 2 poll slots, the persistent loop's abort word
  // pinned, coherent host word the stages writing kst post its status to, and
  // its device address (nullptr: the host copies the status instead)
  int32_t* mirrorHost = nullptr;
  int32_t* mirrorDev = nullptr;
  // barrier words of the persistent inner loop (XcdSync: flags, XCD ids, abort, epoch)
  Buf<uint32_t> xsync;
  int64_t persistLaunches = 0;   // persistent inner loops launched
  int64_t persistAbandoned = 0;  // ... of which abandoned (the launched loop ran instead)
  int64_t persistLastAbort = 0;  // abort word of the last abandoned launch (XcdSync)
  int64_t localLaunches = 0;     // inner loops run in one workgroup (k_pcg_local, k_bicg_local)
  int simAbandon = 0;            // persistent launches still to report as abandoned
  hipEvent_t evPoll[2] = {nullptr, nullptr};
  hipEvent_t evT0 = nullptr, evT1 = nullptr;

  hipfoam_schedule sched{2, 0, 2, -1, 0, 0, -1, -1};
  std::map<GraphKey, hipGraphExec_t> graphs;
  KernelTimer timer;
  // host helper threads (created on first use): staged copies, addressing hash
  std::unique_ptr<CopyPool> pool;
  // pinned staging of large host arrays (created on first use)
  std::unique_ptr<HostStager> stager;
  // topology check of hipfoam_set_topology: HIPFOAM_TOPO_CONTENT (default) or
  // HIPFOAM_TOPO_POINTERS
  int topoCheck = HIPFOAM_TOPO_CONTENT;
  // device address of this device's scratch array (hipGetSymbolAddress at
  // create), used by hipfoam_solve_device_noalloc only
  char* scratchDev = nullptr;
};

namespace {


// Launch shape of the row kernels.  blockSize 0 (auto, the default), by the
// mesh's number of 256-row tiles:
//  - at most kAutoSmallTiles: one 256-thread workgroup per tile;
//  - up to kAutoMidTiles: kAutoMidGrid workgroups of kAutoMidBlock threads,
//    each walking a contiguous run of tiles;
//  - beyond: kAutoBigPerCU workgroups of kAutoMidBlock threads per CU, with
//    separate finalise launches.
// The first two stay within kMergeMaxBlocks, so every reduction's second
// stage runs inside the next row kernel.  Against 8 x 256 per CU (same box,
// lognormal Poisson PCG unless noted): 128^3 16.2 -> 15.7 ms, 160^3 44.0 ->
// 36.5 ms, 192^3 88.6 -> 83.3 ms, 224^3 170.3 -> 167.0 ms, 128^3 BiCGStab
// 4.16 -> 3.86 ms, 96^3 BiCGStab 2.91 -> 2.80 ms; 96^3 7.26 vs 7.28 ms.  At
// 256^3 (65,536 tiles) the merged grid is 3-6 % slower than full occupancy,
// where every pass streams from HBM; there 16 x 512 per CU beats 8 x 256
// per CU: 240^3 229.9 -> 225.4 ms, 256^3 273.1 -> 265.0 ms, 288^3 454.2 ->
// 447.0 ms, 320^3 820.7 -> 791.8 ms (the same shape measured within 0.5 %
// at 160^3 and 224^3 and 4 % slower at 192^3, so the mid rule stays).  An
// explicit blockSize launches min(ceil(n / blockSize), gridLimit)
// workgroups (gridLimit 0: 8 per CU).
// Rows from which the automatic graph chunk (-1) launches eagerly.
constexpr int kEagerRows = 1 << 19;
constexpr int kMergeMaxBlocks = 512;
constexpr int kAutoSmallTiles = 512;
constexpr int kAutoMidTiles = 49152;
constexpr int kAutoMidBlock = 512;
constexpr int kAutoMidGrid = 512;
constexpr int kAutoBigPerCU = 16;

inline int64_t tiles256(const hipfoam_context* c) { return ((int64_t)c->n + 255) / 256; }

// Grid of at most cap workgroups over tiles row tiles with every workgroup
// taking the same number of grid-stride passes (rounded up to 8 workgroups
// for the XCD mapping): the pass count of a grid of cap, without the
// workgroups that run one pass fewer and leave the tail of the launch to
// the others.  Refined lognormal Poisson PCG against the unbalanced grid of
// cap workgroups (same box, alternating runs, same solutions): 72^3 3.46 ->
// 3.28 ms, 112^3 11.82 -> 10.45 ms, 144^3 27.0 -> 26.5 ms, 160^3 35.9 ->
// 35.3 ms; 96^3 and 128^3 (already balanced) and BiCGStab 80^3-128^3
// within 1 %.
inline int balanced_grid(int tiles, int cap) {
  if (tiles <= cap) return tiles;
  const int passes = (tiles + cap - 1) / cap;
  const int g = (tiles + passes - 1) / passes;
  return std::min(cap, (g + 7) / 8 * 8);
}

inline int row_block(const hipfoam_context* c) {
  if (c->sched.blockSize > 0) return c->sched.blockSize;
  return tiles256(c) <= kAutoSmallTiles ? 256 : kAutoMidBlock;
}

inline int row_grid(const hipfoam_context* c) {
  const int bs = row_block(c);
  const int tiles = std::max(1, (c->n + bs - 1) / bs);
  int cap = c->sched.gridLimit > 0 ? c->sched.gridLimit : 8 * std::max(c->numCUs, 1);
  if (c->sched.blockSize == 0 && c->sched.gridLimit <= 0 && tiles256(c) > kAutoSmallTiles) {
    cap = tiles256(c) <= kAutoMidTiles ? kAutoMidGrid : kAutoBigPerCU * std::max(c->numCUs, 1);
    return balanced_grid(tiles, cap);
  }
  return std::min(tiles, cap);
}

// Launch shape of the row kernels that reduce nothing (the interior
// Chebyshev steps, the plain SpMV): no partials to fold, so the automatic
// schedule is free to size them for occupancy.
// Their results do not depend on the launch shape.
// Automatic shape: 256-thread workgroups, 16 per CU (two rounds of resident
// workgroups, 2 passes of rows each at 128^3), and from kStreamBigTiles
// 256-row tiles on 512-thread workgroups, 8 per CU.  Measured on the
// interior Chebyshev step against 256 x 8 per CU (same box): 128^3 15.7 ->
// 15.05 us (solve 15.75 -> 15.10 ms), 96^3 8.9 -> 8.2 us (7.26 -> 6.93 ms),
// 256^3 160 -> 142 us (294 -> 274 ms).
constexpr int64_t kStreamBigTiles = 32768;

inline int stream_block(const hipfoam_context* c) {
  if (c->sched.blockSize > 0) return c->sched.blockSize;
  return tiles256(c) > kStreamBigTiles ? 512 : 256;
}

inline int stream_grid(const hipfoam_context* c) {
  const int bs = stream_block(c);
  const int tiles = std::max(1, (c->n + bs - 1) / bs);
  const int perCU = (c->sched.blockSize == 0 && tiles256(c) <= kStreamBigTiles) ? 16 : 8;
  if (c->sched.gridLimit > 0) return std::min(tiles, c->sched.gridLimit);
  const int cap = perCU * std::max(c->numCUs, 1);
  return c->sched.blockSize == 0 ? balanced_grid(tiles, cap) : std::min(tiles, cap);
}

// Block-to-XCD mapping of the row kernels.  Hardware dispatch places
// workgroup b on XCD b mod 8; xcdRemap 1 gives each XCD one contiguous row
// range instead.  Auto (2): keep the hardware order when it already keeps
// gathers on one XCD's L2, i.e. when the mesh's largest stride is a
// multiple of 8 workgroups and its second stride is at most an eighth of a
// workgroup's rows (128^3: strides 16384 and 128 at 1024 rows per
// workgroup); else contiguous ranges.  The mapping moves no arithmetic:
// partials are indexed by logical block.
inline int effective_remap(const hipfoam_context* c, bool stream = false) {
  if (c->sched.xcdRemap != 2) return c->sched.xcdRemap;
  const int64_t bs = stream ? stream_block(c) : row_block(c);
  const int64_t tiles = std::max<int64_t>(1, (c->n + bs - 1) / bs);
  const int64_t g = stream ? stream_grid(c) : row_grid(c);
  const int64_t rows = (tiles + g - 1) / g * bs;  // rows per workgroup
  const bool local = c->strideMajor > 0 && c->strideMajor % (8 * rows) == 0 &&
                     c->strideMinor * 8 <= rows;
  return local ? 0 : 1;
}

template <int NV, bool LASTMAX, typename F>
void launch_fin(hipfoam_context* c, int nblk, bool checkStatus, int runMask, F f) {
  hipLaunchKernelGGL((k_finalize<NV, LASTMAX, F>), dim3(1), dim3(256), 0, c->stream,
                     c->partials.p, c->pstride, nblk,
                     checkStatus ? &c->kst.p->status : (const int32_t*)nullptr, runMask,
                     checkStatus ? c->mirrorDev : (int32_t*)nullptr, f);
}

// Second stage of a reduction: one workgroup folds the partials of the
// kernel just enqueued (nblk workgroups) and applies the finaliser f.  With
// checkStatus, it runs only when the Krylov status bit is set in runMask.
constexpr int kRunning = 1 << ST_RUNNING;

// The same for a refinement stage: skipped while *guard (NULL: always runs)
// is RUNNING, posting no status.
template <int NV, typename F>
void launch_fin_guarded(hipfoam_context* c, int nblk, const int32_t* guard, F f) {
  hipLaunchKernelGGL((k_finalize<NV, false, F>), dim3(1), dim3(256), 0, c->stream,
                     c->partials.p, c->pstride, nblk, guard, ~kRunning, (int32_t*)nullptr, f);
}
constexpr int kRunningOrEarly = (1 << ST_RUNNING) | (1 << ST_EARLY);

template <int NV, bool LASTMAX = false, typename F>
void launch_fin(hipfoam_context* c, int nblk, bool checkStatus, int runMask, F f);

#define CK(expr)                                                              \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      c->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
      return HIPFOAM_ERR_DEVICE;                                              \
    }                                                                         \
  } while (0)

#define CK_C(ctx, expr)                                                       \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return HIPFOAM_ERR_DEVICE;                                              \
    }                                                                         \
  } while (0)

#define CKL()                                                                 \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      c->err = std::string("kernel launch: ") + hipGetErrorString(e_);        \
      return HIPFOAM_ERR_DEVICE;                                              \
    }                                                                         \
  } while (0)

// Copy of `bytes` on the context stream.  Host <-> device copies of at least
// HostStager::kMin bytes go through the pinned staging ring: host to device
// returns once the host array has been read, device to host once the host
// array holds the data (after the work already on the stream).  Other copies
// are hipMemcpyAsync.
// The context's host helper threads: up to 7 besides the caller's.
CopyPool& host_pool(hipfoam_context* c) {
  if (!c->pool) {
    const int hw = (int)std::thread::hardware_concurrency();
    c->pool = std::make_unique<CopyPool>(std::max(0, std::min(7, hw - 1)));
  }
  return *c->pool;
}

int xfer(hipfoam_context* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  const bool staged = (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToHost) &&
                      bytes >= HostStager::kMin;
  if (!staged) {
    CK(hipMemcpyAsync(dst, src, bytes, kind, c->stream));
    return HIPFOAM_OK;
  }
  if (!c->stager) {
    auto st = std::make_unique<HostStager>();
    CK(st->init(&host_pool(c)));
    c->stager = std::move(st);
  }
  if (kind == hipMemcpyHostToDevice)
    CK(c->stager->h2d(dst, src, bytes, c->stream));
  else
    CK(c->stager->d2h(dst, src, bytes, c->stream));
  return HIPFOAM_OK;
}

#define XF(dst, src, bytes, kind)                    \
  do {                                               \
    const int rc_ = xfer(c, (dst), (src), (bytes), (kind)); \
    if (rc_) return rc_;                             \
  } while (0)

int fail(hipfoam_context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

inline int blocks_for(int n, int bs) { return (n + bs - 1) / bs; }

void clear_graphs(hipfoam_context* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

// slotCoef: row-slot coefficients of this operator (nullptr: none staged); used
// under spmvSchedule 2 when the mesh allows the row-slot layout.
template <typename T>
LduOp<T> make_op(hipfoam_context* c, const T* diag, const T* upper, const T* lowerL,
                 const T* lowerF, const T* slotCoef) {
  LduOp<T> op;
  op.ownStart = c->ownStart.p;
  op.losortStart = c->losortStart.p;
  op.uppAddr = c->uppAddr.p;
  op.losort = c->losort.p;
  op.diag = diag;
  op.upper = upper;
  op.n = c->n;
  op.xcdRemap = effective_remap(c);
  op.shortRows = c->maxOwn <= kShortRow && c->maxNbr <= kShortRow && c->sched.rowForm == 0;
  op.direct = c->sched.spmvSchedule == 1;
  op.slots = slotCoef && c->sched.spmvSchedule == 2 && c->slotOK;
  op.slotCoef = op.slots ? slotCoef : nullptr;
  op.slotOff = op.slots ? c->slotOff.p : nullptr;
  op.slotTile = op.slots ? c->slotTile.p : nullptr;
  op.slotRow = op.slots ? c->slotRow.p : nullptr;
  op.slotSym = op.slots && lowerF == upper;
  if (op.direct) {
    op.nbrAddr = c->lowAddr.p;
    op.lowerL = lowerF;
  } else {
    op.nbrAddr = c->losortLow.p;
    op.lowerL = lowerL;
  }
  return op;
}

// The operator with the block-to-XCD mapping of the stream launch shape.
template <typename T>
LduOp<T> stream_op(LduOp<T> A, const hipfoam_context* c) {
  A.xcdRemap = effective_remap(c, true);
  return A;
}

LduOp<float> op32(hipfoam_context* c) {
  return make_op<float>(c, c->diag32.p, c->upper32.p, c->lowerL32.p,
                        c->sym32 ? c->upper32.p : c->lower32.p, c->slotCoef32.p);
}

LduOp<double> op64(hipfoam_context* c) {
  return make_op<double>(c, c->diag64.p, c->upper64.p, c->lowerL64.p,
                         c->sym64 ? c->upper64.p : c->lower64.p, c->slotCoef64.p);
}


KStateWords state_words(const KState& s) {
  KStateWords w;
  std::memcpy(&w, &s, sizeof(KState));
  return w;
}

template <typename T>
Work<T>& work(hipfoam_context* c);
template <>
Work<float>& work<float>(hipfoam_context* c) { return c->w32; }
template <>
Work<double>& work<double>(hipfoam_context* c) { return c->w64; }

// -------------------------------------------------------------------------
// kernel timing hooks
// -------------------------------------------------------------------------

int timer_flush(hipfoam_context* c) {
  KernelTimer& t = c->timer;
  if (t.used == 0) return HIPFOAM_OK;
  CK(hipEventSynchronize(t.ev[t.used - 1]));
  for (size_t k = 0; k + 1 < t.used; k += 2) {
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]));
    t.seconds += 1e-3 * ms;
    t.launches += t.count[k / 2];
  }
  t.used = 0;
  return HIPFOAM_OK;
}

// Kernel-class timing: a timed launch goes through hipExtLaunchKernelGGL with
// a start/stop event pair that the dispatch packet itself records, so the
// elapsed time is the kernel's execution, not the gap between launches.
bool timer_wants(hipfoam_context* c, int kclass, bool capturing) {
  const KernelTimer& t = c->timer;
  return t.on && !capturing && t.kclass == kclass;
}

std::pair<hipEvent_t, hipEvent_t> timer_slot(hipfoam_context* c) {
  KernelTimer& t = c->timer;
  if (t.used + 2 > t.ev.size()) timer_flush(c);
  auto e = std::make_pair(t.ev[t.used], t.ev[t.used + 1]);
  t.count[t.used / 2] = 1;
  t.used += 2;
  return e;
}

#define LAUNCH_CLASSED(kclass, capturing, kernel, grid, block, ...)                             \
  do {                                                                                          \
    if (timer_wants(c, (kclass), (capturing))) {                                                \
      const auto ev_ = timer_slot(c);                                                           \
      hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, ev_.first, ev_.second, 0,        \
                            __VA_ARGS__);                                                       \
    } else {                                                                                    \
      hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, __VA_ARGS__);                       \
    }                                                                                           \
  } while (0)

// -------------------------------------------------------------------------
// one Krylov iteration, enqueued on the context stream
// -------------------------------------------------------------------------

// Second stages merged into the next kernel (Merge) need workgroups of at
// least 256 threads: the fold always has k_finalize's 256-thread tree, so
// merged and separate stages give the same bits.  Every workgroup of the
// merging kernel reads all
// partials, so merging pays only on small grids: 10k cells (40 workgroups)
// 2.57 -> 2.44 ms; 128^3 (2048) unchanged; 96^3 BiCGStab (1728, three
// values per partial) 3.19 -> 3.49 ms.
inline bool merge_stages(const hipfoam_context* c) {
  return row_block(c) >= kFoldThreads && row_grid(c) <= kMergeMaxBlocks;
}

// One PCG iteration.  With Chebyshev (order > 1):
//   k_cheb_first   previous iteration's update (delta, r), first step
//   k_cheb_step    steps 2 .. order-1 (the last also reduces rho)
//   FinResBeta     previous iteration's checks, beta
//   k_pcg_p_spmv   p = wA + beta p, A p, p.Ap
//   FinAlpha       alpha (the update is left pending)
// so an iteration is order + 2 launches; r and p ping-pong with the
// iteration parity.  On small grids FinResBeta runs inside k_pcg_p_spmv and
// FinAlpha inside the next k_cheb_first (state kst -> kst2 -> kst, partials
// P -> partials2 -> P).  With Jacobi (order 1) the update is a kernel of its
// own (k_pcg_update) followed by FinPcgRes.
template <typename T>
void enqueue_pcg_iter(hipfoam_context* c, const LduOp<T>& A, const T* rD, int order,
                      int parity, bool capturing) {
  Work<T>& w = work<T>(c);
  const int bs = row_block(c);
  const dim3 grid(row_grid(c)), block(bs);
  const int nb = grid.x;
  KState* st = c->kst.p;
  double* P = c->partials.p;
  const int ps = c->pstride;
  const bool merge = merge_stages(c);
  const Merge none{nullptr, nullptr, 0};
  T* pold = parity ? w.p1.p : w.p0.p;
  T* pnew = parity ? w.p0.p : w.p1.p;
  if (order > 1) {
    const T* rin = parity ? w.r1.p : w.r.p;
    T* rout = parity ? w.r.p : w.r1.p;
    KState* st2 = c->kst2.p;
    double* P2 = c->partials2.p;
    T* zout = order == 2 ? w.wA.p : w.z0.p;
    if (merge) {
      if (order == 2)
        LAUNCH_CLASSED(3, capturing, (k_cheb_first<T, true, FinAlpha<T>, true>), grid, block, A,
                       rin, rout, rD, pold, w.Ap.p, w.delta.p, w.deltaBest.p, w.w.p, zout, st2,
                       P, ps, Merge{st, P2, nb, c->mirrorDev}, FinAlpha<T>{nullptr});
      else
        LAUNCH_CLASSED(3, capturing, (k_cheb_first<T, false, FinAlpha<T>, true>), grid, block,
                       A, rin, rout, rD, pold, w.Ap.p, w.delta.p, w.deltaBest.p, w.w.p, zout,
                       st2, P, ps, Merge{st, P2, nb, c->mirrorDev}, FinAlpha<T>{nullptr});
    } else {
      if (order == 2)
        LAUNCH_CLASSED(3, capturing, (k_cheb_first<T, true, FinAlpha<T>, false>), grid, block,
                       A, rin, rout, rD, pold, w.Ap.p, w.delta.p, w.deltaBest.p, w.w.p, zout,
                       st, P, ps, none, FinAlpha<T>{nullptr});
      else
        LAUNCH_CLASSED(3, capturing, (k_cheb_first<T, false, FinAlpha<T>, false>), grid, block,
                       A, rin, rout, rD, pold, w.Ap.p, w.delta.p, w.deltaBest.p, w.w.p, zout,
                       st, P, ps, none, FinAlpha<T>{nullptr});
    }
    const T* zin = zout;
    // the interior steps reduce nothing: the stream launch shape
    const LduOp<T> As = stream_op(A, c);
    const dim3 sgrid(stream_grid(c)), sblock(stream_block(c));
    for (int s = 2; s < order; ++s) {
      const bool last = (s == order - 1);
      T* zo = last ? w.wA.p : ((s & 1) ? w.z0.p : w.z1.p);
      if (last)
        LAUNCH_CLASSED(4, capturing, (k_cheb_step<T, true>), grid, block, A, rout, rD,
                       zin, w.w.p, zo, st, P, ps, s);
      else
        LAUNCH_CLASSED(0, capturing, (k_cheb_step<T, false>), sgrid, sblock, As, rout,
                       rD, zin, w.w.p, zo, st, P, ps, s);
      zin = zo;
    }
    if (merge) {
      LAUNCH_CLASSED(1, capturing, (k_pcg_p_spmv<T, FinResBeta<T>, true, 2>), grid, block, A,
                     w.wA.p, pold, pnew, w.Ap.p, st, P2, ps, Merge{st2, P, nb},
                     FinResBeta<T>{nullptr});
    } else {
      launch_fin<2>(c, nb, true, kRunning, FinResBeta<T>{st});
      LAUNCH_CLASSED(1, capturing, (k_pcg_p_spmv<T, FinResBeta<T>, false, 2>), grid, block, A,
                     w.wA.p, pold, pnew, w.Ap.p, st, P, ps, none, FinResBeta<T>{nullptr});
      launch_fin<1>(c, nb, true, kRunning, FinAlpha<T>{st});
    }
    return;
  }
  // merged: beta's stage runs in k_pcg_p_spmv (state kst -> kst2, partials
  // P -> partials2) and alpha's in k_pcg_update (kst2 -> kst, partials2 -> P)
  hipLaunchKernelGGL((k_jacobi_dot<T>), grid, block, 0, c->stream, c->n, A.xcdRemap, w.r.p, rD,
                     w.wA.p, st, P, ps);
  if (!merge) launch_fin<1>(c, nb, true, kRunning, FinBeta<T>{st});
  if (merge) {
    KState* st2 = c->kst2.p;
    double* P2 = c->partials2.p;
    LAUNCH_CLASSED(1, capturing, (k_pcg_p_spmv<T, FinBeta<T>, true>), grid, block, A, w.wA.p,
                   pold, pnew, w.Ap.p, st, P2, ps, Merge{st2, P, nb}, FinBeta<T>{nullptr});
    hipLaunchKernelGGL((k_pcg_update<T, FinAlpha<T>, true>), grid, block, 0, c->stream, c->n,
                       A.xcdRemap, w.delta.p, w.deltaBest.p, w.r.p, pnew, w.Ap.p, st2, P, ps,
                       Merge{st, P2, nb, c->mirrorDev}, FinAlpha<T>{nullptr});
  } else {
    LAUNCH_CLASSED(1, capturing, (k_pcg_p_spmv<T, FinBeta<T>, false>), grid, block, A, w.wA.p,
                   pold, pnew, w.Ap.p, st, P, ps, none, FinBeta<T>{nullptr});
    launch_fin<1>(c, nb, true, kRunning, FinAlpha<T>{st});
    hipLaunchKernelGGL((k_pcg_update<T, FinAlpha<T>, false>), grid, block, 0, c->stream, c->n,
                       A.xcdRemap, w.delta.p, w.deltaBest.p, w.r.p, pnew, w.Ap.p, st, P, ps,
                       none, FinAlpha<T>{nullptr});
  }
  launch_fin<1>(c, nb, true, kRunning, FinPcgRes<T>{st});
}

template <typename T>
void enqueue_bicg_iter(hipfoam_context* c, const LduOp<T>& A, const T* rD, int parity,
                       bool capturing) {
  Work<T>& w = work<T>(c);
  const int bs = row_block(c);
  const dim3 grid(row_grid(c)), block(bs);
  KState* st = c->kst.p;
  double* P = c->partials.p;
  const int ps = c->pstride;
  const int n = c->n, rm = A.xcdRemap;
  const int nb = grid.x;
  T* pold = parity ? w.p1.p : w.p0.p;
  T* pnew = parity ? w.p0.p : w.p1.p;
  T* AyAold = parity ? w.AyA1.p : w.AyA.p;
  T* AyAnew = parity ? w.AyA.p : w.AyA1.p;
  // merged: alpha's stage runs in k_bicg_sz_spmv (state kst -> kst2,
  // partials P -> partials2) and the |sA| test and omega in
  // k_bicg_update_rho (kst2 -> kst, partials2 -> P)
  const bool merge = merge_stages(c);
  KState* st2 = merge ? c->kst2.p : st;
  double* P2 = merge ? c->partials2.p : P;
  const Merge none{nullptr, nullptr, 0};
  LAUNCH_CLASSED(2, capturing, (k_bicg_py_spmv<T>), grid, block, A, w.r.p, rD, pold, pnew,
                 AyAold, AyAnew, w.yA.p, w.rA0.p, st, P, ps);
  if (merge) {
    LAUNCH_CLASSED(2, capturing, (k_bicg_sz_spmv<T, FinBicgAlpha<T>, true>), grid, block, A,
                   w.r.p, rD, AyAnew, w.sA.p, w.zA.p, w.tA.p, st, P2, ps, Merge{st2, P, nb},
                   FinBicgAlpha<T>{nullptr});
    hipLaunchKernelGGL((k_bicg_update_rho<T, FinBicgSOmega<T>, true>), grid, block, 0,
                       c->stream, n, rm, w.delta.p, w.deltaBest.p, w.yA.p, w.zA.p, w.r.p,
                       w.sA.p, w.tA.p, w.rA0.p, st2, P, ps, Merge{st, P2, nb, c->mirrorDev},
                       FinBicgSOmega<T>{nullptr});
  } else {
    launch_fin<1>(c, nb, true, kRunning, FinBicgAlpha<T>{st});
    LAUNCH_CLASSED(2, capturing, (k_bicg_sz_spmv<T, FinBicgAlpha<T>, false>), grid, block, A,
                   w.r.p, rD, AyAnew, w.sA.p, w.zA.p, w.tA.p, st, P, ps, none,
                   FinBicgAlpha<T>{nullptr});
    launch_fin<3>(c, nb, true, kRunning, FinBicgSOmega<T>{st});
    hipLaunchKernelGGL((k_bicg_update_rho<T, FinBicgSOmega<T>, false>), grid, block, 0,
                       c->stream, n, rm, w.delta.p, w.deltaBest.p, w.yA.p, w.zA.p, w.r.p,
                       w.sA.p, w.tA.p, w.rA0.p, st, P, ps, none, FinBicgSOmega<T>{nullptr});
  }
  launch_fin<2>(c, nb, true, kRunningOrEarly, FinBicgUpdateRho<T>{st});
}

// rho = rA0.rA and beta of the first BiCGStab iteration (later iterations
// reduce rho in their update kernel); *kst must be initialised.
template <typename T>
int enqueue_bicg_start(hipfoam_context* c, const LduOp<T>& A) {
  Work<T>& w = work<T>(c);
  const dim3 grid(row_grid(c)), block(row_block(c));
  hipLaunchKernelGGL((k_bicg_rho<T>), grid, block, 0, c->stream, c->n, A.xcdRemap, w.rA0.p,
                     w.r.p, c->kst.p, c->partials.p, c->pstride);
  launch_fin<1>(c, grid.x, true, kRunning, FinBicgRho<T>{c->kst.p});
  CKL();
  return HIPFOAM_OK;
}

// Iteration `first + k` uses p-buffer parity (first + k) & 1; graphs are
// captured with first = 0 and an even chunk, so every replay starts even.
template <typename T>
void enqueue_chunk(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
                   int chunk, bool capturing, long first = 0) {
  for (int k = 0; k < chunk; ++k) {
    if (kind == KIND_PCG)
      enqueue_pcg_iter<T>(c, A, rD, order, (int)((first + k) & 1), capturing);
    else
      enqueue_bicg_iter<T>(c, A, rD, (int)((first + k) & 1), capturing);
  }
}

template <typename T>
int get_graph(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
              int chunk, hipGraphExec_t* out) {
  GraphKey key{(int)kind, (int)sizeof(T), order, chunk, A.direct + 2 * A.slots + 4 * A.slotSym,
               A.xcdRemap,
               row_block(c), row_grid(c), A.shortRows, c->gen};
  auto it = c->graphs.find(key);
  if (it != c->graphs.end()) {
    *out = it->second;
    return HIPFOAM_OK;
  }
  hipGraph_t g = nullptr;
  CK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_chunk<T>(c, kind, A, rD, order, chunk, true);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return fail(c, HIPFOAM_ERR_DEVICE, std::string("capture: ") + hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, HIPFOAM_ERR_DEVICE, std::string("instantiate: ") + hipGetErrorString(e));
  c->graphs[key] = ex;
  *out = ex;
  return HIPFOAM_OK;
}

// Runs the device-resident iteration loop until the device state leaves
// RUNNING.  Chunks of `chunk` iterations are launched back to back; the host
// polls the status of the chunk before the one just launched, so one chunk
// stays queued behind the GPU.  The status is the one the chunk's kernels
// posted to the context's pinned mirror word (every iteration has a stage
// that posts it), read once the chunk's event has completed; without a
// mirror it is copied after each chunk.  In kernel-timing mode every
// iteration is launched eagerly and polled synchronously, so no timed launch
// is a no-op.  Vectors and *kst must be initialised.
template <typename T>
int run_loop(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
             int maxIter, double* gpuSeconds) {
  const bool timing = c->timer.on;
  // graphChunk -1 (auto): eager launches from kEagerRows rows on, where the
  // kernels outlast their launches and the host stays ahead of the GPU
  // without a graph launch's boundary every few iterations; graphs of 4
  // iterations below, where the launches themselves would set the pace
  const bool eager = c->sched.graphChunk == 0 || timing ||
                     (c->sched.graphChunk < 0 && c->n >= kEagerRows);
  const int chunk = timing ? 1 : (c->sched.graphChunk > 0 ? c->sched.graphChunk : 4);
  hipGraphExec_t ex = nullptr;
  if (!eager) {
    int rc = get_graph<T>(c, kind, A, rD, order, chunk, &ex);
    if (rc) return rc;
  }
  CK(hipEventRecord(c->evT0, c->stream));
  const long maxChunks = (long)std::max(maxIter, 1) / chunk + 3;
  int slot = 0;
  bool havePrev = false;
  for (long k = 0; k < maxChunks + 2; ++k) {
    if (eager) {
      enqueue_chunk<T>(c, kind, A, rD, order, chunk, false, k * chunk);
      CKL();
    } else {
      CK(hipGraphLaunch(ex, c->stream));
    }
    if (!c->mirrorHost)
      CK(hipMemcpyAsync(&c->hstatus[slot], &c->kst.p->status, sizeof(int32_t),
                        hipMemcpyDeviceToHost, c->stream));
    CK(hipEventRecord(c->evPoll[slot], c->stream));
    // the status as of the chunk whose event completed last (or later)
    auto polled = [&](int s) {
      return c->mirrorHost ? __atomic_load_n(c->mirrorHost, __ATOMIC_ACQUIRE) : c->hstatus[s];
    };
    if (timing) {
      CK(hipEventSynchronize(c->evPoll[slot]));
      if (c->timer.used + 64 > c->timer.ev.size()) {
        int rc = timer_flush(c);
        if (rc) return rc;
      }
      if (polled(slot) != ST_RUNNING) break;
      continue;
    }
    if (havePrev) {
      CK(hipEventSynchronize(c->evPoll[slot ^ 1]));
      if (polled(slot ^ 1) != ST_RUNNING) break;
    }
    havePrev = true;
    slot ^= 1;
  }
  CK(hipEventRecord(c->evT1, c->stream));
  CK(hipEventSynchronize(c->evT1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->evT0, c->evT1));
  if (gpuSeconds) *gpuSeconds += 1e-3 * ms;
  if (timing) {
    int rc = timer_flush(c);
    if (rc) return rc;
  }
  CK(hipMemcpy(c->hk, c->kst.p, sizeof(KState), hipMemcpyDeviceToHost));
  if (c->hk->status == ST_RUNNING)
    return fail(c, HIPFOAM_ERR_STATE, "inner loop did not terminate within maxIter chunks");
  return HIPFOAM_OK;
}

// Meshes of at most this many 256-row tiles run the Chebyshev PCG loop as one
// persistent launch on one XCD by default (schedule.persistent -1), for the
// FP32 inner loop and the plain-FP64 loop.  Refined / plain-FP64 solve,
// launched against persistent, lognormal Poisson, same box: 10k cells (40
// tiles) 2.01 / 1.12 ms and 2.77 / 2.00 ms; 24^3 (54) 1.26 / 0.71 and
// 1.50 / 1.10; 28^3 (86) 1.46 / 1.01 and 1.74 / 1.55; 32^3 (128) 1.47 / 1.11
// and 1.96 / 2.01; 36^3 (183) 1.62 / 1.55 and 2.12 / 3.07; 40^3 (250) 1.73 /
// 1.99 and 2.39 / 4.45.  One XCD's share of the chip bounds the persistent
// loop; the launched one pays a fixed ~4.4 us per kernel.
constexpr int kPersistAutoTiles32 = 160;
constexpr int kPersistAutoTiles64 = 96;

// Whether the inner loop runs as one persistent launch (k_pcg_persistent,
// k_bicg_persistent): Chebyshev PCG or BiCGStab on the per-tile launch shape
// (one 256-thread workgroup per tile, second stages merged), whose rows,
// partials and fold order it reproduces, up to kPersistAutoTiles32/64 tiles
// (persistent -1) or every per-tile mesh (1); never under kernel timing,
// whose per-launch events need the launched kernels.
template <typename T>
inline bool persistent_ok(const hipfoam_context* c, Kind kind, int order) {
  if (c->sched.persistent == 0 || !c->xsync.p || c->timer.on) return false;
  if (kind == KIND_PCG && order < 2) return false;
  const int ntiles = (c->n + kSlotTile - 1) / kSlotTile;
  if (row_block(c) != kSlotTile || row_grid(c) != ntiles || !merge_stages(c)) return false;
  const int autoTiles = sizeof(T) == sizeof(float) ? kPersistAutoTiles32 : kPersistAutoTiles64;
  return ntiles <= (c->sched.persistent > 0 ? kAutoSmallTiles : autoTiles);
}

// Whether the inner loop runs in one workgroup (k_pcg_local, k_bicg_local):
// Chebyshev PCG or BiCGStab on the per-tile launch shape, whose rows,
// partials and fold order it reproduces, on a mesh of at most
// local_max_rows<T>() (PCG) or local_bicg_max_rows<T>() rows whose tiles are
// all row-slot tiles (schedule.localLoop -1, the default); never under kernel
// timing.
template <typename T>
inline bool local_ok(const hipfoam_context* c, Kind kind, int order, const LduOp<T>& A) {
  if (c->sched.localLoop == 0 || c->timer.on || (kind == KIND_PCG && order < 2)) return false;
  const int maxRows = kind == KIND_PCG ? local_max_rows<T>() : local_bicg_max_rows<T>();
  if (!A.slots || !c->slotOK || c->slotFaceRows != 0 || c->n > maxRows) return false;
  const int ntiles = (c->n + kSlotTile - 1) / kSlotTile;
  return row_block(c) == kSlotTile && row_grid(c) == ntiles && merge_stages(c);
}

// The whole iteration loop in one workgroup (BiCGStab: its start, then the
// loop), enqueued between the timing events evT0 and evT1; vectors and the
// state must be initialised.  The state leaves RUNNING within maxIter + 4
// iterations.
template <typename T>
int local_enqueue(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
                  int maxIter) {
  Work<T>& w = work<T>(c);
  const int loops = std::max(maxIter, 1) + 4;
  if (kind == KIND_BICG) {
    int rc = enqueue_bicg_start<T>(c, A);
    if (rc) return rc;
  }
  c->localLaunches += 1;
  CK(hipEventRecord(c->evT0, c->stream));
  if (kind == KIND_PCG) {
    const PcgVectors<T> V{w.r.p,     w.r1.p,        w.p0.p, w.p1.p, w.wA.p, w.Ap.p,
                          w.delta.p, w.deltaBest.p, w.z0.p, w.z1.p, w.w.p};
    const auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(1), dim3(kLocalThreads), 0, c->stream, A, rD, V, c->kst.p,
                         c->kst2.p, c->partials.p, c->partials2.p, c->pstride, order, loops);
    };
    if constexpr (sizeof(T) == sizeof(float)) {
      if (local_pcg_passes<T>(c->n) == 4)
        launch(k_pcg_local<T, 4>);
      else
        launch(k_pcg_local<T, 6>);
    } else {
      launch(k_pcg_local<T, 3>);
    }
  } else {
    const BicgVectors<T> V{w.r.p,  w.p0.p, w.p1.p, w.AyA.p, w.AyA1.p,  w.yA.p,
                           w.rA0.p, w.sA.p, w.zA.p, w.tA.p,  w.delta.p, w.deltaBest.p};
    hipLaunchKernelGGL((k_bicg_local<T>), dim3(1), dim3(kLocalThreads), 0, c->stream, A, rD,
                       V, c->kst.p, c->kst2.p, c->partials.p, c->partials2.p, c->pstride,
                       loops);
  }
  CKL();
  CK(hipEventRecord(c->evT1, c->stream));
  return HIPFOAM_OK;
}

// After evT1 has completed and the state has been copied to c->hk: the GPU
// time of the loop, and the check that it ended.
int local_finish(hipfoam_context* c, double* gpuSeconds) {
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->evT0, c->evT1));
  if (gpuSeconds) *gpuSeconds += 1e-3 * ms;
  if (c->hk->status != ST_RUNNING) return HIPFOAM_OK;
  return fail(c, HIPFOAM_ERR_STATE, "inner loop did not terminate within maxIter iterations");
}

template <typename T>
int local_loop(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
               int maxIter, double* gpuSeconds) {
  int rc = local_enqueue<T>(c, kind, A, rD, order, maxIter);
  if (rc) return rc;
  CK(hipEventSynchronize(c->evT1));
  CK(hipMemcpy(c->hk, c->kst.p, sizeof(KState), hipMemcpyDeviceToHost));
  return local_finish(c, gpuSeconds);
}

// One persistent launch of the whole iteration loop, enqueued between the
// timing events evT0 and evT1; vectors and the state must be initialised
// (and the BiCGStab start enqueued).
template <typename T>
int persistent_enqueue(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
                       int maxIter) {
  Work<T>& w = work<T>(c);
  const int ntiles = (c->n + kSlotTile - 1) / kSlotTile;
  const int K = std::min(ntiles, kPersistMaxGroups);
  const int loops = std::max(maxIter, 1) + 4;
  uint32_t* xs = c->xsync.p;
  const XcdSync S{xs, xs + kPersistMaxGroups, xs + 2 * kPersistMaxGroups,
                  xs + 2 * kPersistMaxGroups + 1};
  const dim3 grid(8 * K), block(kSlotTile);
  c->persistLaunches += 1;
  CK(hipEventRecord(c->evT0, c->stream));
  if (kind == KIND_PCG) {
    const PcgVectors<T> V{w.r.p,     w.r1.p,        w.p0.p, w.p1.p, w.wA.p, w.Ap.p,
                          w.delta.p, w.deltaBest.p, w.z0.p, w.z1.p, w.w.p};
    hipLaunchKernelGGL((k_pcg_persistent<T>), grid, block, 0, c->stream, A, rD, V, c->kst.p,
                       c->kst2.p, c->partials.p, c->partials2.p, c->pstride, order, K, loops, S);
  } else {
    const BicgVectors<T> V{w.r.p,  w.p0.p, w.p1.p, w.AyA.p, w.AyA1.p,  w.yA.p,
                           w.rA0.p, w.sA.p, w.zA.p, w.tA.p,  w.delta.p, w.deltaBest.p};
    hipLaunchKernelGGL((k_bicg_persistent<T>), grid, block, 0, c->stream, A, rD, V, c->kst.p,
                       c->kst2.p, c->partials.p, c->partials2.p, c->pstride, K, loops, S);
  }
  if (c->simAbandon > 0) {
    c->simAbandon -= 1;
    hipLaunchKernelGGL(k_simulate_abandon, dim3(1), dim3(64), 0, c->stream, c->kst.p, S.abortw);
  }
  CKL();
  CK(hipEventRecord(c->evT1, c->stream));
  return HIPFOAM_OK;
}

// After evT1 has completed and the state has been copied to c->hk: the GPU
// time of the launch, and whether it ended the loop.  A state left running
// means the launch was abandoned (*abandoned: the barrier words are reset
// and the caller solves again, from the restored start) or ran out of
// iterations (an error).
int persistent_finish(hipfoam_context* c, double* gpuSeconds, bool* abandoned) {
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, c->evT0, c->evT1));
  if (gpuSeconds) *gpuSeconds += 1e-3 * ms;
  if (c->hk->status != ST_RUNNING) return HIPFOAM_OK;
  uint32_t* xs = c->xsync.p;
  int32_t* ab = &c->hstatus[2];
  CK(hipMemcpy(ab, xs + 2 * kPersistMaxGroups, sizeof(int32_t), hipMemcpyDeviceToHost));
  CK(hipMemsetAsync(xs, 0, sizeof(uint32_t) * kXcdSyncWords, c->stream));
  CK(hipStreamSynchronize(c->stream));
  if (*ab != 0) {
    c->persistAbandoned += 1;
    c->persistLastAbort = (int64_t)(uint32_t)*ab;
    *abandoned = true;
    return HIPFOAM_OK;
  }
  return fail(c, HIPFOAM_ERR_STATE, "inner loop did not terminate within maxIter iterations");
}

// The persistent launch, waited for.
template <typename T>
int persistent_loop(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
                    int maxIter, double* gpuSeconds, bool* abandoned) {
  int rc = persistent_enqueue<T>(c, kind, A, rD, order, maxIter);
  if (rc) return rc;
  CK(hipEventSynchronize(c->evT1));
  CK(hipMemcpy(c->hk, c->kst.p, sizeof(KState), hipMemcpyDeviceToHost));
  return persistent_finish(c, gpuSeconds, abandoned);
}

// The iteration loop of an initialised solve (BiCGStab: its start, then the
// loop): one launch in one workgroup where local_ok allows, else one
// persistent launch where persistent_ok allows, else launched iterations
// (run_loop).  If the persistent launch was abandoned, reinit() restores the
// vectors and the state the loop starts from, and the launched loop runs.
template <typename T, typename Reinit>
int iterate(hipfoam_context* c, Kind kind, const LduOp<T>& A, const T* rD, int order,
            int maxIter, double* gpuSeconds, Reinit reinit) {
  if (local_ok<T>(c, kind, order, A))
    return local_loop<T>(c, kind, A, rD, order, maxIter, gpuSeconds);
  if (persistent_ok<T>(c, kind, order)) {
    if (kind == KIND_BICG) {
      int rc = enqueue_bicg_start<T>(c, A);
      if (rc) return rc;
    }
    bool abandoned = false;
    int rc = persistent_loop<T>(c, kind, A, rD, order, maxIter, gpuSeconds, &abandoned);
    if (rc || !abandoned) return rc;
    rc = reinit();
    if (rc) return rc;
  }
  if (kind == KIND_BICG) {
    int rc = enqueue_bicg_start<T>(c, A);
    if (rc) return rc;
  }
  return run_loop<T>(c, kind, A, rD, order, maxIter, gpuSeconds);
}

// An inner solve left running on the stream (inner_solve with a pending
// record): its loop (one workgroup, or persistent on one XCD), the
// best-iterate restore and the copy of its final state are enqueued;
// inner_finish reads its statistics once the caller has synchronised the
// stream.
struct InnerPending {
  bool on = false;
  bool persist = false;  // the loop is a persistent launch, which may be abandoned
  double tw0 = 0.0;
};

// Inner correction solve in reference-backend semantics
// (metalFoamBackend::pcg / ::bicgstab): r64 -> FP32 residual (minus *mean,
// times proj when given, when mean is given), delta0 = 0, device loop,
// best-iterate restore.  With pending (not NULL), a loop that runs as one
// launch (one workgroup, or persistent on one XCD) is only enqueued,
// together with the restore (on the device) and the state copy: the caller
// queues its next steps behind it, guarded by the state's status, and calls
// inner_finish after its next synchronisation.
int inner_solve(hipfoam_context* c, Kind kind, const double* r64, const double* mean,
                const double* proj, double target, int maxIter, int ips, int order,
                double beta, double ratio, hipfoam_stats* stats,
                InnerPending* pending = nullptr) {
  const double tw0 = now_seconds();
  hipfoam_stats st{};
  if (!c->good || c->n == 0) {
    st.singular = 1;
    if (stats) *stats = st;
    return HIPFOAM_OK;
  }
  if (order > kMaxCheb + 1) return fail(c, HIPFOAM_ERR_ARG, "chebyshevOrder > 33");
  Work<float>& w = c->w32;
  const int n = c->n;
  KState s{};
  s.mode = MODE_INNER;
  s.target = target;
  s.maxIter = maxIter;
  s.ips = ips < 1 ? 1 : ips;
  s.batchSize = std::min(s.ips, std::max(maxIter, 0));
  s.batchLeft = s.batchSize;
  s.bestRes = 1.0e300;
  s.status = maxIter > 0 ? ST_RUNNING : ST_MAXITER;
  if (kind == KIND_PCG && order > 1) cheb_constants(s, order, beta, ratio, true);
  // vectors and both state buffers (merged stages read the first iteration's
  // state from kst2)
  auto init = [&]() -> int {
    const dim3 g1(blocks_for(n, 256)), b1(256);
    hipLaunchKernelGGL((k_inner_init<float>), g1, b1, 0, c->stream, n, r64, mean, proj, w.r.p,
                       kind == KIND_BICG ? w.rA0.p : nullptr, w.delta.p, w.p0.p, w.p1.p,
                       kind == KIND_BICG ? w.AyA.p : nullptr, state_words(s), c->kst.p,
                       c->kst2.p);
    CKL();
    return HIPFOAM_OK;
  };
  int rc0 = init();
  if (rc0) return rc0;
  double gpu = 0;
  if (s.status == ST_RUNNING) {
    const LduOp<float> A = op32(c);
    const int ord = kind == KIND_PCG ? order : 1;
    const bool local = pending && local_ok<float>(c, kind, ord, A);
    const bool persist = pending && !local && persistent_ok<float>(c, kind, ord);
    if (local || persist) {
      int rc = HIPFOAM_OK;
      if (local) {
        rc = local_enqueue<float>(c, kind, A, c->rD32.p, ord, maxIter);
      } else {
        if (kind == KIND_BICG) rc = enqueue_bicg_start<float>(c, A);
        if (!rc) rc = persistent_enqueue<float>(c, kind, A, c->rD32.p, ord, maxIter);
      }
      if (rc) return rc;
      hipLaunchKernelGGL((k_restore_best<float>), dim3(blocks_for(n, 256)), dim3(256), 0,
                         c->stream, n, c->kst.p, w.delta.p, w.deltaBest.p);
      CKL();
      CK(hipMemcpyAsync(c->hk, c->kst.p, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
      pending->on = true;
      pending->persist = persist;
      pending->tw0 = tw0;
      return HIPFOAM_OK;
    }
    int rc = iterate<float>(c, kind, A, c->rD32.p, ord, maxIter, &gpu, init);
    if (rc) return rc;
  } else {
    CK(hipStreamSynchronize(c->stream));
    *c->hk = s;
  }
  const KState& h = *c->hk;
  st.nIterations = h.it;
  st.finalResSum = (s.status == ST_RUNNING) ? h.res : 0.0;
  st.converged = h.status == ST_CONVERGED;
  st.singular = h.status == ST_SINGULAR;
  st.stagnated = h.status == ST_STAGNATED;
  // best iterate (metalFoamBackend.C:718-727); the |sA| exit keeps delta
  if (s.status == ST_RUNNING && h.bestRes < st.finalResSum) {
    CK(hipMemcpyAsync(w.delta.p, w.deltaBest.p, sizeof(float) * n, hipMemcpyDeviceToDevice,
                      c->stream));
    st.finalResSum = h.bestRes;
  }
  st.tGpuSeconds = gpu;
  st.tWallSeconds = now_seconds() - tw0;
  if (stats) *stats = st;
  return HIPFOAM_OK;
}

// Statistics of a pending inner solve (its loop ran from RUNNING; the state
// is in c->hk and the restore has run), once the stream has been
// synchronised: the same as inner_solve reports.  *redo: the persistent
// launch was abandoned, so nothing queued behind it ran (its guards saw the
// state still RUNNING); the caller solves again with inner_solve.
int inner_finish(hipfoam_context* c, InnerPending& pending, hipfoam_stats* stats, bool* redo) {
  hipfoam_stats st{};
  double gpu = 0;
  pending.on = false;
  *redo = false;
  int rc = pending.persist ? persistent_finish(c, &gpu, redo) : local_finish(c, &gpu);
  if (rc || *redo) return rc;
  const KState& h = *c->hk;
  st.nIterations = h.it;
  st.finalResSum = h.res;
  st.converged = h.status == ST_CONVERGED;
  st.singular = h.status == ST_SINGULAR;
  st.stagnated = h.status == ST_STAGNATED;
  if (h.bestRes < st.finalResSum) st.finalResSum = h.bestRes;
  st.tGpuSeconds = gpu;
  st.tWallSeconds = now_seconds() - pending.tw0;
  if (stats) *stats = st;
  return HIPFOAM_OK;
}

// -------------------------------------------------------------------------
// staging
// -------------------------------------------------------------------------

int ensure_vectors(hipfoam_context* c) {
  const int n = c->n;
  CK(c->w32.ensure(n));
  CK(c->psi0.ensure(n));
  CK(c->psi1.ensure(n));
  CK(c->b.ensure(n));
  CK(c->r64.ensure(n));
  CK(c->u64.ensure(n));
  CK(c->scratch64.ensure(n));
  CK(c->diag32.ensure(n));
  CK(c->rD32.ensure(n));
  CK(c->diag64.ensure(n));
  CK(c->rD64.ensure(n));
  return HIPFOAM_OK;
}

// Stage an FP64 matrix (copy into library buffers, losort-ordered lower,
// FP32 faces).  kind: hipMemcpyHostToDevice or hipMemcpyDeviceToDevice.
int stage_matrix(hipfoam_context* c, const double* diag, const double* upper,
                 const double* lower, hipMemcpyKind kind) {
  const int n = c->n, nf = c->nf;
  if (!diag || !upper) return fail(c, HIPFOAM_ERR_ARG, "diag/upper must not be NULL");
  CK(c->diag64.ensure(n));
  CK(c->upper64.ensure(nf));
  CK(c->lowerL64.ensure(nf));
  CK(c->upper32.ensure(nf));
  CK(c->lowerL32.ensure(nf));
  XF(c->diag64.p, diag, sizeof(double) * n, kind);
  XF(c->upper64.p, upper, sizeof(double) * nf, kind);
  c->sym64 = (lower == nullptr);
  c->sym32 = c->sym64;
  if (lower) {
    CK(c->lower64.ensure(nf));
    CK(c->lower32.ensure(nf));
    XF(c->lower64.p, lower, sizeof(double) * nf, kind);
  }
  if (nf > 0) {
    hipLaunchKernelGGL(k_stage_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, c->stream, nf,
                       c->losort.p, c->upper64.p, lower ? c->lower64.p : nullptr,
                       c->lowerL64.p, c->upper32.p, c->lowerL32.p,
                       lower ? c->lower32.p : nullptr);
    CKL();
  }
  if (c->slotOK && n > 0) {
    CK(c->slotCoef32.ensure((size_t)2 * kShortRow * n));
    CK(c->slotCoef64.ensure((size_t)2 * kShortRow * n));
    const int sym = lower == nullptr;
    hipLaunchKernelGGL((k_build_slot_coeffs<float>), dim3(blocks_for(n, 256)), dim3(256), 0,
                       c->stream, n, nullptr, c->ownStart.p, c->losortStart.p, c->slotOff.p,
                       c->slotTile.p, c->slotRow.p, sym, c->upper32.p, c->lowerL32.p,
                       c->slotCoef32.p);
    hipLaunchKernelGGL((k_build_slot_coeffs<double>), dim3(blocks_for(n, 256)), dim3(256), 0,
                       c->stream, n, nullptr, c->ownStart.p, c->losortStart.p, c->slotOff.p,
                       c->slotTile.p, c->slotRow.p, sym, c->upper64.p, c->lowerL64.p,
                       c->slotCoef64.p);
    CKL();
  }
  c->have64 = true;
  return HIPFOAM_OK;
}

// FP32 inner operator diagonal: diag* = deflate ? diag - u : diag (+ rD, and
// the Gershgorin bound when wantBeta).  Writes os->bail / os->beta*.
int prepare_inner(hipfoam_context* c, int deflate, int wantBeta) {
  const int bs = row_block(c);
  LduOp<double> A = op64(c);
  hipLaunchKernelGGL((k_prepare_inner<float>), dim3(row_grid(c)), dim3(bs), 0, c->stream,
                     A, c->upper64.p, c->losort.p, c->u64.p, deflate, wantBeta, c->diag32.p,
                     c->rD32.p, c->ost.p, c->partials.p, c->pstride);
  launch_fin<2, true>(c, row_grid(c), false, 0, FinPrepare{c->ost.p, deflate, wantBeta, c->n});
  CKL();
  c->have32 = true;
  return HIPFOAM_OK;
}

int read_ostate(hipfoam_context* c) {
  CK(hipMemcpyAsync(c->ho, c->ost.p, sizeof(OState), hipMemcpyDeviceToHost, c->stream));
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

// initial FP64 residual + normFactor (metalFoamSolver.C:101-129): psi in psiCur;
// read: wait for it and copy the outer state to the host.
int initial_residual(hipfoam_context* c, const double* psiCur, const double* src,
                     bool read = true) {
  const int bs = row_block(c);
  LduOp<double> A = op64(c);
  const dim3 g(row_grid(c)), bl(bs);
  hipLaunchKernelGGL(k_rowsum, g, bl, 0, c->stream, A, psiCur, c->u64.p, c->ost.p, c->partials.p,
                     c->pstride);
  launch_fin<4>(c, g.x, false, 0, FinRowsum{c->ost.p, c->n});
  hipLaunchKernelGGL(k_initial_residual, g, bl, 0, c->stream, A, psiCur, src, c->u64.p,
                     c->r64.p, c->ost.p, c->partials.p, c->pstride);
  launch_fin<3>(c, g.x, false, 0, FinInitialResidual{c->ost.p, c->n});
  CKL();
  return read ? read_ostate(c) : HIPFOAM_OK;
}

bool of_converged_h(double res, double res0, double tol, double relTol) {
  return res < tol || (relTol > kSmall && res < relTol * res0);
}

int check_ready(hipfoam_context* c) {
  if (!c) return HIPFOAM_ERR_ARG;
  if (!c->good) return fail(c, HIPFOAM_ERR_DEVICE, "no usable device: " + c->desc);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, HIPFOAM_ERR_DEVICE, "hipSetDevice");
  if (!c->ownStart.p) return fail(c, HIPFOAM_ERR_STATE, "set_topology not called");
  return HIPFOAM_OK;
}

// -------------------------------------------------------------------------
// metalFoamSolver::solveImpl with the FP64 steps on the device
// -------------------------------------------------------------------------

// kind: where psi, source and residual live; matKind: where diag, upper and
// lower live (default: the same)
int refined_solve(hipfoam_context* c, int solver, const double* diag, const double* upper,
                  const double* lower, double* psi, const double* source,
                  const hipfoam_controls* ctl, hipfoam_performance* perf, double* residual,
                  hipMemcpyKind kind, hipMemcpyKind matKind = hipMemcpyDefault) {
  if (matKind == hipMemcpyDefault) matKind = kind;
  const double t0 = now_seconds();
  int rc = check_ready(c);
  if (rc) return rc;
  if (!psi || !source || !ctl) return fail(c, HIPFOAM_ERR_ARG, "NULL psi/source/controls");
  if (solver != HIPFOAM_PCG && solver != HIPFOAM_PBICGSTAB)
    return fail(c, HIPFOAM_ERR_ARG, "unknown solver");
  const bool useBiCG = solver == HIPFOAM_PBICGSTAB;
  const int n = c->n;
  hipfoam_performance pf{};
  if (ctl->deflate < 0 || ctl->deflate > 3)
    return fail(c, HIPFOAM_ERR_ARG, "deflate must be 0 (off), 1 (on), 2 (auto) or 3 (pinned)");
  rc = ensure_vectors(c);
  if (rc) return rc;
  // lower is used only by PBiCGStab on asymmetric matrices (metalFoamSolver.C:214-219)
  rc = stage_matrix(c, diag, upper, lower, matKind);
  if (rc) return rc;
  const double tStage = now_seconds() - t0;
  // device arrays: refine in the caller's psi and read its source in place
  // (psi ping-pongs with psi1; the last copy back is needed only when the
  // solution ends in psi1); host arrays, or psi overlapping source: copies
  const bool inPlace =
      kind == hipMemcpyDeviceToDevice && (psi + n <= source || source + n <= psi);
  double* psiCur = inPlace ? psi : c->psi0.p;
  double* psiNxt = c->psi1.p;
  const double* src = inPlace ? source : c->b.p;
  if (!inPlace) {
    XF(c->psi0.p, psi, sizeof(double) * n, kind);
    XF(c->b.p, source, sizeof(double) * n, kind);
  }
  // deflation (metalFoamSolver.C:170-190) is decided on the device from the
  // row sums (deflate 2 and 3: automatic; 3 also projects the inner
  // right-hand side along u instead of the constant, no reference
  // counterpart), so the inner operator is prepared before the first read
  // and one round trip returns the initial residual and the preparation;
  // a solve converged at start discards the preparation
  const bool pinnedProj = ctl->deflate == 3;
  const int mode = useBiCG || ctl->deflate == 0 ? 0 : (ctl->deflate == 1 ? 1 : kDeflateAuto);
  const int chebOrder = useBiCG ? 1 : ctl->chebyshevOrder;
  const bool orderOK = chebOrder >= 1 && chebOrder <= kMaxCheb + 1;
  const bool wantBeta = chebOrder > 1;
  rc = initial_residual(c, psiCur, src, !orderOK);
  if (rc) return rc;
  const double tc0 = now_seconds();
  if (orderOK) {
    // metalFoamSolver.C:192-249: diag0 = diag - u, sign-flip bail, setCoeffs, Gershgorin
    rc = prepare_inner(c, mode, wantBeta);
    if (rc) return rc;
    rc = read_ostate(c);
    if (rc) return rc;
  }
  const OState& o = *c->ho;
  const double normFactor = o.normFactor;
  pf.normFactor = normFactor;
  pf.initialResidual = o.initRes;
  pf.finalResidual = o.initRes;
  if (ctl->minIter > 0 ||
      !of_converged_h(pf.finalResidual, pf.initialResidual, ctl->tolerance, ctl->relTol)) {
    if (!orderOK) return fail(c, HIPFOAM_ERR_ARG, "chebyshevOrder must be in [1, 33]");
    bool deflate = o.deflate != 0;
    if (deflate && o.bail > 0.0) {
      deflate = false;
      rc = prepare_inner(c, 0, wantBeta);
      if (rc) return rc;
      rc = read_ostate(c);
      if (rc) return rc;
    }
    pf.tConvert = tStage + (now_seconds() - tc0);
    pf.deflated = deflate;
    const double chebRatio = ctl->chebyshevRatio;
    const double chebBeta = wantBeta ? (deflate ? o.betaDefl : o.betaExact) : 2.0;
    pf.chebBeta = chebBeta;
    // metalFoamSolver.C:251-257
    const double target = std::max(ctl->tolerance, ctl->relTol * pf.initialResidual) * normFactor;
    const double uu = o.uu;
    int totalIter = 0, outer = 0;
    bool singular = false;
    double prevRes = pf.initialResidual;
    const int bs = row_block(c);
    const dim3 g(row_grid(c)), bl(bs);
    for (outer = 0; outer <= ctl->maxRestarts; ++outer) {
      hipfoam_stats st{};
      // zero-mean projection of the inner RHS (metalFoamSolver.C:278-287);
      // an inner solve that runs as one launch stays pending, so the
      // refinement below queues behind it (guarded by its status) and one
      // round trip ends the restart
      const double* mean =
          deflate ? (pinnedProj ? &c->ost.p->rproj : &c->ost.p->rmean) : nullptr;
      const double* proj = deflate && pinnedProj ? c->u64.p : nullptr;
      const int innerMax = ctl->maxIter - totalIter;
      // psi += delta - mean(delta); r = b - A psi (metalFoamSolver.C:318-340),
      // the new psi in psiNxt; the constant-mode correction
      // (metalFoamSolver.C:342-361)
      auto refine = [&](const int32_t* guard) -> int {
        LduOp<double> A = op64(c);
        if (deflate) {
          hipLaunchKernelGGL((k_delta_mean<float>), g, bl, 0, c->stream, n, A.xcdRemap,
                             c->w32.delta.p, c->ost.p, c->partials.p, c->pstride, guard);
          launch_fin_guarded<1>(c, g.x, guard, FinDeltaMean{c->ost.p, n});
        }
        hipLaunchKernelGGL((k_refine_residual<float>), g, bl, 0, c->stream, A, psiCur,
                           c->w32.delta.p, psiNxt, src, c->u64.p, c->r64.p, deflate ? 1 : 0,
                           c->ost.p, c->partials.p, c->pstride, guard);
        launch_fin_guarded<3>(c, g.x, guard, FinRefine{c->ost.p, n});
        if (deflate && uu > kVSmall) {
          hipLaunchKernelGGL(k_const_shift, g, bl, 0, c->stream, n, A.xcdRemap, psiNxt,
                             c->r64.p, c->u64.p, c->ost.p, c->partials.p, c->pstride, guard);
          launch_fin_guarded<2>(c, g.x, guard, FinShift{c->ost.p, n});
        }
        CKL();
        return read_ostate(c);
      };
      InnerPending pending;
      rc = inner_solve(c, useBiCG ? KIND_BICG : KIND_PCG, c->r64.p, mean, proj, target,
                       innerMax, ctl->iterationsPerSync, chebOrder, chebBeta, chebRatio, &st,
                       &pending);
      if (rc) return rc;
      rc = refine(pending.on ? &c->kst.p->status : nullptr);
      if (rc) return rc;
      if (pending.on) {
        bool redo = false;
        rc = inner_finish(c, pending, &st, &redo);
        if (rc) return rc;
        if (redo) {
          rc = inner_solve(c, useBiCG ? KIND_BICG : KIND_PCG, c->r64.p, mean, proj, target,
                           innerMax, ctl->iterationsPerSync, chebOrder, chebBeta, chebRatio,
                           &st);
          if (rc) return rc;
          rc = refine(nullptr);
          if (rc) return rc;
        }
      }
      std::swap(psiCur, psiNxt);
      totalIter += st.nIterations;
      pf.tGpu += st.tGpuSeconds;
      pf.tInner += st.tWallSeconds;
      singular = st.singular;
      pf.finalResidual = o.finalRes;
      pf.nIterations = totalIter;
      // hybrid polish (metalFoamSolver.C:371-391): exact operator, beta kept
      if (deflate && pf.finalResidual > prevRes / 1.5 && !st.singular) {
        deflate = false;
        pf.polished = 1;
        rc = prepare_inner(c, 0, 0);
        if (rc) return rc;
        if (ctl->verbose)
          printf("    %s switching to exact operator (polish)\n",
                 useBiCG ? "hipPBiCGStab" : "hipPCG");
      }
      prevRes = pf.finalResidual;
      if (ctl->verbose)
        printf("    %s restart %d: inner=%d fp32ResSum=%g fp64Res=%g%s%s\n",
               useBiCG ? "hipPBiCGStab" : "hipPCG", outer, st.nIterations, st.finalResSum,
               pf.finalResidual, st.stagnated ? " (stagnated)" : "",
               st.singular ? " (breakdown)" : "");
      // metalFoamSolver.C:404-414
      if (singular) break;
      if (totalIter >= ctl->maxIter) break;
      if (pf.nIterations >= ctl->minIter &&
          of_converged_h(pf.finalResidual, pf.initialResidual, ctl->tolerance, ctl->relTol))
        break;
    }
    pf.restarts = outer;
    pf.singular = singular;
  }
  pf.converged =
      of_converged_h(pf.finalResidual, pf.initialResidual, ctl->tolerance, ctl->relTol);
  const hipMemcpyKind back =
      kind == hipMemcpyHostToDevice ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (psiCur != psi) XF(psi, psiCur, sizeof(double) * n, back);
  if (residual) XF(residual, c->r64.p, sizeof(double) * n, back);
  CK(hipStreamSynchronize(c->stream));
  pf.tTotal = now_seconds() - t0;
  if (perf) *perf = pf;
  return HIPFOAM_OK;
}

// Plain FP64: OpenFOAM PCG / PBiCGStab loops on the device.
int fp64_solve(hipfoam_context* c, int solver, const double* diag, const double* upper,
               const double* lower, double* psi, const double* source,
               const hipfoam_controls* ctl, hipfoam_performance* perf, double* residual,
               hipMemcpyKind kind, hipMemcpyKind matKind = hipMemcpyDefault) {
  if (matKind == hipMemcpyDefault) matKind = kind;
  const double t0 = now_seconds();
  int rc = check_ready(c);
  if (rc) return rc;
  if (!psi || !source || !ctl) return fail(c, HIPFOAM_ERR_ARG, "NULL psi/source/controls");
  if (solver != HIPFOAM_PCG && solver != HIPFOAM_PBICGSTAB)
    return fail(c, HIPFOAM_ERR_ARG, "unknown solver");
  const bool useBiCG = solver == HIPFOAM_PBICGSTAB;
  const int n = c->n;
  hipfoam_performance pf{};
  rc = ensure_vectors(c);
  if (rc) return rc;
  CK(c->w64.ensure(n));
  rc = stage_matrix(c, diag, upper, useBiCG ? lower : nullptr, matKind);
  if (rc) return rc;
  XF(c->psi0.p, psi, sizeof(double) * n, kind);
  XF(c->b.p, source, sizeof(double) * n, kind);
  rc = initial_residual(c, c->psi0.p, c->b.p);
  if (rc) return rc;
  const OState& o = *c->ho;
  pf.normFactor = o.normFactor;
  pf.initialResidual = o.initRes;
  pf.finalResidual = o.initRes;
  Work<double>& w = c->w64;
  const int order = useBiCG ? 1 : ctl->chebyshevOrder;
  if (order > kMaxCheb + 1 || order < 1)
    return fail(c, HIPFOAM_ERR_ARG, "chebyshevOrder must be in [1, 33]");
  double* result = c->psi0.p;
  if (ctl->minIter > 0 ||
      !of_converged_h(pf.finalResidual, pf.initialResidual, ctl->tolerance, ctl->relTol)) {
    const int bs = row_block(c);
    const dim3 g1(blocks_for(n, 256)), b1(256);
    // Gershgorin bound of the exact operator for Chebyshev
    double beta = 2.0;
    if (order > 1) {
      rc = prepare_inner(c, 0, 1);
      if (rc) return rc;
      rc = read_ostate(c);
      if (rc) return rc;
      beta = o.betaExact;
    }
    pf.chebBeta = beta;
    hipLaunchKernelGGL(k_recip, g1, b1, 0, c->stream, n, c->diag64.p, c->rD64.p);
    KState s{};
    s.mode = MODE_OPENFOAM;
    s.normFactor = pf.normFactor;
    s.initRes = pf.initialResidual;
    s.tol = ctl->tolerance;
    s.relTol = ctl->relTol;
    s.maxIter = ctl->maxIter;
    s.minIter = ctl->minIter;
    s.status = ST_RUNNING;
    if (order > 1) cheb_constants(s, order, beta, ctl->chebyshevRatio, false);
    // vectors (delta starts from psi) and both state buffers
    auto init = [&]() -> int {
      hipLaunchKernelGGL((k_inner_init<double>), g1, b1, 0, c->stream, n, c->r64.p,
                         (const double*)nullptr, (const double*)nullptr, w.r.p,
                         useBiCG ? w.rA0.p : nullptr, (double*)nullptr, w.p0.p, w.p1.p,
                         useBiCG ? w.AyA.p : nullptr, state_words(s), c->kst.p, c->kst2.p);
      hipLaunchKernelGGL((k_copy<double>), g1, b1, 0, c->stream, n, c->psi0.p, w.delta.p);
      CKL();
      return HIPFOAM_OK;
    };
    rc = init();
    if (rc) return rc;
    const LduOp<double> A = op64(c);
    double gpu = 0;
    const double tw = now_seconds();
    rc = iterate<double>(c, useBiCG ? KIND_BICG : KIND_PCG, A, c->rD64.p, order,
                         std::max(ctl->maxIter, ctl->minIter), &gpu, init);
    if (rc) return rc;
    pf.tInner = now_seconds() - tw;
    pf.tGpu = gpu;
    pf.nIterations = c->hk->it;
    pf.finalResidual = c->hk->res;
    pf.singular = c->hk->status == ST_SINGULAR;
    if (pf.singular && c->hk->it == 0) pf.finalResidual = pf.initialResidual;
    result = w.delta.p;
    (void)bs;
  }
  pf.converged =
      of_converged_h(pf.finalResidual, pf.initialResidual, ctl->tolerance, ctl->relTol);
  const hipMemcpyKind back =
      kind == hipMemcpyHostToDevice ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  XF(psi, result, sizeof(double) * n, back);
  if (residual) {
    // Chebyshev PCG: the residual of the last update went to the buffer of
    // parity it (k_cheb_first of iteration it wrote it)
    const double* rlast = (!useBiCG && order > 1 && (c->hk->it & 1) == 0) ? w.r1.p : w.r.p;
    const double* rr = (result == c->psi0.p) ? c->r64.p : rlast;
    XF(residual, rr, sizeof(double) * n, back);
  }
  CK(hipStreamSynchronize(c->stream));
  pf.tTotal = now_seconds() - t0;
  if (perf) *perf = pf;
  return HIPFOAM_OK;
}

int amul_impl(hipfoam_context* c, int precision, const void* x, void* y, bool device) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!x || !y) return fail(c, HIPFOAM_ERR_ARG, "NULL x/y");
  const int n = c->n, bs = stream_block(c);
  const dim3 g(stream_grid(c)), bl(bs);
  if (precision == HIPFOAM_FP32) {
    if (!c->have32) return fail(c, HIPFOAM_ERR_STATE, "no FP32 operator staged");
    const float* xd = static_cast<const float*>(x);
    float* yd = static_cast<float*>(y);
    if (!device) {
      CK(c->w32.ensure(n));
      XF(c->w32.z0.p, x, sizeof(float) * n, hipMemcpyHostToDevice);
      xd = c->w32.z0.p;
      yd = c->w32.z1.p;
    }
    LAUNCH_CLASSED(5, false, (k_spmv<float>), g, bl, stream_op(op32(c), c), xd, yd);
    CKL();
    if (!device)
      XF(y, yd, sizeof(float) * n, hipMemcpyDeviceToHost);
  } else if (precision == HIPFOAM_FP64) {
    if (!c->have64) return fail(c, HIPFOAM_ERR_STATE, "no FP64 matrix staged");
    const double* xd = static_cast<const double*>(x);
    double* yd = static_cast<double*>(y);
    if (!device) {
      CK(c->psi1.ensure(n));
      CK(c->scratch64.ensure(n));
      XF(c->psi1.p, x, sizeof(double) * n, hipMemcpyHostToDevice);
      xd = c->psi1.p;
      yd = c->scratch64.p;
    }
    LAUNCH_CLASSED(5, false, (k_spmv<double>), g, bl, stream_op(op64(c), c), xd, yd);
    CKL();
    if (!device)
      XF(y, yd, sizeof(double) * n, hipMemcpyDeviceToHost);
  } else {
    return fail(c, HIPFOAM_ERR_ARG, "precision must be HIPFOAM_FP32 or HIPFOAM_FP64");
  }
  if (!device || !c->timer.on) CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int set_matrix_impl(hipfoam_context* c, const double* diag, const double* upper,
                    const double* lower, hipMemcpyKind kind) {
  int rc = check_ready(c);
  if (rc) return rc;
  rc = ensure_vectors(c);
  if (rc) return rc;
  rc = stage_matrix(c, diag, upper, lower, kind);
  if (rc) return rc;
  rc = prepare_inner(c, 0, 0);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

// Adjoint of a solve (all pointers on the device): the transpose solve
// A^T lambda = g runs through the same refined path with upper and lower
// swapped on the same addressing (lduMatrix::Tmul semantics), then the face,
// diagonal and source gradients are formed from x and lambda.
// x, g, lambda and the gradients on the device; diag, upper and lower where
// matKind says (hipMemcpyHostToDevice: staged from host memory by the solve)
int adjoint_device(hipfoam_context* c, int solver, int precision, const double* diag,
                   const double* upper, const double* lower, const double* x, const double* g,
                   double* lambda, double* dUpper, double* dLower, double* dDiag,
                   double* dSource, const hipfoam_controls* ctl, hipfoam_performance* perf,
                   hipMemcpyKind matKind = hipMemcpyDeviceToDevice) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!x || !g || !lambda || !ctl) return fail(c, HIPFOAM_ERR_ARG, "NULL x/g/lambda/controls");
  const int n = c->n, nf = c->nf;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, 0.0, lambda);
  CKL();
  const double* upT = lower ? lower : upper;
  const double* loT = lower ? upper : nullptr;
  if (precision != HIPFOAM_FP32 && precision != HIPFOAM_FP64)
    return fail(c, HIPFOAM_ERR_ARG, "precision must be HIPFOAM_FP32 or HIPFOAM_FP64");
  rc = precision == HIPFOAM_FP64
           ? fp64_solve(c, solver, diag, upT, loT, lambda, g, ctl, perf, nullptr,
                        hipMemcpyDeviceToDevice, matKind)
           : refined_solve(c, solver, diag, upT, loT, lambda, g, ctl, perf, nullptr,
                           hipMemcpyDeviceToDevice, matKind);
  if (rc) return rc;
  if (nf > 0 && (dUpper || dLower)) {
    hipLaunchKernelGGL(k_adjoint_faces, dim3(blocks_for(nf, 256)), dim3(256), 0, c->stream, nf,
                       c->lowAddr.p, c->uppAddr.p, x, lambda, lower ? 0 : 1, dUpper,
                       lower ? dLower : nullptr);
  }
  if (n > 0 && (dDiag || dSource)) {
    hipLaunchKernelGGL(k_adjoint_cells, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, x,
                       lambda, dDiag, dSource);
  }
  CKL();
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int adjoint_host(hipfoam_context* c, int solver, int precision, const double* diag,
                 const double* upper, const double* lower, const double* x, const double* g,
                 double* lambda, double* dUpper, double* dLower, double* dDiag,
                 double* dSource, const hipfoam_controls* ctl, hipfoam_performance* perf) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!diag || !upper || !x || !g || !lambda || !ctl)
    return fail(c, HIPFOAM_ERR_ARG, "NULL argument");
  const int n = c->n, nf = c->nf;
  const hipMemcpyKind h2d = hipMemcpyHostToDevice, d2h = hipMemcpyDeviceToHost;
  CK(c->adjX.ensure(n));
  CK(c->adjG.ensure(n));
  CK(c->adjLam.ensure(n));
  CK(c->adjDD.ensure(n));
  CK(c->adjDB.ensure(n));
  CK(c->adjDU.ensure(nf));
  CK(c->adjDL.ensure(nf));
  rc = ensure_vectors(c);
  if (rc) return rc;
  XF(c->adjX.p, x, sizeof(double) * n, h2d);
  XF(c->adjG.p, g, sizeof(double) * n, h2d);
  // the coefficients go from host memory straight into the staged matrix
  rc = adjoint_device(c, solver, precision, diag, upper, lower, c->adjX.p, c->adjG.p,
                      c->adjLam.p, dUpper ? c->adjDU.p : nullptr,
                      (dLower && lower) ? c->adjDL.p : nullptr, dDiag ? c->adjDD.p : nullptr,
                      dSource ? c->adjDB.p : nullptr, ctl, perf, h2d);
  if (rc == HIPFOAM_OK) {
    XF(lambda, c->adjLam.p, sizeof(double) * n, d2h);
    if (dUpper) XF(dUpper, c->adjDU.p, sizeof(double) * nf, d2h);
    if (dLower && lower)
      XF(dLower, c->adjDL.p, sizeof(double) * nf, d2h);
    if (dDiag) XF(dDiag, c->adjDD.p, sizeof(double) * n, d2h);
    if (dSource) XF(dSource, c->adjDB.p, sizeof(double) * n, d2h);
  }
  CK(hipStreamSynchronize(c->stream));
  return rc;
}

#include "hipfoam_batch.inc"

}  // namespace

// ---------------------------------------------------------------------------
// Scratch array of hipfoam_solve_device_noalloc
// ---------------------------------------------------------------------------

// Every per-solve device buffer of an allocation-free refined solve (work
// vectors, staged FP64/FP32 operator, row-slot coefficients, reduction
// partials, Krylov and outer states) is carved from this array.  Each GPU
// the library runs on holds its own copy, for the life of the process.  In
// host code `scratch` names a host-side stand-in: the device address comes
// from hipGetSymbolAddress, once per device with that device current, and
// only that address is passed to kernels and copies.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchInitMu;
char* g_scratchDev[kMaxScratchDevices] = {};
// one allocation-free solve at a time per device in this process: the calls
// share the device's array
std::mutex g_scratchUseMu[kMaxScratchDevices];

// Called at hipfoam_create with `device` current.
hipError_t scratch_address(int device, char** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratchInitMu);
  if (!g_scratchDev[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchDev[device] = static_cast<char*>(ptr);
  }
  *out = g_scratchDev[device];
  return hipSuccess;
}

__global__ void k_zero_words(size_t nw, uint32_t* __restrict__ p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) p[i] = 0u;
}

// Points buffers of the context into the scratch array (or, with base ==
// nullptr, only counts the bytes) and puts them back on destruction.
class ScratchCarve {
 public:
  ScratchCarve(char* base, size_t cap) : base_(base), cap_(cap) {}
  ScratchCarve(const ScratchCarve&) = delete;
  ScratchCarve& operator=(const ScratchCarve&) = delete;
  ~ScratchCarve() { restore(); }

  template <typename T>
  void take(Buf<T>& b, size_t n) {
    if (n == 0) n = 1;
    const size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
    if (base_ && used_ + bytes <= cap_) {
      saved_.push_back({reinterpret_cast<void**>(&b.p), &b.cap, &b.borrowed,
                        static_cast<void*>(b.p), b.cap, b.borrowed});
      b.p = reinterpret_cast<T*>(base_ + used_);
      b.cap = n;
      b.borrowed = true;
    }
    used_ += bytes;
  }
  size_t used() const { return used_; }
  void restore() {
    for (auto it = saved_.rbegin(); it != saved_.rend(); ++it) {
      *it->pp = it->p;
      *it->pcap = it->cap;
      *it->pb = it->borrowed;
    }
    saved_.clear();
  }

 private:
  struct Saved {
    void** pp;
    size_t* pcap;
    bool* pb;
    void* p;
    size_t cap;
    bool borrowed;
  };
  char* base_;
  size_t cap_;
  size_t used_ = 0;
  std::vector<Saved> saved_;
};

// The buffers refined_solve touches on a device-array solve, at the sizes it
// asks for: ensure_vectors, stage_matrix (lower staged when given), the
// partials and the state structs.  The states come first, so that they can
// be cleared in one fill.
void carve_refined_solve(hipfoam_context* c, bool withLower, ScratchCarve& sc) {
  const size_t n = (size_t)c->n, nf = (size_t)c->nf;
  sc.take(c->kst, 1);
  sc.take(c->kst2, 1);
  sc.take(c->ost, 1);
  sc.take(c->partials, (size_t)c->pstride * 4);
  sc.take(c->partials2, (size_t)c->pstride * 4);
  Buf<float>* w32[] = {&c->w32.r,   &c->w32.r1,  &c->w32.p0,  &c->w32.p1,  &c->w32.wA,
                       &c->w32.Ap,  &c->w32.delta, &c->w32.deltaBest, &c->w32.z0,
                       &c->w32.z1,  &c->w32.w,   &c->w32.rA0, &c->w32.yA,  &c->w32.AyA,
                       &c->w32.AyA1, &c->w32.sA, &c->w32.zA,  &c->w32.tA};
  for (auto* b : w32) sc.take(*b, n);
  Buf<double>* v64[] = {&c->psi0, &c->psi1, &c->b, &c->r64, &c->u64, &c->scratch64,
                        &c->diag64, &c->rD64};
  for (auto* b : v64) sc.take(*b, n);
  sc.take(c->diag32, n);
  sc.take(c->rD32, n);
  sc.take(c->upper64, nf);
  sc.take(c->lowerL64, nf);
  sc.take(c->upper32, nf);
  sc.take(c->lowerL32, nf);
  if (withLower) {
    sc.take(c->lower64, nf);
    sc.take(c->lower32, nf);
  }
  if (c->slotOK && n > 0) {
    sc.take(c->slotCoef32, (size_t)2 * kShortRow * n);
    sc.take(c->slotCoef64, (size_t)2 * kShortRow * n);
  }
}

// Context state an allocation-free solve changes besides its carved buffers;
// put back afterwards, so the context's own staged operator stays valid.
struct ContextStateSave {
  hipfoam_context* c;
  size_t w32n;
  bool have64, sym64, have32, sym32, timerOn;
  int graphChunk;
  explicit ContextStateSave(hipfoam_context* cc)
      : c(cc), w32n(cc->w32.n), have64(cc->have64), sym64(cc->sym64), have32(cc->have32),
        sym32(cc->sym32), timerOn(cc->timer.on), graphChunk(cc->sched.graphChunk) {}
  ~ContextStateSave() {
    c->w32.n = w32n;
    c->have64 = have64;
    c->sym64 = sym64;
    c->have32 = have32;
    c->sym32 = sym32;
    c->timer.on = timerOn;
    c->sched.graphChunk = graphChunk;
  }
};

struct NoAllocFlag {
  NoAllocFlag() { t_noAlloc = true; }
  ~NoAllocFlag() { t_noAlloc = false; }
};

int solve_device_noalloc(hipfoam_context* c, int solver, const double* diag,
                         const double* upper, const double* lower, double* psi,
                         const double* source, const hipfoam_controls* ctl,
                         hipfoam_performance* perf, double* residual) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->scratchDev)
    return fail(c, HIPFOAM_ERR_DEVICE, "no scratch array address on device " +
                                           std::to_string(c->device));
  {
    ScratchCarve count(nullptr, 0);
    carve_refined_solve(c, lower != nullptr, count);
    if (count.used() > kScratchBytes)
      return fail(c, HIPFOAM_ERR_SIZE,
                  "allocation-free solve needs " + std::to_string(count.used()) +
                      " bytes of scratch for " + std::to_string(c->n) + " cells and " +
                      std::to_string(c->nf) + " faces; the array holds " +
                      std::to_string(kScratchBytes));
  }
  std::lock_guard<std::mutex> lk(g_scratchUseMu[c->device]);
  ContextStateSave keep(c);
  rc = HIPFOAM_OK;
  {
    ScratchCarve sc(c->scratchDev, kScratchBytes);
    carve_refined_solve(c, lower != nullptr, sc);
    c->w32.n = (size_t)c->n;
    c->sched.graphChunk = 0;  // graph instantiation allocates: eager launches
    c->timer.on = false;
    {
      NoAllocFlag guard;
      // the previous call's states are still in the array: clear them
      const size_t stateBytes = (reinterpret_cast<char*>(c->partials.p) -
                                 reinterpret_cast<char*>(c->kst.p));
      const size_t nw = stateBytes / sizeof(uint32_t);
      hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                         c->stream, nw, reinterpret_cast<uint32_t*>(c->kst.p));
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess)
        rc = fail(c, HIPFOAM_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
      else
        rc = refined_solve(c, solver, diag, upper, lower, psi, source, ctl, perf, residual,
                           hipMemcpyDeviceToDevice);
    }
    // nothing may still run on the array when the next call takes it
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (rc == HIPFOAM_OK && es != hipSuccess)
      rc = fail(c, HIPFOAM_ERR_DEVICE, std::string("hipStreamSynchronize: ") +
                                           hipGetErrorString(es));
  }
  return rc;
}

}  // namespace

// ===========================================================================
// extern "C" exports
// ===========================================================================

extern "C" {

void hipfoam_default_controls(hipfoam_controls* c) {
  if (!c) return;
  c->tolerance = 1e-6;
  c->relTol = 0.0;
  c->maxIter = 1000;
  c->minIter = 0;
  c->deflate = 2;
  c->chebyshevOrder = 8;
  c->chebyshevRatio = 30.0;
  c->maxRestarts = 8;
  c->iterationsPerSync = 1;
  c->verbose = 0;
}

int32_t hipfoam_create(int32_t device, hipfoam_context** out) {
  FpTrapGuard fpe;
  if (!out) return HIPFOAM_ERR_ARG;
  hipfoam_context* c = new (std::nothrow) hipfoam_context();
  *out = c;
  if (!c) return HIPFOAM_ERR_DEVICE;
  c->device = device;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) {
    c->desc = "hipfoam: no HIP device " + std::to_string(device);
    c->err = c->desc;
    return HIPFOAM_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->desc = "hipfoam: cannot initialise HIP device " + std::to_string(device);
    c->err = c->desc;
    return HIPFOAM_ERR_DEVICE;
  }
  c->numCUs = prop.multiProcessorCount;
  if (hipHostMalloc(reinterpret_cast<void**>(&c->hk), sizeof(KState)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->ho), sizeof(OState)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->hstatus), 4 * sizeof(int32_t)) != hipSuccess ||
      c->kst.ensure(1) != hipSuccess || c->kst2.ensure(1) != hipSuccess ||
      c->ost.ensure(1) != hipSuccess || c->xsync.ensure(kXcdSyncWords) != hipSuccess ||
      hipEventCreateWithFlags(&c->evPoll[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->evPoll[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&c->evT0) != hipSuccess || hipEventCreate(&c->evT1) != hipSuccess) {
    c->desc = "hipfoam: device allocation failed";
    c->err = c->desc;
    return HIPFOAM_ERR_DEVICE;
  }
  (void)hipMemset(c->kst.p, 0, sizeof(KState));
  (void)hipMemset(c->kst2.p, 0, sizeof(KState));
  (void)hipMemset(c->ost.p, 0, sizeof(OState));
  (void)hipMemset(c->xsync.p, 0, sizeof(uint32_t) * kXcdSyncWords);
  // the status mirror: coherent pinned memory the kernels store to directly
  // (without it the host copies the status after each chunk)
  if (hipHostMalloc(reinterpret_cast<void**>(&c->mirrorHost), 64, hipHostMallocCoherent) ==
          hipSuccess &&
      hipHostGetDevicePointer(reinterpret_cast<void**>(&c->mirrorDev), c->mirrorHost, 0) ==
          hipSuccess) {
    *c->mirrorHost = ST_RUNNING;
  } else {
    if (c->mirrorHost) (void)hipHostFree(c->mirrorHost);
    c->mirrorHost = nullptr;
    c->mirrorDev = nullptr;
  }
  (void)hipDeviceSynchronize();
  // the device address of this device's scratch array (allocation-free solve)
  if (scratch_address(device, &c->scratchDev) != hipSuccess) c->scratchDev = nullptr;
  char buf[512];
  snprintf(buf, sizeof(buf), "hipfoam: device %d %s(%s, %d CUs, %.0f GiB HBM)", device,
           prop.name[0] ? (std::string(prop.name) + " ").c_str() : "",
           prop.gcnArchName, prop.multiProcessorCount,
           (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
  c->desc = buf;
  c->good = true;
  return HIPFOAM_OK;
}

void hipfoam_destroy(hipfoam_context* c) {
  FpTrapGuard fpe;
  if (!c) return;
  if (c->good) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  clear_graphs(c);
  for (auto& e : c->timer.ev) (void)hipEventDestroy(e);
  Buf<int32_t>* ib[] = {&c->ownStart, &c->losortStart, &c->uppAddr,
                        &c->lowAddr, &c->losort, &c->losortLow};
  for (auto* b : ib) b->release();
  Buf<double>* db[] = {&c->diag64, &c->upper64, &c->lower64, &c->lowerL64, &c->rD64,
                       &c->psi0, &c->psi1, &c->b, &c->r64, &c->u64, &c->scratch64,
                       &c->partials, &c->adjX, &c->adjG, &c->adjLam, &c->adjDU,
                       &c->adjDL, &c->adjDD, &c->adjDB};
  for (auto* b : db) b->release();
  Buf<float>* fb[] = {&c->diag32, &c->upper32, &c->lowerL32, &c->lower32, &c->rD32};
  for (auto* b : fb) b->release();
  c->w32.release();
  c->w64.release();
  c->kst.release();
  c->ost.release();
  if (c->hk) (void)hipHostFree(c->hk);
  if (c->ho) (void)hipHostFree(c->ho);
  if (c->hstatus) (void)hipHostFree(c->hstatus);
  if (c->mirrorHost) (void)hipHostFree(c->mirrorHost);
  for (hipEvent_t e : {c->evPoll[0], c->evPoll[1], c->evT0, c->evT1})
    if (e) (void)hipEventDestroy(e);
  c->stager.reset();  // waits for its DMAs, frees the pinned chunks, joins its threads
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int32_t hipfoam_good(const hipfoam_context* c) { return c && c->good ? 1 : 0; }

const char* hipfoam_description(const hipfoam_context* c) {
  return c ? c->desc.c_str() : "hipfoam: NULL context";
}

const char* hipfoam_last_error(const hipfoam_context* c) {
  return c ? c->err.c_str() : "NULL context";
}

// Fingerprint of a mesh's addressing: the sizes and 1024 evenly spaced
// entries (the first and last included) of each addressing array, mixed
// with splitmix64.  A few thousand loads, so it is taken on every call.
// 64-bit hash of a byte range: four independent multiply-rotate lanes over
// 8-byte words (the accumulate step of the published xxHash64 round, written
// here from its description), a tail loop, and the murmur3 64-bit finaliser.
uint64_t hash_bytes(const unsigned char* p, size_t len, uint64_t seed) {
  constexpr uint64_t P1 = 0x9e3779b185ebca87ull, P2 = 0xc2b2ae3d27d4eb4full;
  auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
  uint64_t a0 = seed + P1, a1 = seed ^ P2, a2 = seed * P1 + 1, a3 = seed - P1;
  size_t i = 0;
  for (; i + 32 <= len; i += 32) {
    uint64_t w[4];
    std::memcpy(w, p + i, 32);
    a0 = rotl(a0 + w[0] * P2, 31) * P1;
    a1 = rotl(a1 + w[1] * P2, 31) * P1;
    a2 = rotl(a2 + w[2] * P2, 31) * P1;
    a3 = rotl(a3 + w[3] * P2, 31) * P1;
  }
  uint64_t h = rotl(a0, 1) + rotl(a1, 7) + rotl(a2, 12) + rotl(a3, 18) + len;
  for (; i < len; ++i) h = (h ^ p[i]) * P1;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  return h ^ (h >> 33);
}

// Hash of the whole addressing content (every entry of the five arrays and
// the sizes), so any rewrite of it is detected, up to 64-bit collisions.
// The arrays are cut into 1 MiB chunks hashed independently (on the
// context's host threads when the addressing is large) and combined in
// order: the result does not depend on the thread count.  Measured on the
// MI355X host at 128^3 (92 MB of addressing): see hipfoam_set_topology_check.
uint64_t addressing_hash(hipfoam_context* c, int32_t n, int32_t nf, const int32_t* low,
                         const int32_t* upp, const int32_t* ownStart, const int32_t* losort,
                         const int32_t* losortStart) {
  constexpr size_t kChunk = size_t(1) << 20;
  constexpr size_t kThreadedMin = size_t(4) << 20;
  const size_t fb = sizeof(int32_t) * (size_t)std::max(nf, 0);
  const size_t cb = sizeof(int32_t) * ((size_t)std::max(n, 0) + 1);
  const unsigned char* arr[5] = {reinterpret_cast<const unsigned char*>(low),
                                 reinterpret_cast<const unsigned char*>(upp),
                                 reinterpret_cast<const unsigned char*>(losort),
                                 reinterpret_cast<const unsigned char*>(ownStart),
                                 reinterpret_cast<const unsigned char*>(losortStart)};
  const size_t bytes[5] = {fb, fb, fb, cb, cb};
  std::vector<std::pair<int, size_t>> jobs;
  size_t total = 0;
  for (int k = 0; k < 5; ++k) {
    if (!arr[k]) continue;
    total += bytes[k];
    for (size_t o = 0; o < bytes[k]; o += kChunk) jobs.push_back({k, o});
  }
  std::vector<uint64_t> part(jobs.size());
  auto work = [&](int k, int parts) {
    for (size_t j = (size_t)k; j < jobs.size(); j += (size_t)parts) {
      const int a = jobs[j].first;
      const size_t o = jobs[j].second, len = std::min(kChunk, bytes[a] - o);
      part[j] = hash_bytes(arr[a] + o, len, ((uint64_t)a << 40) ^ o);
    }
  };
  if (total >= kThreadedMin)
    host_pool(c).run(work);
  else
    work(0, 1);
  auto mix = [](uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  };
  uint64_t h = mix(((uint64_t)(uint32_t)n << 32) | (uint32_t)nf);
  for (size_t j = 0; j < part.size(); ++j) h = mix(h ^ part[j]) + j;
  return h;
}

int32_t hipfoam_set_topology(hipfoam_context* c, int32_t nCells, int32_t nFaces,
                             const int32_t* lowAddr, const int32_t* uppAddr,
                             const int32_t* ownStart, const int32_t* losort,
                             const int32_t* losortStart) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  if (!c->good) return fail(c, HIPFOAM_ERR_DEVICE, "no usable device: " + c->desc);
  if (nCells < 0 || nFaces < 0) return fail(c, HIPFOAM_ERR_ARG, "negative size");
  if (nFaces > 0 && (!lowAddr || !uppAddr || !losort))
    return fail(c, HIPFOAM_ERR_ARG, "NULL addressing");
  if (!ownStart || !losortStart) return fail(c, HIPFOAM_ERR_ARG, "NULL start arrays");
  CK(hipSetDevice(c->device));
  // no-op when unchanged.  HIPFOAM_TOPO_POINTERS keys on the sizes and the
  // addressing pointers, as the reference does (metalFoamBackend.C:469-477);
  // HIPFOAM_TOPO_CONTENT (the default) on the sizes and a hash of every
  // addressing entry, so addressing rewritten in place is uploaded again and
  // the same mesh at new addresses is not
  uint64_t hash = 0;
  if (c->topoCheck == HIPFOAM_TOPO_CONTENT) {
    hash = addressing_hash(c, nCells, nFaces, lowAddr, uppAddr, ownStart, losort, losortStart);
    if (c->ownStart.p && c->topoHashValid && c->n == nCells && c->nf == nFaces &&
        c->topoHash == hash) {
      c->lastLow = lowAddr;
      c->lastUpp = uppAddr;
      return HIPFOAM_OK;
    }
  } else if (c->ownStart.p && c->n == nCells && c->nf == nFaces && c->lastLow == lowAddr &&
             c->lastUpp == uppAddr) {
    return HIPFOAM_OK;
  }
  CK(hipStreamSynchronize(c->stream));
  clear_graphs(c);
  c->gen += 1;
  c->n = nCells;
  c->nf = nFaces;
  c->maxOwn = c->maxNbr = 0;
  for (int32_t i = 0; i < nCells; ++i) {
    c->maxOwn = std::max(c->maxOwn, ownStart[i + 1] - ownStart[i]);
    c->maxNbr = std::max(c->maxNbr, losortStart[i + 1] - losortStart[i]);
  }
  SlotLayout lay;
  c->slotOK = build_slots(nCells, ownStart, uppAddr, losortStart, losort, lowAddr, lay);
  c->slotPatterned = c->slotOK ? lay.patternedRows : 0;
  c->slotFaceRows = c->slotOK ? lay.faceRows : 0;
  c->strideMajor = c->slotOK ? lay.strideMajor : 0;
  c->strideMinor = c->slotOK ? lay.strideMinor : 0;
  // the cache key is set once the upload has succeeded
  c->lastLow = c->lastUpp = nullptr;
  c->topoHashValid = false;
  c->have32 = c->have64 = false;
  const size_t nf = (size_t)nFaces, n1 = (size_t)nCells + 1;
  CK(c->ownStart.ensure(n1));
  CK(c->losortStart.ensure(n1));
  CK(c->uppAddr.ensure(nf));
  CK(c->lowAddr.ensure(nf));
  CK(c->losort.ensure(nf));
  CK(c->losortLow.ensure(nf));
  CK(hipMemcpyAsync(c->ownStart.p, ownStart, sizeof(int32_t) * n1, hipMemcpyHostToDevice,
                    c->stream));
  CK(hipMemcpyAsync(c->losortStart.p, losortStart, sizeof(int32_t) * n1, hipMemcpyHostToDevice,
                    c->stream));
  if (nFaces > 0) {
    CK(hipMemcpyAsync(c->uppAddr.p, uppAddr, sizeof(int32_t) * nf, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync(c->lowAddr.p, lowAddr, sizeof(int32_t) * nf, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync(c->losort.p, losort, sizeof(int32_t) * nf, hipMemcpyHostToDevice,
                      c->stream));
    hipLaunchKernelGGL(k_gather_addr, dim3(blocks_for(nFaces, 256)), dim3(256), 0, c->stream,
                       nFaces, c->losort.p, c->lowAddr.p, c->losortLow.p);
    CKL();
  }
  if (c->slotOK && nCells > 0) {
    CK(c->slotOff.ensure(lay.words.size()));
    CK(c->slotTile.ensure(lay.tile.size()));
    CK(c->slotRow.ensure(lay.row.size()));
    CK(hipMemcpyAsync(c->slotTile.p, lay.tile.data(), sizeof(int32_t) * lay.tile.size(),
                      hipMemcpyHostToDevice, c->stream));
    CK(hipMemcpyAsync(c->slotOff.p, lay.words.data(), sizeof(uint32_t) * lay.words.size(),
                      hipMemcpyHostToDevice, c->stream));
    CK(hipMemcpyAsync(c->slotRow.p, lay.row.data(), lay.row.size(), hipMemcpyHostToDevice,
                      c->stream));
    // the host layout must outlive the asynchronous copies
    CK(hipStreamSynchronize(c->stream));
  }
  c->pstride = std::max(1, blocks_for(nCells, 64));
  CK(c->partials.ensure((size_t)c->pstride * 4));
  CK(c->partials2.ensure((size_t)c->pstride * 4));
  c->w32.release();
  c->w64.release();
  int rc = ensure_vectors(c);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->stream));
  c->lastLow = lowAddr;
  c->lastUpp = uppAddr;
  c->topoHash = hash;
  c->topoHashValid = c->topoCheck == HIPFOAM_TOPO_CONTENT;
  return HIPFOAM_OK;
}

int32_t hipfoam_set_topology_check(hipfoam_context* c, int32_t mode) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  if (mode != HIPFOAM_TOPO_POINTERS && mode != HIPFOAM_TOPO_CONTENT)
    return fail(c, HIPFOAM_ERR_ARG, "topology check must be HIPFOAM_TOPO_POINTERS (0) or "
                                    "HIPFOAM_TOPO_CONTENT (1)");
  c->topoCheck = mode;
  return HIPFOAM_OK;
}

int32_t hipfoam_set_coeffs(hipfoam_context* c, const double* diag, const double* upper,
                           const double* lower, double* seconds) {
  FpTrapGuard fpe;
  const double t0 = now_seconds();
  int rc = set_matrix_impl(c, diag, upper, lower, hipMemcpyHostToDevice);
  if (seconds) *seconds = now_seconds() - t0;
  return rc;
}

int32_t hipfoam_set_matrix(hipfoam_context* c, const double* diag, const double* upper,
                           const double* lower) {
  FpTrapGuard fpe;
  return set_matrix_impl(c, diag, upper, lower, hipMemcpyHostToDevice);
}

int32_t hipfoam_set_matrix_device(hipfoam_context* c, const double* diag, const double* upper,
                                  const double* lower) {
  FpTrapGuard fpe;
  return set_matrix_impl(c, diag, upper, lower, hipMemcpyDeviceToDevice);
}

static int32_t inner_entry(hipfoam_context* c, Kind kind, const double* r, double target,
                           int32_t maxIter, int32_t ips, int32_t order, double beta,
                           double ratio, hipfoam_stats* stats) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!r) return fail(c, HIPFOAM_ERR_ARG, "NULL r");
  if (!c->have32) return fail(c, HIPFOAM_ERR_STATE, "set_coeffs not called");
  if (order < 1 || order > kMaxCheb + 1)
    return fail(c, HIPFOAM_ERR_ARG, "chebOrder must be in [1, 33]");
  XF(c->r64.p, r, sizeof(double) * c->n, hipMemcpyHostToDevice);
  rc = inner_solve(c, kind, c->r64.p, nullptr, nullptr, target, maxIter, ips, order, beta, ratio,
                   stats);
  if (rc) return rc;
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_pcg(hipfoam_context* c, const double* r, double targetResSum, int32_t maxIter,
                    int32_t iterationsPerSync, int32_t chebOrder, double chebBeta,
                    double chebRatio, hipfoam_stats* stats) {
  FpTrapGuard fpe;
  return inner_entry(c, KIND_PCG, r, targetResSum, maxIter, iterationsPerSync, chebOrder,
                     chebBeta, chebRatio, stats);
}

int32_t hipfoam_bicgstab(hipfoam_context* c, const double* r, double targetResSum,
                         int32_t maxIter, hipfoam_stats* stats) {
  FpTrapGuard fpe;
  return inner_entry(c, KIND_BICG, r, targetResSum, maxIter, 1, 1, 2.0, 30.0, stats);
}

int32_t hipfoam_delta(hipfoam_context* c, float* out) {
  FpTrapGuard fpe;
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out) return fail(c, HIPFOAM_ERR_ARG, "NULL out");
  if (!c->w32.delta.p) return fail(c, HIPFOAM_ERR_STATE, "no inner solve yet");
  XF(out, c->w32.delta.p, sizeof(float) * c->n, hipMemcpyDeviceToHost);
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_solve(hipfoam_context* c, int32_t solver, const double* diag,
                      const double* upper, const double* lower, double* psi,
                      const double* source, const hipfoam_controls* controls,
                      hipfoam_performance* perf, double* residual) {
  FpTrapGuard fpe;
  return refined_solve(c, solver, diag, upper, lower, psi, source, controls, perf, residual,
                       hipMemcpyHostToDevice);
}

int32_t hipfoam_solve_device(hipfoam_context* c, int32_t solver, const double* diag,
                             const double* upper, const double* lower, double* psi,
                             const double* source, const hipfoam_controls* controls,
                             hipfoam_performance* perf, double* residual) {
  FpTrapGuard fpe;
  return refined_solve(c, solver, diag, upper, lower, psi, source, controls, perf, residual,
                       hipMemcpyDeviceToDevice);
}

int32_t hipfoam_solve_device_noalloc(hipfoam_context* c, int32_t solver, const double* diag,
                                     const double* upper, const double* lower, double* psi,
                                     const double* source, const hipfoam_controls* controls,
                                     hipfoam_performance* perf, double* residual) {
  FpTrapGuard fpe;
  return solve_device_noalloc(c, solver, diag, upper, lower, psi, source, controls, perf,
                              residual);
}

int32_t hipfoam_solve_fp64(hipfoam_context* c, int32_t solver, const double* diag,
                           const double* upper, const double* lower, double* psi,
                           const double* source, const hipfoam_controls* controls,
                           hipfoam_performance* perf, double* residual) {
  FpTrapGuard fpe;
  return fp64_solve(c, solver, diag, upper, lower, psi, source, controls, perf, residual,
                    hipMemcpyHostToDevice);
}

int32_t hipfoam_solve_fp64_device(hipfoam_context* c, int32_t solver, const double* diag,
                                  const double* upper, const double* lower, double* psi,
                                  const double* source, const hipfoam_controls* controls,
                                  hipfoam_performance* perf, double* residual) {
  FpTrapGuard fpe;
  return fp64_solve(c, solver, diag, upper, lower, psi, source, controls, perf, residual,
                    hipMemcpyDeviceToDevice);
}

int32_t hipfoam_adjoint(hipfoam_context* c, int32_t solver, int32_t precision,
                        const double* diag, const double* upper, const double* lower,
                        const double* x, const double* g, double* lambda, double* dUpper,
                        double* dLower, double* dDiag, double* dSource,
                        const hipfoam_controls* controls, hipfoam_performance* perf) {
  FpTrapGuard fpe;
  return adjoint_host(c, solver, precision, diag, upper, lower, x, g, lambda, dUpper, dLower,
                      dDiag, dSource, controls, perf);
}

int32_t hipfoam_adjoint_device(hipfoam_context* c, int32_t solver, int32_t precision,
                               const double* diag, const double* upper, const double* lower,
                               const double* x, const double* g, double* lambda,
                               double* dUpper, double* dLower, double* dDiag, double* dSource,
                               const hipfoam_controls* controls, hipfoam_performance* perf) {
  FpTrapGuard fpe;
  return adjoint_device(c, solver, precision, diag, upper, lower, x, g, lambda, dUpper, dLower,
                        dDiag, dSource, controls, perf);
}

int32_t hipfoam_batch_create(hipfoam_context* c, int32_t solver, int32_t nSystems,
                             const hipfoam_system* systems, int32_t nRhs, hipfoam_batch** out) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  return batch_create_impl(c, solver, nSystems, systems, nRhs, out);
}

int32_t hipfoam_batch_set_rhs(hipfoam_batch* bt, int32_t system, const double* source,
                              const double* psi) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  hipfoam_context* c = bt->c;
  if (system < 0 || system >= bt->nsys || !source)
    return fail(c, HIPFOAM_ERR_ARG, "bad system index or NULL source");
  CK(hipSetDevice(c->device));
  const int n = bt->sysRow0[system + 1] - bt->sysRow0[system];
  double* psiCur = bt->psiIs0 ? bt->psi0.p : bt->psi1.p;
  for (int r = 0; r < bt->nrhs; ++r) {
    const long off = (long)r * bt->N + bt->sysRow0[system];
    CK(hipMemcpyAsync(bt->b.p + off, source + (long)r * n, sizeof(double) * n,
                      hipMemcpyHostToDevice, c->stream));
    if (psi)
      CK(hipMemcpyAsync(psiCur + off, psi + (long)r * n, sizeof(double) * n,
                        hipMemcpyHostToDevice, c->stream));
    else
      hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, 0.0,
                         psiCur + off);
  }
  CKL();
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_batch_reset_psi(hipfoam_batch* bt) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  hipfoam_context* c = bt->c;
  CK(hipSetDevice(c->device));
  const long NR = (long)bt->N * bt->nrhs;
  double* psiCur = bt->psiIs0 ? bt->psi0.p : bt->psi1.p;
  if (NR > 0)
    hipLaunchKernelGGL(k_fill, dim3(blocks_for((int)NR, 256)), dim3(256), 0, c->stream, (int)NR,
                       0.0, psiCur);
  CKL();
  return HIPFOAM_OK;
}

int32_t hipfoam_batch_solve(hipfoam_batch* bt, const hipfoam_controls* controls,
                            hipfoam_performance* perf) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  CK_C(bt->c, hipSetDevice(bt->c->device));
  return batch_solve_impl(bt, controls, perf);
}

int32_t hipfoam_batch_set_rhs_group(hipfoam_batch* bt, int32_t group) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  if (group < 1 || group > kRhsGroup)
    return fail(bt->c, HIPFOAM_ERR_ARG,
                "rhs group must be 1.." + std::to_string(kRhsGroup));
  if (group != bt->group) {
    bt->group = group;
    bclear_graphs(bt);
  }
  return HIPFOAM_OK;
}

int32_t hipfoam_batch_autotune(hipfoam_batch* bt, hipfoam_schedule* best, int32_t* group) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  CK_C(bt->c, hipSetDevice(bt->c->device));
  return batch_autotune_impl(bt, best, group);
}

int32_t hipfoam_batch_get_psi(hipfoam_batch* bt, int32_t system, double* psi) {
  FpTrapGuard fpe;
  if (!bt) return HIPFOAM_ERR_ARG;
  hipfoam_context* c = bt->c;
  if (system < 0 || system >= bt->nsys || !psi)
    return fail(c, HIPFOAM_ERR_ARG, "bad system index or NULL psi");
  CK(hipSetDevice(c->device));
  const int n = bt->sysRow0[system + 1] - bt->sysRow0[system];
  const double* psiCur = bt->psiIs0 ? bt->psi0.p : bt->psi1.p;
  for (int r = 0; r < bt->nrhs; ++r)
    XF(psi + (long)r * n, psiCur + (long)r * bt->N + bt->sysRow0[system],
       sizeof(double) * n, hipMemcpyDeviceToHost);
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

void hipfoam_batch_destroy(hipfoam_batch* bt) {
  FpTrapGuard fpe;
  batch_destroy_impl(bt);
}

int32_t hipfoam_amul(hipfoam_context* c, int32_t precision, const void* x, void* y) {
  FpTrapGuard fpe;
  return amul_impl(c, precision, x, y, false);
}

int32_t hipfoam_amul_device(hipfoam_context* c, int32_t precision, const void* x, void* y) {
  FpTrapGuard fpe;
  return amul_impl(c, precision, x, y, true);
}

int32_t hipfoam_device_alloc(hipfoam_context* c, uint64_t bytes, void** ptr) {
  FpTrapGuard fpe;
  if (!c || !ptr) return HIPFOAM_ERR_ARG;
  if (!c->good) return fail(c, HIPFOAM_ERR_DEVICE, "no usable device");
  CK(hipSetDevice(c->device));
  CK(hipMalloc(ptr, bytes ? bytes : 1));
  return HIPFOAM_OK;
}

int32_t hipfoam_device_free(hipfoam_context* c, void* ptr) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  if (!ptr) return HIPFOAM_OK;
  CK(hipSetDevice(c->device));
  CK(hipStreamSynchronize(c->stream));
  CK(hipFree(ptr));
  return HIPFOAM_OK;
}

int32_t hipfoam_copy_to_device(hipfoam_context* c, void* dst, const void* src, uint64_t bytes) {
  FpTrapGuard fpe;
  if (!c || (!dst && bytes) || (!src && bytes)) return HIPFOAM_ERR_ARG;
  CK(hipSetDevice(c->device));
  XF(dst, src, bytes, hipMemcpyHostToDevice);
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_copy_to_host(hipfoam_context* c, void* dst, const void* src, uint64_t bytes) {
  FpTrapGuard fpe;
  if (!c || (!dst && bytes) || (!src && bytes)) return HIPFOAM_ERR_ARG;
  CK(hipSetDevice(c->device));
  XF(dst, src, bytes, hipMemcpyDeviceToHost);
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_copy_on_device(hipfoam_context* c, void* dst, const void* src,
                               uint64_t bytes) {
  FpTrapGuard fpe;
  if (!c || (!dst && bytes) || (!src && bytes)) return HIPFOAM_ERR_ARG;
  CK(hipSetDevice(c->device));
  CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_synchronize(hipfoam_context* c) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  if (!c->good) return HIPFOAM_ERR_DEVICE;
  CK(hipStreamSynchronize(c->stream));
  return HIPFOAM_OK;
}

int32_t hipfoam_get_schedule(const hipfoam_context* c, hipfoam_schedule* s) {
  if (!c || !s) return HIPFOAM_ERR_ARG;
  *s = c->sched;
  return HIPFOAM_OK;
}

int32_t hipfoam_active_schedule(const hipfoam_context* c, int32_t* spmvSchedule) {
  if (!c || !spmvSchedule) return HIPFOAM_ERR_ARG;
  const int s = c->sched.spmvSchedule;
  *spmvSchedule = (s == 2 && !c->slotOK) ? 0 : s;
  return HIPFOAM_OK;
}

int32_t hipfoam_slot_layout(const hipfoam_context* c, int64_t* patternedRows,
                            int32_t* symmetricPlanes) {
  if (!c || !patternedRows || !symmetricPlanes) return HIPFOAM_ERR_ARG;
  *patternedRows = c->slotOK ? c->slotPatterned : 0;
  *symmetricPlanes = c->slotOK ? kShortRow : 0;
  return HIPFOAM_OK;
}

int32_t hipfoam_slot_tiles(const hipfoam_context* c, int64_t* patternedRows,
                           int64_t* explicitRows, int64_t* faceRows) {
  if (!c) return HIPFOAM_ERR_ARG;
  const int64_t pat = c->slotOK ? c->slotPatterned : 0;
  const int64_t fac = c->slotOK ? c->slotFaceRows : (int64_t)c->n;
  if (patternedRows) *patternedRows = pat;
  if (faceRows) *faceRows = fac;
  if (explicitRows) *explicitRows = (int64_t)c->n - pat - fac;
  return HIPFOAM_OK;
}

int32_t hipfoam_set_schedule(hipfoam_context* c, const hipfoam_schedule* s) {
  FpTrapGuard fpe;
  if (!c || !s) return HIPFOAM_ERR_ARG;
  if (s->spmvSchedule < 0 || s->spmvSchedule > 2)
    return fail(c, HIPFOAM_ERR_ARG, "spmvSchedule must be 0, 1 or 2");
  if (s->blockSize != 0 && (s->blockSize < 64 || s->blockSize > 1024 || s->blockSize % 64))
    return fail(c, HIPFOAM_ERR_ARG, "blockSize must be 0 (auto) or a multiple of 64 in [64, 1024]");
  if (s->graphChunk < -1 || s->graphChunk > 4096 || (s->graphChunk > 0 && (s->graphChunk & 1)))
    return fail(c, HIPFOAM_ERR_ARG, "graphChunk must be -1 (auto), 0 (eager) or even in [2, 4096]");
  if (s->gridLimit < 0) return fail(c, HIPFOAM_ERR_ARG, "gridLimit must be >= 0");
  if (s->rowForm < 0 || s->rowForm > 1) return fail(c, HIPFOAM_ERR_ARG, "rowForm must be 0 or 1");
  if (s->persistent < -1 || s->persistent > 1)
    return fail(c, HIPFOAM_ERR_ARG, "persistent must be -1 (auto), 0 (off) or 1 (on)");
  if (s->localLoop < -1 || s->localLoop > 0)
    return fail(c, HIPFOAM_ERR_ARG, "localLoop must be -1 (auto) or 0 (off)");
  if (c->good) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  if (s->xcdRemap < 0 || s->xcdRemap > 2)
    return fail(c, HIPFOAM_ERR_ARG, "xcdRemap must be 0, 1 or 2");
  c->sched = *s;
  clear_graphs(c);
  return HIPFOAM_OK;
}

int32_t hipfoam_persistent_stats(const hipfoam_context* c, int64_t* launches,
                                 int64_t* abandoned, int64_t* lastAbort) {
  if (!c) return HIPFOAM_ERR_ARG;
  if (launches) *launches = c->persistLaunches;
  if (abandoned) *abandoned = c->persistAbandoned;
  if (lastAbort) *lastAbort = c->persistLastAbort;
  return HIPFOAM_OK;
}

int32_t hipfoam_simulate_abandon(hipfoam_context* c, int32_t count) {
  if (!c || count < 0) return HIPFOAM_ERR_ARG;
  c->simAbandon = count;
  return HIPFOAM_OK;
}

int32_t hipfoam_local_stats(const hipfoam_context* c, int64_t* launches) {
  if (!c || !launches) return HIPFOAM_ERR_ARG;
  *launches = c->localLaunches;
  return HIPFOAM_OK;
}

int32_t hipfoam_autotune(hipfoam_context* c, hipfoam_schedule* best) {
  FpTrapGuard fpe;
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->have32) return fail(c, HIPFOAM_ERR_STATE, "stage a matrix first");
  const int n = c->n;
  CK(c->w32.ensure(n));
  CK(hipMemsetAsync(c->w32.z0.p, 0, sizeof(float) * n, c->stream));
  const hipfoam_schedule keep = c->sched;
  hipfoam_schedule bestS = keep;
  double bestT = 1e300;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int reps = 20;
  const int cus = std::max(c->numCUs, 1);
  for (int form = 0; form <= 1; ++form)
  for (int sch = 0; sch <= (c->slotOK ? 2 : 1); ++sch)
    for (int bs : {0, 128, 256, 512, 1024})
      for (int rm = 0; rm <= 1; ++rm)
        for (int gl : {0, 4 * cus, 16 * cus, 1 << 30}) {
        c->sched.rowForm = form;
        c->sched.spmvSchedule = sch;
        c->sched.blockSize = bs;
        c->sched.xcdRemap = rm;
        c->sched.gridLimit = gl;
        const LduOp<float> A = op32(c);
        const dim3 g(row_grid(c)), bl(row_block(c));
        auto spmv = [&]() {
          hipLaunchKernelGGL((k_spmv<float>), g, bl, 0, c->stream, A, c->w32.z0.p, c->w32.z1.p);
        };
        spmv();
        CK(hipEventRecord(e0, c->stream));
        for (int k = 0; k < reps; ++k) spmv();
        CK(hipEventRecord(e1, c->stream));
        CK(hipEventSynchronize(e1));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < bestT) {
          bestT = ms;
          bestS = c->sched;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  bestS.graphChunk = keep.graphChunk;
  c->sched = bestS;
  clear_graphs(c);
  if (best) *best = bestS;
  return HIPFOAM_OK;
}

int32_t hipfoam_set_kernel_timing(hipfoam_context* c, int32_t enable, int32_t kclass) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  if (!c->good) return fail(c, HIPFOAM_ERR_DEVICE, "no usable device");
  CK(hipSetDevice(c->device));
  KernelTimer& t = c->timer;
  if (t.ev.empty()) {
    t.ev.resize(4096);
    t.count.assign(2048, 0);
    for (auto& e : t.ev) CK(hipEventCreate(&e));
  }
  t.on = enable != 0;
  t.kclass = kclass;
  t.used = 0;
  t.launches = 0;
  t.seconds = 0;
  return HIPFOAM_OK;
}

int32_t hipfoam_kernel_timing(hipfoam_context* c, int64_t* launches, double* seconds) {
  FpTrapGuard fpe;
  if (!c) return HIPFOAM_ERR_ARG;
  int rc = timer_flush(c);
  if (rc) return rc;
  if (launches) *launches = c->timer.launches;
  if (seconds) *seconds = c->timer.seconds;
  return HIPFOAM_OK;
}

}  // extern "C"
