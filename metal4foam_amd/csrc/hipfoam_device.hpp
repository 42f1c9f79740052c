// hipfoam_device.hpp -- gfx950 device code of the LDU Krylov path.
//
// Matrix layout in HBM (rows = cells, faces = internal faces):
//   ownStart[n+1], uppAddr[F]            owner faces of row i:  upper[f] * x[uppAddr[f]]
//   losortStart[n+1], losortLow[F]       neighbour faces of row i, in losort order:
//                                        lowerL[k] * x[losortLow[k]], with
//                                        losortLow[k] = lowAddr[losort[k]] (built once per mesh)
//                                        lowerL[k]    = lower[losort[k]]   (staged per solve)
//   diag[n]
// Row i of A*x is summed in the reference kernel's order: diag first, owner
// faces ascending, then losort faces ascending (metalFoamKernels.metal:87-104).
// The "direct" schedule reads lower[losort[k]] and lowAddr[losort[k]]
// through losort instead of the staged copies.
//
// Launch shape: one thread per row, 64-wide wavefronts, blockDim rows per
// tile.  A grid of G workgroups covers the ceil(n/blockDim) tiles; logical
// workgroup b owns the contiguous tile range [b*T, (b+1)*T), and the logical
// index is remapped so that each XCD (hardware round-robin over 8) owns one
// contiguous row range, keeping neighbour gathers inside its L2.
//
// Reductions: per-thread values widened to double, wavefront butterfly
// (__shfl_xor), LDS across the block, one double per workgroup into
// `partials`; a one-workgroup finalise kernel folds them in a fixed order --
// so results are deterministic -- and updates the Krylov scalars on the
// device.  (Electing the last workgroup with device-scope atomics instead
// costs ~13 ns per arriving workgroup, serialised: 27 us at 2048 workgroups,
// against ~4.5 us for the extra kernel.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hipfoam_slots.hpp"

namespace hipfoam {

constexpr int kMaxCheb = 32;
constexpr double kVSmall = 1.0e-300;  // OpenFOAM VSMALL (double)
constexpr double kSmall = 1.0e-20;    // solverPerformance small_

enum : int32_t {
  ST_RUNNING = 0,
  ST_CONVERGED = 1,
  ST_SINGULAR = 2,
  ST_STAGNATED = 3,
  ST_MAXITER = 4,
  ST_EARLY = 5  // BiCGStab |sA| exit pending its delta update
};

enum : int32_t {
  MODE_INNER = 0,    // reference backend semantics (FP32 correction solve)
  MODE_OPENFOAM = 1  // OpenFOAM PCG/PBiCGStab semantics (plain solve)
};

// Device-resident Krylov state: the reference's scalars buffer
// (metalFoamKernels.metal:20-32) widened with the host loop's bookkeeping
// (metalFoamBackend.C:599-715) so the whole loop runs without host syncs.
struct alignas(128) KState {
  double rho, rhoOld, wApA, alpha, beta, res;         // PCG (rho = wA.rA)
  double rA0rA, rA0rAold, rA0AyA, omega, tAtA, tAsA;  // BiCGStab
  double bestRes, target;                             // inner mode
  double normFactor, initRes, tol, relTol;            // OpenFOAM mode
  double chebC1, chebC2[kMaxCheb], chebC3[kMaxCheb];
  int32_t it, maxIter, minIter, ips, batchSize, batchLeft, noProgress, status;
  int32_t copyBest, mode;
  // PCG with Chebyshev: alphaDue = a p.Ap reduction follows the last beta;
  // pend = delta and r await the update with the last alpha (applied by the
  // next iteration's first Chebyshev step, k_cheb_first)
  int32_t pend, alphaDue;
};

// A Krylov state as a kernel argument: its words, without KState's 128-byte
// alignment.
struct KStateWords {
  uint32_t w[sizeof(KState) / sizeof(uint32_t)];
};
static_assert(sizeof(KState) % sizeof(uint32_t) == 0, "KState words");

// Outer (FP64 refinement) scalars.
struct alignas(128) OState {
  double sumPsi, xbar, nPinned, uu;  // rowsum pass
  double normFactor, sumMagR, sumR, initRes;
  double rmean;  // mean of the current residual
  double sumU;   // sum of u = A 1
  double rproj;  // sum r / sum u: weight of u in the pinned-cell projection
  double dmean;  // mean of the inner correction
  double ur, c, finalRes;
  double bail, betaDefl, betaExact;
  int32_t n, deflate, pad0, pad1;
};

template <typename T>
struct LduOp {
  const int32_t* __restrict__ ownStart;
  const int32_t* __restrict__ losortStart;
  const int32_t* __restrict__ uppAddr;
  const int32_t* __restrict__ nbrAddr;  // losortLow (staged) or lowAddr (direct)
  const int32_t* __restrict__ losort;   // direct schedule only
  const T* __restrict__ diag;
  const T* __restrict__ upper;
  const T* __restrict__ lowerL;  // staged: lower[losort[k]]; direct: lower[f]
  // Row-slot layout (slots != 0): row i's owner faces in slots 0..2 and its
  // neighbour faces in slots 3..5, in the order the face loops visit them.
  // Each kSlotTile-row tile has a descriptor slotTile[t*kSlotDesc + ...]: one
  // column offset base per slot (words 0..5) and flags (word 6).
  //  - Patterned tile (flags & kSlotPatterned): every face of the tile sits
  //    exactly at its slot's base, so the column of slot q is i + base[q];
  //    slotRow[i] holds the row's presence mask (bit q: slot q holds a face).
  //  - Explicit tile: the column of slot q is i + base[q] + delta, the int16
  //    deltas of slots 2p and 2p+1 packed in slotOff[p*n + i], kSlotEmpty
  //    marking an empty slot.
  // slotCoef[q*n + i] is the coefficient of slot q.  A symmetric operator
  // (slotSym) stores the owner slots only: the coefficient of neighbour slot
  // q is the owner-side coefficient of the same face, slotCoef[s*n + col],
  // where s is the face's owner slot in row col: bits 2(q-3)..2(q-3)+1 of
  // flags >> 8 in a patterned tile, of slotRow[i] in an explicit tile.
  const T* __restrict__ slotCoef;
  const uint32_t* __restrict__ slotOff;
  const int32_t* __restrict__ slotTile;
  const uint8_t* __restrict__ slotRow;
  int32_t slotSym;
  int32_t n;
  int32_t direct;
  int32_t xcdRemap;
  int32_t shortRows;  // every row has <= kShortRow owner and <= kShortRow neighbour faces
  int32_t slots;
};


// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs; remapping gives each
// XCD one contiguous range of logical indices.
__device__ __forceinline__ int logical_block(int xcdRemap) {
  const int b = blockIdx.x;
  if (!xcdRemap) return b;
  const int G = gridDim.x;
  const int xcd = b & 7, local = b >> 3;
  const int q = G >> 3, r = G & 7;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Rows [r0, r1) owned by logical workgroup lb: a contiguous run of tiles.
__device__ __forceinline__ void block_rows(int n, int lb, int& r0, int& r1) {
  const int bs = blockDim.x;
  const int ntiles = (n + bs - 1) / bs;
  const int per = (ntiles + gridDim.x - 1) / gridDim.x;
  const long t0 = (long)lb * per;
  const long a = t0 * bs, e = (t0 + per) * bs;
  r0 = (int)(a < n ? a : n);
  r1 = (int)(e < n ? e : n);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

// Element i of a vector the current launch also writes.  COH (the persistent
// loop, where other workgroups of the same launch wrote it before a barrier):
// a device-scope load, which bypasses the CU's L1 and reads the XCD's L2.
// Otherwise a plain load: a kernel boundary separates writer and reader.
template <bool COH, typename T>
__device__ __forceinline__ T vld(const T* __restrict__ p, int i) {
  if constexpr (COH)
    return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    return p[i];
}


// Sum over the 64 lanes of a wavefront, every lane receiving the total: the
// butterfly v += v[lane ^ m] for m = 32, 16, 8, 4, 2, 1, on VALU lane moves
// rather than __shfl_xor, which lowers to ds_bpermute (two LDS operations per
// step for a double).  The steps 32 and 16 use the gfx950 swaps of the two
// 32-lane halves and of 16-lane rows; 8, 4, 2 and 1 use DPP rotations within
// 16-lane rows.  After the steps for 32, 16 and 8, lanes i and i ^ m hold
// equal values for every m those steps covered, so lane (i + k) mod 16 holds
// the value of lane i ^ k, and every lane adds the same two operands as the
// xor butterfly (a swap hands half of the lanes their operands in the other
// order, which IEEE addition does not see): the sum is bit-identical to the
// xor butterfly's in every lane.  Needs all 64 lanes active: every caller
// reduces whole waves (workgroup sizes are multiples of 64).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ double wave_sum(double v) {
  {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v = __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  }
  {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    v = __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  }
  v = v + dpp_f64<0x128>(v);  // row_ror:8
  v = v + dpp_f64<0x124>(v);  // row_ror:4
  v = v + dpp_f64<0x122>(v);  // row_ror:2
  v = v + dpp_f64<0x121>(v);  // row_ror:1
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-reduce NV values (the last one as a max when LASTMAX) and write one
// value per workgroup into partials[j*pstride + lb] (plain stores: the fold
// runs in the next kernel, after the kernel boundary has made them visible).
template <int NV, bool LASTMAX = false>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials, int pstride,
                                               int lb) {
  __shared__ double sh[NV][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const bool mx = LASTMAX && j == NV - 1;
    const double t = mx ? wave_max(v[j]) : wave_sum(v[j]);
    if (lane == 0) sh[j][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const bool mx = LASTMAX && j == NV - 1;
      double t = sh[j][0];
      for (int w = 1; w < nw; ++w) t = mx ? fmax(t, sh[j][w]) : t + sh[j][w];
      partials[j * pstride + lb] = t;
    }
  }
}

// Fold partials[j*pstride + 0 .. nblk) in a fixed order (deterministic):
// always the tree of 256 threads, whatever the workgroup size (at least 256;
// the other threads only join the barrier), so a stage merged into a row
// kernel of any workgroup size folds exactly as k_finalize does.
constexpr int kFoldThreads = 256;

template <int NV, bool LASTMAX = false, bool COH = false>
__device__ __forceinline__ void fold_partials(const double* partials, int pstride, int nblk,
                                              double (&out)[NV]) {
  __shared__ double sh[NV][kFoldThreads / 64];
  double v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = (LASTMAX && j == NV - 1) ? -1.0e308 : 0.0;
  constexpr int U = 8;  // independent loads in flight per thread
  const int tid = threadIdx.x;
  if (tid < kFoldThreads) {
    for (int b0 = tid; b0 < nblk; b0 += U * kFoldThreads) {
      double t[U][NV];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = b0 + u * kFoldThreads;
#pragma unroll
        for (int j = 0; j < NV; ++j)
          t[u][j] = b < nblk ? vld<COH>(partials, j * pstride + b)
                             : ((LASTMAX && j == NV - 1) ? -1.0e308 : 0.0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < NV; ++j)
          v[j] = (LASTMAX && j == NV - 1) ? fmax(v[j], t[u][j]) : v[j] + t[u][j];
    }
    const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const bool mx = LASTMAX && j == NV - 1;
      const double t = mx ? wave_max(v[j]) : wave_sum(v[j]);
      if (lane == 0) sh[j][wid] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const bool mx = LASTMAX && j == NV - 1;
    double t = sh[j][0];
    for (int w = 1; w < kFoldThreads / 64; ++w) t = mx ? fmax(t, sh[j][w]) : t + sh[j][w];
    out[j] = t;
  }
}

// Host-visible copy of the Krylov status: pinned host memory the host polls
// between launches, written with a system-scope store by the stages that
// write the primary state.
__device__ __forceinline__ void post_status(int32_t* mirror, int32_t status) {
  if (mirror) __hip_atomic_store(mirror, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Second stage of every reduction: one workgroup folds the partials of the
// previous kernel and thread 0 applies the finaliser f (the device-side
// scalar update, reduce_scalars of metalFoamKernels.metal:298-367).  It runs
// only when the Krylov status is in runMask (status == NULL: always); either
// way thread 0 posts the status to mirror (NULL: none), so that every chunk
// of launches posts the current status.
template <int NV, bool LASTMAX, typename F>
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials, int pstride,
                                                  int nblk, const int32_t* status, int runMask,
                                                  int32_t* mirror, F f) {
  if (status && !((runMask >> *status) & 1)) {
    if (threadIdx.x == 0) post_status(mirror, *status);
    return;
  }
  double t[NV];
  fold_partials<NV, LASTMAX>(partials, pstride, nblk, t);
  if (threadIdx.x == 0) {
    f(t);
    if (status) post_status(mirror, *status);
  }
}

// A reduction's second stage merged into the next kernel: the state it
// writes and the previous kernel's partials (nblk workgroups); mirror: where
// to post the resulting status (NULL: none; set when stOut is the primary
// state the host reads).
struct Merge {
  KState* stOut;
  const double* pin;
  int nblk;
  int32_t* mirror;
};

// Every workgroup reads the Krylov state stIn and the partials M.pin, folds
// them exactly as k_finalize does (fold_partials' 256-thread tree, in a
// workgroup of at least 256 threads) and applies the finaliser f, when the state is running, to its own
// LDS copy ks; workgroup 0 writes that copy to M.stOut.  No workgroup writes
// memory another one reads (stIn, M.pin and this kernel's partials are three
// different buffers), so the workgroups need no ordering.  Returns the state
// the kernel's rows use.
template <int NV, typename F>
__device__ __forceinline__ const KState* merged_finalize(const KState* stIn, const Merge& M,
                                                         int pstride, const F& f, KState* ks) {
  constexpr int W = (int)(sizeof(KState) / sizeof(uint32_t));
  static_assert(sizeof(KState) % sizeof(uint32_t) == 0 && W <= 256, "KState word copy");
  const int k = threadIdx.x;
  uint32_t* lds = reinterpret_cast<uint32_t*>(ks);
  const uint32_t word = k < W ? reinterpret_cast<const uint32_t*>(stIn)[k] : 0u;
  double t[NV];
  fold_partials<NV, false>(M.pin, pstride, M.nblk, t);
  if (k < W) lds[k] = word;
  __syncthreads();
  if (k == 0 && ks->status == ST_RUNNING) {
    F g = f;
    g.st = ks;
    g(t);
  }
  __syncthreads();
  if (blockIdx.x == 0 && k < W) reinterpret_cast<uint32_t*>(M.stOut)[k] = lds[k];
  if (blockIdx.x == 0 && k == 0) post_status(M.mirror, ks->status);
  return ks;
}

// ---------------------------------------------------------------------------
// the LDU row
// ---------------------------------------------------------------------------

// Row i of A*x: diag, owner faces ascending, neighbour faces in losort order.
// Short rows (shortRows set): the face loops are unrolled to kShortRow with
// per-lane predication, so every coefficient/address load of the row issues
// before any x gather, and every gather before the sums -- three waves of
// independent memory requests instead of a chain per face.  The terms are
// added in exactly the order of the loop form, so both give identical bits.
// One row of the row-slot layout held in registers: loaded once, applied to
// any number of vectors (the right-hand sides of a batch share the matrix).
template <typename T>
struct RowSlots {
  T c[2 * kShortRow];
  int col[2 * kShortRow];  // a valid row index for every slot (the row itself if empty)
  unsigned m;              // bit q: slot q holds a face
  bool faces;              // the row is in a face tile: it runs the face loops
  // row i of slot tile t (i / kSlotTile, or the batch's work tile)
  __device__ __forceinline__ void load(const LduOp<T>& A, int i, int t) {
    const size_t n = (size_t)A.n;
    // the row's own coefficients first: independent coalesced loads (in
    // bounds for every row; a face tile's are not used)
#pragma unroll
    for (int q = 0; q < 2 * kShortRow; ++q)
      if (q < kShortRow || !A.slotSym) c[q] = A.slotCoef[q * n + i];
    const unsigned rb = A.slotRow[i];
    const int t0 = __builtin_amdgcn_readfirstlane(t);
    if (__all(t == t0))
      decode(A, A.slotTile + (size_t)t0 * kSlotDesc, rb, i);
    else
      decode(A, A.slotTile + (size_t)t * kSlotDesc, rb, i);
  }
  __device__ __forceinline__ void decode(const LduOp<T>& A, const int32_t* d, unsigned rb,
                                         int i) {
    const size_t n = (size_t)A.n;
    const int flags = d[6];
    faces = (flags & kSlotFaces) != 0;
    if (faces) return;
    // symmetric: a neighbour face's coefficient is stored once, in its owner's
    // row, except in tiles whose neighbour planes hold them
    const bool gather = A.slotSym && !(flags & kSlotNbrPlanes);
    if (A.slotSym && !gather) {
#pragma unroll
      for (int q = kShortRow; q < 2 * kShortRow; ++q) c[q] = A.slotCoef[q * n + i];
    }
    unsigned own;
    if (flags & kSlotPatterned) {
      // columns from the tile bases alone: the gathers need not wait for the
      // slot byte, which only selects the terms
#pragma unroll
      for (int q = 0; q < 2 * kShortRow; ++q) col[q] = min(max(i + d[q], 0), A.n - 1);
      m = rb;
      own = (unsigned)flags >> 8;
    } else {
      int o[2 * kShortRow];
      slot_decode(d, rb, A.slotOff, n, i, o, own);
      m = 0;
#pragma unroll
      for (int q = 0; q < 2 * kShortRow; ++q) {
        col[q] = i + o[q];
        m |= (o[q] != 0 ? 1u : 0u) << q;
      }
    }
    if (gather) {
#pragma unroll
      for (int q = kShortRow; q < 2 * kShortRow; ++q) {
        const unsigned sq = (own >> (2 * (q - kShortRow))) & 3u;
        c[q] = A.slotCoef[sq * n + (size_t)col[q]];
      }
    }
  }
  // diag term, then the owner slots, then the neighbour slots; empty slots
  // are skipped, so the terms and their order are those of the face loops
  template <typename XF>
  __device__ __forceinline__ T apply(const T* diag, int i, T xi, XF xf) const {
    return apply_d(diag[i], xi, xf);
  }
  // the same with the row's diagonal coefficient di already at hand
  template <typename XF>
  __device__ __forceinline__ T apply_d(T di, T xi, XF xf) const {
    T s = di * xi;
    T xq[2 * kShortRow];
#pragma unroll
    for (int q = 0; q < 2 * kShortRow; ++q) xq[q] = xf(col[q]);
#pragma unroll
    for (int q = 0; q < 2 * kShortRow; ++q) {
      const T t = s + c[q] * xq[q];
      s = ((m >> q) & 1u) ? t : s;
    }
    return s;
  }
};

// Row i of A x through the LDU face arrays: the schedules without row slots,
// and the rows of face tiles.
template <typename T, typename XF>
__device__ __forceinline__ T ldu_row_faces(const LduOp<T>& A, int i, T xi, XF xf) {
  const int f0 = A.ownStart[i], f1 = A.ownStart[i + 1];
  const int k0 = A.losortStart[i], k1 = A.losortStart[i + 1];
  T s = A.diag[i] * xi;
  if (A.shortRows) {
    T cu[kShortRow], cl[kShortRow], xu[kShortRow], xl[kShortRow];
    int ju[kShortRow], jl[kShortRow];
#pragma unroll
    for (int q = 0; q < kShortRow; ++q) {
      if (f0 + q < f1) {
        cu[q] = A.upper[f0 + q];
        ju[q] = A.uppAddr[f0 + q];
      }
      if (k0 + q < k1) {
        const int k = A.direct ? A.losort[k0 + q] : k0 + q;
        cl[q] = A.lowerL[k];
        jl[q] = A.nbrAddr[k];
      }
    }
#pragma unroll
    for (int q = 0; q < kShortRow; ++q) {
      if (f0 + q < f1) xu[q] = xf(ju[q]);
      if (k0 + q < k1) xl[q] = xf(jl[q]);
    }
#pragma unroll
    for (int q = 0; q < kShortRow; ++q)
      if (f0 + q < f1) s += cu[q] * xu[q];
#pragma unroll
    for (int q = 0; q < kShortRow; ++q)
      if (k0 + q < k1) s += cl[q] * xl[q];
    return s;
  }
  for (int f = f0; f < f1; ++f) s += A.upper[f] * xf(A.uppAddr[f]);
  if (!A.direct) {
    for (int k = k0; k < k1; ++k) s += A.lowerL[k] * xf(A.nbrAddr[k]);
  } else {
    for (int k = k0; k < k1; ++k) {
      const int f = A.losort[k];
      s += A.lowerL[f] * xf(A.nbrAddr[f]);
    }
  }
  return s;
}

// Row i of A x; t is row i's slot tile (ldu_row below: i / kSlotTile).
template <typename T, typename XF>
__device__ __forceinline__ T ldu_row_t(const LduOp<T>& A, int i, int t, T xi, XF xf) {
  if (A.slots) {
    // coefficients and offsets are independent coalesced loads (no row-start
    // indirection)
    RowSlots<T> rs;
    rs.load(A, i, t);
    if (!rs.faces) return rs.apply(A.diag, i, xi, xf);
  }
  return ldu_row_faces(A, i, xi, xf);
}

template <typename T, typename XF>
__device__ __forceinline__ T ldu_row(const LduOp<T>& A, int i, T xi, XF xf) {
  return ldu_row_t(A, i, i / kSlotTile, xi, xf);
}

// ---------------------------------------------------------------------------
// row loop of the single-system kernels
// ---------------------------------------------------------------------------

// Loads through the constant address space at a wave-uniform address compile
// to scalar loads.  The tile descriptors are written once per mesh and only
// read by the kernels.
#define HF_SCALAR_AS __attribute__((address_space(4)))

// A slot tile's descriptor in scalar registers.
struct TileDesc {
  int base[2 * kShortRow];
  int flags;
};

__device__ __forceinline__ TileDesc tile_desc(const int32_t* slotTile, int t) {
  const HF_SCALAR_AS int32_t* d =
      (const HF_SCALAR_AS int32_t*)(slotTile + (size_t)t * kSlotDesc);
  TileDesc s;
#pragma unroll
  for (int q = 0; q < 2 * kShortRow; ++q) s.base[q] = d[q];
  s.flags = d[6];
  return s;
}

// Row i of A x in a slot tile with descriptor d (PAT: a patterned tile).
// Straight-line code: the row's coefficient, slot-byte, diagonal and x loads
// are independent of one another, so they issue together; empty slots are
// masked by selects.  The terms and their order are those of ldu_row_t.
template <typename T, bool PAT, typename XF>
__device__ __forceinline__ T slot_row(const LduOp<T>& A, const TileDesc& d, int i, T xi, XF xf) {
  constexpr int K = kShortRow, S = 2 * kShortRow;
  const size_t n = (size_t)A.n;
  const T* __restrict__ C = A.slotCoef;
  const bool sym = A.slotSym != 0;
  T c[S];
  int col[S];
  unsigned m;
#pragma unroll
  for (int q = 0; q < K; ++q) c[q] = C[q * n + i];
  const unsigned rb = A.slotRow[i];
  if (PAT) {
    // columns from the tile bases alone (an absent slot is clamped to a valid
    // row); the neighbour faces' owner slots are the tile's (flags >> 8)
    const unsigned own = (unsigned)d.flags >> 8;
    const bool gather = sym && !(d.flags & kSlotNbrPlanes);
#pragma unroll
    for (int q = 0; q < S; ++q) col[q] = min(max(i + d.base[q], 0), A.n - 1);
#pragma unroll
    for (int q = K; q < S; ++q) {
      const int plane = gather ? (int)((own >> (2 * (q - K))) & 3u) : q;
      c[q] = C[(size_t)plane * n + (gather ? col[q] : i)];
    }
    m = rb;
  } else {
    uint32_t w[K];
#pragma unroll
    for (int p = 0; p < K; ++p) w[p] = A.slotOff[p * n + i];
    m = 0;
#pragma unroll
    for (int p = 0; p < K; ++p) {
      const int d0 = (int)(int16_t)(w[p] & 0xffffu), d1 = (int)(int16_t)(w[p] >> 16);
      const int o0 = d0 == kSlotEmpty ? 0 : d.base[2 * p] + d0;
      const int o1 = d1 == kSlotEmpty ? 0 : d.base[2 * p + 1] + d1;
      col[2 * p] = i + o0;
      col[2 * p + 1] = i + o1;
      m |= (o0 != 0 ? 1u : 0u) << (2 * p);
      m |= (o1 != 0 ? 1u : 0u) << (2 * p + 1);
    }
    // symmetric tiles flagged kSlotNbrPlanes hold their neighbour
    // coefficients in their own planes 3..5, as asymmetric operators do
    const bool gather = sym && !(d.flags & kSlotNbrPlanes);
#pragma unroll
    for (int q = K; q < S; ++q) {
      const size_t plane = gather ? ((rb >> (2 * (q - K))) & 3u) : (unsigned)q;
      c[q] = C[plane * n + (gather ? col[q] : i)];
    }
  }
  T s = A.diag[i] * xi;
  T xq[S];
#pragma unroll
  for (int q = 0; q < S; ++q) xq[q] = xf(col[q]);
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const T t = s + c[q] * xq[q];
    s = ((m >> q) & 1u) ? t : s;
  }
  return s;
}

// Calls fn(i, row) for every row i in [r0, r1), where row(xi, xf) returns row
// i of A x for xi = x[i] and xf(j) = x[j].  With row slots the rows go in
// passes of blockDim rows: a wavefront's 64 rows of a pass lie in one slot
// tile (blockDim and the slot tile are multiples of 64, r0 a multiple of
// blockDim), whose descriptor is read once per pass into scalar registers.
template <typename T, typename Fn>
__device__ __forceinline__ void for_rows(const LduOp<T>& A, int r0, int r1, Fn fn) {
  if (A.slots) {
    for (int base = r0; base < r1; base += blockDim.x) {
      const int w0 = __builtin_amdgcn_readfirstlane(base + (int)(threadIdx.x & ~63u));
      if (w0 >= r1) continue;
      const TileDesc d = tile_desc(A.slotTile, w0 / kSlotTile);
      const int i = base + (int)threadIdx.x;
      if (i >= r1) continue;
      if (d.flags & kSlotPatterned)
        fn(i, [&](T xi, auto xf) { return slot_row<T, true>(A, d, i, xi, xf); });
      else if (d.flags & kSlotFaces)
        fn(i, [&](T xi, auto xf) { return ldu_row_faces(A, i, xi, xf); });
      else
        fn(i, [&](T xi, auto xf) { return slot_row<T, false>(A, d, i, xi, xf); });
    }
  } else {
    for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x)
      fn(i, [&](T xi, auto xf) { return ldu_row_t(A, i, i / kSlotTile, xi, xf); });
  }
}

// for_rows over the tiles of logical workgroup lb in an XCD march: on a grid
// whose XCD mapping gives each XCD one contiguous run of workgroups
// (xcdRemap 1, a multiple of 8 workgroups), the XCD's share of the mesh is
// one contiguous slab of tiles, and its q workgroups walk the slab together:
// pass p of local workgroup j is tile slab + p q + j.  The XCD's workgroups
// then work on one window of q tiles at a time instead of q windows spread
// over the slab, so the neighbour planes a pass gathers were loaded by the
// pass before, in the same L2.  The rows and their terms are those of
// for_rows; only which workgroup sums which rows, and so the grouping of
// the FP64 partial sums, changes.  With one tile per workgroup (the
// per-tile launch shape) every workgroup gets the tile it has contiguously.
template <typename T, typename Fn>
__device__ __forceinline__ void for_rows_xcd(const LduOp<T>& A, int lb, int r0, int r1, Fn fn) {
  const int G = (int)gridDim.x;
  if ((G & 7) != 0 || A.xcdRemap != 1) {
    for_rows(A, r0, r1, fn);
    return;
  }
  const int bs = (int)blockDim.x, q = G >> 3;
  const int ntiles = (A.n + bs - 1) / bs;
  const int xcd = lb / q, j = lb - xcd * q;
  const int per = (ntiles + 7) >> 3;
  const int s0 = xcd * per, s1 = min(s0 + per, ntiles);
  for (int t = s0 + j; t < s1; t += q) for_rows(A, t * bs, min((t + 1) * bs, A.n), fn);
}

// ---------------------------------------------------------------------------
// scalar finalisation (runs in thread 0 of the last workgroup)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool of_converged(const KState* s, double res) {
  return res < s->tol || (s->relTol > kSmall && res < s->relTol * s->initRes);
}

// End of one PCG/BiCGStab iteration in inner mode: the host-loop checks of
// metalFoamBackend.C:676-715 / 908-943, evaluated at sync-batch boundaries.
__device__ __forceinline__ void inner_iteration_end(KState* s, double checkB) {
  s->copyBest = 0;
  s->it += 1;
  s->batchLeft -= 1;
  if (s->batchLeft > 0) return;
  const double res = s->res;
  if (!isfinite(res) || !isfinite(checkB)) {
    s->status = ST_SINGULAR;
    return;
  }
  if (res < s->bestRes) {
    s->bestRes = res;
    s->noProgress = 0;
    s->copyBest = 1;
  } else {
    s->noProgress += s->batchSize;
  }
  if (res <= s->target) {
    s->status = ST_CONVERGED;
    return;
  }
  if (s->noProgress >= 50) {
    s->status = ST_STAGNATED;
    return;
  }
  if (s->it >= s->maxIter) {
    s->status = ST_MAXITER;
    return;
  }
  const int left = s->maxIter - s->it;
  s->batchSize = s->ips < left ? s->ips : left;
  s->batchLeft = s->batchSize;
}

// Chebyshev constants (encCheb, metalFoamBackend.C:349-372), in FP64, rounded
// to FP32 when round32.
__host__ __device__ inline void cheb_constants(KState& s, int order, double beta, double ratio,
                                               bool round32) {
  const double alpha = beta / ratio;
  const double theta = 0.5 * (beta + alpha);
  const double delta = 0.5 * (beta - alpha);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  auto rd = [&](double x) { return round32 ? (double)(float)x : x; };
  s.chebC1 = rd(1.0 / theta);
  for (int i = 1; i < order && i <= kMaxCheb; ++i) {
    const double rhoNew = 1.0 / (2.0 * sigma - rho);
    s.chebC2[i - 1] = rd(rhoNew * rho);
    s.chebC3[i - 1] = rd(2.0 * rhoNew / delta);
    rho = rhoNew;
  }
}

// OpenFOAM loop condition: while ((++nIter < maxIter && !converged) || nIter < minIter)
__device__ __forceinline__ void openfoam_iteration_end(KState* s) {
  s->it += 1;
  const bool conv = of_converged(s, s->res);
  const bool cont = (s->it < s->maxIter && !conv) || s->it < s->minIter;
  if (!cont) s->status = conv ? ST_CONVERGED : ST_MAXITER;
}

// rho = wA.rA and beta (reduce_scalars mode 0, metalFoamKernels.metal:336-340)
template <typename T>
__device__ __forceinline__ void finalize_beta(KState* st, double t) {
  const bool first = (st->it == 0);
  st->rhoOld = st->rho;
  if (st->mode == MODE_INNER) {
    const T rho = (T)t;
    st->rho = rho;
    st->beta = first ? 0.0 : (double)(rho / (T)st->rhoOld);
  } else {
    st->rho = t;
    st->beta = first ? 0.0 : t / st->rhoOld;
  }
  st->alphaDue = 1;
}

// ---------------------------------------------------------------------------
// PCG row bodies: one row of each PCG kernel, shared by the launched kernels
// and the persistent loop (hipfoam_persist.hpp), so both evaluate the same
// float expressions in the same order.
// ---------------------------------------------------------------------------

// One Chebyshev step after the first, fused with its SpMV (ldu_spmv +
// cheb_step of metalFoamKernels.metal:112-194):
//   v = A z_{s-1};  w_s = c2 w_{s-1} + c3 rD (r - v);  z_s = z_{s-1} + w_s
// LAST:  z_s is the preconditioned residual wA; *dot += wA . rA.
template <typename T, bool LAST, bool COH>
struct ChebStepRow {
  const T* __restrict__ r;
  const T* __restrict__ rD;
  const T* __restrict__ zin;
  T* __restrict__ w;
  T* __restrict__ zout;
  T c2, c3;
  double* dot;
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    // the row's own operands load before the gathers, in one batch
    const T ri = vld<COH>(r, i), rDi = rD[i];
    const T zi = vld<COH>(zin, i), wold = vld<COH>(w, i);
    const T v = row(zi, [&](int j) { return vld<COH>(zin, j); });
    const T wi = c2 * wold + c3 * rDi * (ri - v);
    const T zi_new = zi + wi;
    if (!LAST) w[i] = wi;
    zout[i] = zi_new;
    if (LAST) *dot += (double)(zi_new * ri);
  }
};

// First Chebyshev step of a PCG iteration (k_cheb_first below).  PEND: the
// previous iteration's update is applied first: rn = r - alpha Ap (gathered
// on the fly at the neighbours), delta += alpha p, deltaBest = delta when
// copy; v[1] += |rn|.  Then z_0 = c1 rD rn and the first step; LAST: v[0] +=
// wA . rn.
template <typename T, bool LAST, bool COH, bool PEND>
struct ChebFirstRow {
  const T* __restrict__ rin;
  T* __restrict__ rout;
  const T* __restrict__ rD;
  const T* __restrict__ p;
  const T* __restrict__ Ap;
  T* __restrict__ delta;
  T* __restrict__ deltaBest;
  T* __restrict__ w;
  T* __restrict__ zout;
  T c1, c2, c3, alpha;
  bool copy;
  double* v;  // [2]: rho, sum|rn|
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    if constexpr (PEND) {
      const T ri = vld<COH>(rin, i) - alpha * vld<COH>(Ap, i);
      const T di = vld<COH>(delta, i), pi = vld<COH>(p, i), rDi = rD[i];
      const T z0 = c1 * rDi * ri;
      const T y = row(z0, [&](int j) {
        return c1 * rD[j] * (vld<COH>(rin, j) - alpha * vld<COH>(Ap, j));
      });
      const T wi = c2 * z0 + c3 * rDi * (ri - y);
      const T zn = z0 + wi;
      if (copy) deltaBest[i] = di;
      delta[i] = di + alpha * pi;
      rout[i] = ri;
      if (!LAST) w[i] = wi;
      zout[i] = zn;
      v[1] += (double)fabs(ri);
      if (LAST) v[0] += (double)(zn * ri);
    } else {
      const T ri = vld<COH>(rin, i), rDi = rD[i];
      const T z0 = c1 * rDi * ri;
      const T y = row(z0, [&](int j) { return c1 * rD[j] * vld<COH>(rin, j); });
      const T wi = c2 * z0 + c3 * rDi * (ri - y);
      const T zn = z0 + wi;
      rout[i] = ri;
      if (!LAST) w[i] = wi;
      zout[i] = zn;
      if (LAST) v[0] += (double)(zn * ri);
    }
  }
};

// p = wA + beta p_old formed at the gathers, Ap = A p, *dot += p . Ap
// (k_pcg_p_spmv below).
template <typename T, bool COH>
struct PcgPRow {
  const T* __restrict__ wA;
  const T* __restrict__ pold;
  T* __restrict__ pnew;
  T* __restrict__ Ap;
  T beta;
  double* dot;
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    const T pi = vld<COH>(wA, i) + beta * vld<COH>(pold, i);
    const T y = row(pi, [&](int j) { return vld<COH>(wA, j) + beta * vld<COH>(pold, j); });
    pnew[i] = pi;
    Ap[i] = y;
    *dot += (double)(pi * y);
  }
};

// ---------------------------------------------------------------------------
// PCG kernels
// ---------------------------------------------------------------------------

// One Chebyshev step after the first, fused with its SpMV (ChebStepRow).
// LAST:  z_s is the preconditioned residual wA; fuses dot(wA, rA) and beta.
// (The first step, z_0 = w_0 = c1 rD r, is k_cheb_first.)
template <typename T, bool LAST>
__global__ __launch_bounds__(1024) void k_cheb_step(LduOp<T> A, const T* __restrict__ r,
                                                    const T* __restrict__ rD,
                                                    const T* __restrict__ zin,
                                                    T* __restrict__ w, T* __restrict__ zout,
                                                    KState* st, double* partials, int pstride,
                                                    int step) {
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const T c2 = (T)st->chebC2[step - 1];
  const T c3 = (T)st->chebC3[step - 1];
  double dot = 0.0;
  const ChebStepRow<T, LAST, false> row_step{r, rD, zin, w, zout, c2, c3, &dot};
  // the last step runs on the reducing launch shape: an XCD march
  if (LAST)
    for_rows_xcd(A, lb, r0, r1, row_step);
  else
    for_rows(A, r0, r1, row_step);
  if (LAST) {
    double v[1] = {dot};
    block_partials<1>(v, partials, pstride, lb);
  }
}

// First Chebyshev step of a PCG iteration fused with the previous
// iteration's update (pcg_update_sol_res + cheb_init + ldu_spmv + cheb_step,
// metalFoamKernels.metal:112-194, 250-276).  When the state has an update
// pending (pend): deltaBest = delta if the last check improved the residual,
// delta += alpha p, and the updated residual rn = r - alpha Ap is written to
// rout; neighbour gathers form rn on the fly from r and Ap, so r ping-pongs
// (rin -> rout) between iterations.  Partial sum|rn| goes to row 1 of the
// partials (FinResBeta checks it with the next beta).  Without a pending
// update (first iteration) rout = rin.  Then z_0 = c1 rD rn and the first
// step as in k_cheb_step; LAST also reduces rho = dot(wA, rn) into row 0.
// MERGED: alpha's finaliser (FinAlpha of p.Ap) runs here.
template <typename T, bool LAST, typename F, bool MERGED>
__global__ __launch_bounds__(1024) void k_cheb_first(
    LduOp<T> A, const T* __restrict__ rin, T* __restrict__ rout, const T* __restrict__ rD,
    const T* __restrict__ p, const T* __restrict__ Ap, T* __restrict__ delta,
    T* __restrict__ deltaBest, T* __restrict__ w, T* __restrict__ zout, const KState* stIn,
    double* partials, int pstride, Merge M, F fin) {
  const KState* st = stIn;
  if constexpr (MERGED) {
    __shared__ KState ks;
    st = merged_finalize<1>(stIn, M, pstride, fin, &ks);
  }
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const T c2 = (T)st->chebC2[0];
  const T c3 = (T)st->chebC3[0];
  const T c1 = (T)st->chebC1;
  double v[2] = {0.0, 0.0};  // rho, sum|rn|
  if (st->pend) {
    const T alpha = (T)st->alpha;
    const bool copy = st->copyBest != 0;
    for_rows(A, r0, r1,
             ChebFirstRow<T, LAST, false, true>{rin, rout, rD, p, Ap, delta, deltaBest, w, zout,
                                                c1, c2, c3, alpha, copy, v});
  } else {
    for_rows(A, r0, r1,
             ChebFirstRow<T, LAST, false, false>{rin, rout, rD, p, Ap, delta, deltaBest, w,
                                                 zout, c1, c2, c3, (T)0, false, v});
  }
  block_partials<2>(v, partials, pstride, lb);
}

// Jacobi apply fused with dot(wA, rA) (pcg_precond_dot, metalFoamKernels.metal:211-233)
template <typename T>
__global__ __launch_bounds__(1024) void k_jacobi_dot(int n, int xcdRemap, const T* __restrict__ r,
                                                     const T* __restrict__ rD,
                                                     T* __restrict__ wA, KState* st,
                                                     double* partials, int pstride) {
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  double dot = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const T wv = r[i] * rD[i];
    wA[i] = wv;
    dot += (double)(wv * r[i]);
  }
  double v[1] = {dot};
  block_partials<1>(v, partials, pstride, lb);
}

// p = wA + beta p fused into A p (pcg_update_p + ldu_spmv_dotw with w = p,
// metalFoamKernels.metal:62-109, 236-247); p ping-pongs (pold -> pnew) so
// neighbour gathers read the previous p.  Finalises wApA and alpha
// (reduce_scalars mode 1) or, in OpenFOAM mode, the singularity check.
// MERGED: beta's finaliser runs here (merged_finalize over NVM partial
// rows): FinBeta of the Jacobi dot, or FinResBeta of the last Chebyshev
// step's rho and k_cheb_first's sum|r|.
template <typename T, typename F, bool MERGED, int NVM = 1>
__global__ __launch_bounds__(1024) void k_pcg_p_spmv(LduOp<T> A, const T* __restrict__ wA,
                                                     const T* __restrict__ pold,
                                                     T* __restrict__ pnew, T* __restrict__ Ap,
                                                     const KState* stIn, double* partials,
                                                     int pstride, Merge M, F fin) {
  const KState* st = stIn;
  if constexpr (MERGED) {
    __shared__ KState ks;
    st = merged_finalize<NVM>(stIn, M, pstride, fin, &ks);
  }
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const T beta = (T)st->beta;
  double dot = 0.0;
  for_rows_xcd(A, lb, r0, r1, PcgPRow<T, false>{wA, pold, pnew, Ap, beta, &dot});
  double v[1] = {dot};
  block_partials<1>(v, partials, pstride, lb);
}

// delta += alpha p; r -= alpha Ap; partial sum|r| (pcg_update_sol_res,
// metalFoamKernels.metal:250-276).  Before the update, copies delta to
// deltaBest when the previous sync batch improved the residual.
// MERGED: alpha's finaliser (FinAlpha of p.Ap) runs here.
template <typename T, typename F, bool MERGED>
__global__ __launch_bounds__(1024) void k_pcg_update(int n, int xcdRemap, T* __restrict__ delta,
                                                     T* __restrict__ deltaBest,
                                                     T* __restrict__ r, const T* __restrict__ p,
                                                     const T* __restrict__ Ap, const KState* stIn,
                                                     double* partials, int pstride, Merge M,
                                                     F fin) {
  const KState* st = stIn;
  if constexpr (MERGED) {
    __shared__ KState ks;
    st = merged_finalize<1>(stIn, M, pstride, fin, &ks);
  }
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  const T alpha = (T)st->alpha;
  const bool copy = st->copyBest != 0;
  double m = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const T d = delta[i];
    if (copy) deltaBest[i] = d;
    delta[i] = d + alpha * p[i];
    const T ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    m += (double)fabs(ri);
  }
  double v[1] = {m};
  block_partials<1>(v, partials, pstride, lb);
}

// ---------------------------------------------------------------------------
// BiCGStab kernels (metalFoamKernels.metal:279-458; metalFoamBackend.C:784-906)
// ---------------------------------------------------------------------------

// BiCGStab row bodies, shared by the launched kernels and the persistent
// loop (COH: vectors the same launch writes are read with device-scope loads).

// p = rA + beta (p_old - omega AyA_old) and y = rD p formed at the gathers,
// AyA = A y; *d += rA0 . AyA (k_bicg_py_spmv).
template <typename T, bool COH>
struct BicgPyRow {
  const T* __restrict__ r;
  const T* __restrict__ rD;
  const T* __restrict__ pold;
  T* __restrict__ pnew;
  const T* __restrict__ AyAold;
  T* __restrict__ AyAnew;
  T* __restrict__ yA;
  const T* __restrict__ rA0;
  T beta, omega;
  double* d;
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    const T a0 = rA0[i];
    const T pi = vld<COH>(r, i) + beta * (vld<COH>(pold, i) - omega * vld<COH>(AyAold, i));
    const T yi = rD[i] * pi;
    const T y = row(yi, [&](int j) {
      const T pj = vld<COH>(r, j) + beta * (vld<COH>(pold, j) - omega * vld<COH>(AyAold, j));
      return rD[j] * pj;
    });
    pnew[i] = pi;
    yA[i] = yi;
    AyAnew[i] = y;
    *d += (double)(a0 * y);
  }
};

// sA = rA - alpha AyA and zA = rD sA formed at the gathers, tA = A zA;
// d += |sA|, tA . sA, tA . tA (k_bicg_sz_spmv).
template <typename T, bool COH>
struct BicgSzRow {
  const T* __restrict__ r;
  const T* __restrict__ rD;
  const T* __restrict__ AyA;
  T* __restrict__ sA;
  T* __restrict__ zA;
  T* __restrict__ tA;
  T alpha;
  double* d;  // [3]
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    const T si = vld<COH>(r, i) - alpha * vld<COH>(AyA, i);
    const T zi = rD[i] * si;
    const T t = row(zi, [&](int j) {
      const T sj = vld<COH>(r, j) - alpha * vld<COH>(AyA, j);
      return rD[j] * sj;
    });
    sA[i] = si;
    zA[i] = zi;
    tA[i] = t;
    d[0] += (double)fabs(si);
    d[1] += (double)(si * t);
    d[2] += (double)(t * t);
  }
};

// Rows [r0, r1), stride blockDim, of the BiCGStab update (k_bicg_update_rho):
// deltaBest = delta when copy, delta += alpha yA + omega zA, rA = sA - omega
// tA; v[0] += |rA|, v[1] += rA0 . rA.  Four rows per thread at a time:
// every row's loads issue before any row's arithmetic (one memory round trip
// instead of four); the rows and their sums go in the order of the plain loop.
template <typename T, bool COH>
__device__ __forceinline__ void bicg_update_rows(int r0, int r1, T* __restrict__ delta,
                                                 T* __restrict__ deltaBest,
                                                 const T* __restrict__ yA,
                                                 const T* __restrict__ zA, T* __restrict__ r,
                                                 const T* __restrict__ sA,
                                                 const T* __restrict__ tA,
                                                 const T* __restrict__ rA0, T alpha, T omega,
                                                 bool copy, double (&v)[2]) {
  constexpr int U = 4;
  const int bd = blockDim.x;
  int i0 = r0 + threadIdx.x;
  for (; i0 + (U - 1) * bd < r1; i0 += U * bd) {
    T d[U], y[U], z[U], s[U], t[U], a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * bd;
      d[u] = vld<COH>(delta, i);
      y[u] = vld<COH>(yA, i);
      z[u] = vld<COH>(zA, i);
      s[u] = vld<COH>(sA, i);
      t[u] = vld<COH>(tA, i);
      a[u] = rA0[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * bd;
      if (copy) deltaBest[i] = d[u];
      delta[i] = d[u] + (alpha * y[u] + omega * z[u]);
      const T ri = s[u] - omega * t[u];
      r[i] = ri;
      v[0] += (double)fabs(ri);
      v[1] += (double)(a[u] * ri);
    }
  }
  for (int i = i0; i < r1; i += bd) {
    const T d = vld<COH>(delta, i);
    if (copy) deltaBest[i] = d;
    delta[i] = d + (alpha * vld<COH>(yA, i) + omega * vld<COH>(zA, i));
    const T ri = vld<COH>(sA, i) - omega * vld<COH>(tA, i);
    r[i] = ri;
    v[0] += (double)fabs(ri);
    v[1] += (double)(rA0[i] * ri);
  }
}

// After an |sA| exit only delta += alpha yA is applied (bicg_early_update,
// metalFoamKernels.metal:415-426).
template <typename T, bool COH>
__device__ __forceinline__ void bicg_early_rows(int r0, int r1, T* __restrict__ delta,
                                                const T* __restrict__ yA, T alpha) {
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x)
    delta[i] = vld<COH>(delta, i) + alpha * vld<COH>(yA, i);
}

// rho = rA0.rA, beta (reduce_scalars mode 3).
template <typename T>
__global__ __launch_bounds__(1024) void k_bicg_rho(int n, int xcdRemap, const T* __restrict__ rA0,
                                                   const T* __restrict__ r, KState* st,
                                                   double* partials, int pstride) {
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  double d = 0.0;
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) d += (double)(rA0[i] * r[i]);
  double v[1] = {d};
  block_partials<1>(v, partials, pstride, lb);
}

// One BiCGStab iteration is three row kernels, each followed by a one-
// workgroup finalise:
//   k_bicg_py_spmv   p = rA + beta (p_old - omega AyA_old), yA = rD p,
//                    AyA = A yA, partial rA0.AyA          -> FinBicgAlpha
//   k_bicg_sz_spmv   sA = rA - alpha AyA, zA = rD sA, tA = A zA,
//                    partials |sA|, tA.sA, tA.tA          -> FinBicgSOmega
//   k_bicg_update_rho delta += alpha yA + omega zA, rA = sA - omega tA,
//                    partials |rA|, rA0.rA                -> FinBicgUpdateRho
// (bicg_update_p, emul, ldu_spmv_dotw, bicg_s_res, dot_partial,
// bicg_update_sol_res and reduce_scalars modes 3-5 of
// metalFoamKernels.metal:279-458).  The SpMV kernels form the gathered
// neighbour values on the fly from the vectors they derive from, so p and
// AyA ping-pong between iterations; rho of the next iteration is reduced in
// the update kernel.  Every value is the same float expression the separate
// kernels evaluate, so the results are the same.

template <typename T>
__global__ __launch_bounds__(1024) void k_bicg_py_spmv(LduOp<T> A, const T* __restrict__ r,
                                                       const T* __restrict__ rD,
                                                       const T* __restrict__ pold,
                                                       T* __restrict__ pnew,
                                                       const T* __restrict__ AyAold,
                                                       T* __restrict__ AyAnew,
                                                       T* __restrict__ yA,
                                                       const T* __restrict__ rA0, KState* st,
                                                       double* partials, int pstride) {
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const T beta = (T)st->beta, omega = (T)st->omega;
  double d = 0.0;
  for_rows_xcd(A, lb, r0, r1,
               BicgPyRow<T, false>{r, rD, pold, pnew, AyAold, AyAnew, yA, rA0, beta, omega, &d});
  double v[1] = {d};
  block_partials<1>(v, partials, pstride, lb);
}

// MERGED: alpha's finaliser (FinBicgAlpha of rA0.AyA) runs here.
template <typename T, typename F, bool MERGED>
__global__ __launch_bounds__(1024) void k_bicg_sz_spmv(LduOp<T> A, const T* __restrict__ r,
                                                       const T* __restrict__ rD,
                                                       const T* __restrict__ AyA,
                                                       T* __restrict__ sA, T* __restrict__ zA,
                                                       T* __restrict__ tA, const KState* stIn,
                                                       double* partials, int pstride, Merge M,
                                                       F fin) {
  const KState* st = stIn;
  if constexpr (MERGED) {
    __shared__ KState ks;
    st = merged_finalize<1>(stIn, M, pstride, fin, &ks);
  }
  if (st->status != ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const T alpha = (T)st->alpha;
  double d[3] = {0.0, 0.0, 0.0};
  for_rows_xcd(A, lb, r0, r1, BicgSzRow<T, false>{r, rD, AyA, sA, zA, tA, alpha, d});
  block_partials<3>(d, partials, pstride, lb);
}

// After an |sA| exit (status EARLY) only delta += alpha yA is applied
// (bicg_early_update, metalFoamKernels.metal:415-426).
// MERGED: the |sA| test and omega (FinBicgSOmega) run here.
template <typename T, typename F, bool MERGED>
__global__ __launch_bounds__(1024) void k_bicg_update_rho(int n, int xcdRemap,
                                                          T* __restrict__ delta,
                                                          T* __restrict__ deltaBest,
                                                          const T* __restrict__ yA,
                                                          const T* __restrict__ zA,
                                                          T* __restrict__ r,
                                                          const T* __restrict__ sA,
                                                          const T* __restrict__ tA,
                                                          const T* __restrict__ rA0,
                                                          const KState* stIn, double* partials,
                                                          int pstride, Merge M, F fin) {
  const KState* st = stIn;
  if constexpr (MERGED) {
    __shared__ KState ks;
    st = merged_finalize<3>(stIn, M, pstride, fin, &ks);
  }
  const int status = st->status;
  if (status != ST_RUNNING && status != ST_EARLY) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  const T alpha = (T)st->alpha;
  if (status == ST_EARLY) {
    bicg_early_rows<T, false>(r0, r1, delta, yA, alpha);
    return;  // FinBicgUpdateRho completes the |sA| exit
  }
  double v[2] = {0.0, 0.0};
  bicg_update_rows<T, false>(r0, r1, delta, deltaBest, yA, zA, r, sA, tA, rA0, alpha,
                             (T)st->omega, st->copyBest != 0, v);
  block_partials<2>(v, partials, pstride, lb);
}

// ---------------------------------------------------------------------------
// inner-solve initialisation
// ---------------------------------------------------------------------------

// Row i of the inner-solve initialisation (k_inner_init); useMean: m applies.
template <typename T>
__device__ __forceinline__ void inner_init_row(int i, const double* __restrict__ r64,
                                               bool useMean, double m,
                                               const double* __restrict__ proj,
                                               T* __restrict__ rA, T* __restrict__ rA0,
                                               T* __restrict__ delta, T* __restrict__ p0,
                                               T* __restrict__ p1, T* __restrict__ AyA) {
  const T v = (T)(useMean ? (proj ? r64[i] - m * proj[i] : r64[i] - m) : r64[i]);
  rA[i] = v;
  if (rA0) rA0[i] = v;
  if (delta) delta[i] = (T)0;
  if (p0) p0[i] = (T)0;
  if (p1) p1[i] = (T)0;
  if (AyA) AyA[i] = (T)0;
}

// rA = T(r64 - m v) with m = *mean and v = proj (or 1 without proj): the
// zero-mean projection of the deflated inner right-hand side, or with
// proj = u = A 1 the projection along the pinned rows; delta = p0 = p1 =
// AyA = 0; rA0 = rA (BiCGStab).  Workgroup 0 also writes the solve's start
// state s into *kst and *kst2 (the merged stages read the first iteration's
// state from kst2).
template <typename T>
__global__ __launch_bounds__(1024) void k_inner_init(int n, const double* __restrict__ r64,
                                                     const double* mean,
                                                     const double* __restrict__ proj,
                                                     T* __restrict__ rA, T* __restrict__ rA0,
                                                     T* __restrict__ delta, T* __restrict__ p0,
                                                     T* __restrict__ p1, T* __restrict__ AyA,
                                                     KStateWords s, KState* kst, KState* kst2) {
  if (blockIdx.x == 0) {
    constexpr int W = (int)(sizeof(KStateWords) / sizeof(uint32_t));
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
      reinterpret_cast<uint32_t*>(kst)[k] = s.w[k];
      reinterpret_cast<uint32_t*>(kst2)[k] = s.w[k];
    }
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  inner_init_row<T>(i, r64, mean != nullptr, mean ? *mean : 0.0, proj, rA, rA0, delta, p0, p1,
                    AyA);
}

// ---------------------------------------------------------------------------
// FP64 refinement kernels (metalFoamSolver.C:101-365 on the device)
// ---------------------------------------------------------------------------

// Rows of the FP64 refinement kernels.

// u = A 1; v += psi, [|u| > 1e-8 |diag|], u.u, u (k_rowsum)
struct RowsumRow {
  const double* __restrict__ psi;
  const double* __restrict__ diag;
  double* __restrict__ u;
  double* v;  // [4]
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    const double pi = psi[i], di = diag[i];
    const double ui = row(1.0, [](int) { return 1.0; });
    u[i] = ui;
    v[0] += pi;
    v[1] += (fabs(ui) > 1e-8 * fabs(di)) ? 1.0 : 0.0;
    v[2] += ui * ui;
    v[3] += ui;
  }
};

// wA = A psi, r = b - wA; v += |r|, |wA - xbar u| + |b - xbar u|, r
// (k_initial_residual)
struct InitialResidualRow {
  const double* __restrict__ psi;
  const double* __restrict__ b;
  const double* __restrict__ u;
  double* __restrict__ r;
  double xbar;
  double* v;  // [3]
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    const double bi = b[i], t = u[i] * xbar;
    const double wA = row(psi[i], [&](int j) { return psi[j]; });
    const double ri = bi - wA;
    r[i] = ri;
    v[0] += fabs(ri);
    v[1] += fabs(wA - t) + fabs(bi - t);
    v[2] += ri;
  }
};

// psiNew = psiOld + (double(delta) - dm), formed at the gathers; r = b - A
// psiNew; v += u.r, |r|, r (k_refine_residual)
template <typename T>
struct RefineRow {
  const double* __restrict__ psiOld;
  const T* __restrict__ delta;
  double* __restrict__ psiNew;
  const double* __restrict__ b;
  const double* __restrict__ u;
  double* __restrict__ r;
  double dm;
  double* v;  // [3]
  template <typename Row>
  __device__ __forceinline__ void operator()(int i, Row row) const {
    auto pf = [&](int j) { return psiOld[j] + ((double)delta[j] - dm); };
    const double bi = b[i], ui = u[i];
    const double pi = pf(i);
    const double wA = row(pi, pf);
    psiNew[i] = pi;
    const double ri = bi - wA;
    r[i] = ri;
    v[0] += ui * ri;
    v[1] += fabs(ri);
    v[2] += ri;
  }
};

// Deflation mode of k_prepare_inner / FinPrepare: 0 off, 1 on, or
// kDeflateAuto: on when at most a tenth of the rows are pinned
// (metalFoamSolver.C:170-190), from the row-sum pass's count os->nPinned,
// which the stream has computed before (FinRowsum).
constexpr int kDeflateAuto = 2;

__device__ __forceinline__ int deflate_on(int mode, const OState* os, int n) {
  return mode == kDeflateAuto ? (10.0 * os->nPinned <= (double)n ? 1 : 0) : mode;
}

// Rows [r0, r1), stride blockDim, of k_prepare_inner: diag* = deflate ? diag
// - u : diag, diagT = T(diag*), rDT = T(1 / diag*); v[0] += sign flips,
// v[1] = max(v[1], 1 + sum |upper| / |diag*|) when wantBeta.
template <typename T>
__device__ __forceinline__ void prepare_rows(const LduOp<double>& A, int r0, int r1,
                                             const double* __restrict__ upperF,
                                             const int32_t* __restrict__ losort,
                                             const double* __restrict__ u, int deflate,
                                             int wantBeta, T* __restrict__ diagT,
                                             T* __restrict__ rDT, double (&v)[2]) {
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const double d = A.diag[i];
    double ds = d;
    if (deflate) {
      ds = d - u[i];
      if (ds * d <= kVSmall * fabs(ds)) v[0] += 1.0;
    }
    diagT[i] = (T)ds;
    rDT[i] = (T)(1.0 / ds);
    if (wantBeta) {
      double so = 0.0;
      for (int f = A.ownStart[i]; f < A.ownStart[i + 1]; ++f) so += fabs(upperF[f]);
      for (int k = A.losortStart[i]; k < A.losortStart[i + 1]; ++k)
        so += fabs(upperF[losort[k]]);
      v[1] = fmax(v[1], 1.0 + so / fabs(ds));
    }
  }
}

// Rows [r0, r1), stride blockDim: v += double(delta) (k_delta_mean)
template <typename T>
__device__ __forceinline__ void delta_mean_rows(int r0, int r1, const T* __restrict__ delta,
                                                double (&v)[1]) {
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) v[0] += (double)delta[i];
}

// Rows [r0, r1), stride blockDim: psi += c, r -= c u; v += |r|, r (k_const_shift)
__device__ __forceinline__ void const_shift_rows(int r0, int r1, double* __restrict__ psi,
                                                 double* __restrict__ r,
                                                 const double* __restrict__ u, double c,
                                                 double (&v)[2]) {
  for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    psi[i] += c;
    const double ri = r[i] - c * u[i];
    r[i] = ri;
    v[0] += fabs(ri);
    v[1] += ri;
  }
}

// u = A 1 (sumA), partials: sum psi, #{|u_i| > 1e-8 |diag_i|}, u.u
__global__ __launch_bounds__(1024) void k_rowsum(LduOp<double> A, const double* __restrict__ psi,
                                                 double* __restrict__ u, OState* os,
                                                 double* partials, int pstride) {
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for_rows(A, r0, r1, RowsumRow{psi, A.diag, u, v});
  block_partials<4>(v, partials, pstride, lb);
}

// wA = A psi; r = b - wA; partials: sum|r|, normFactor sum, sum r.
// normFactor = sum(|A psi - xbar u| + |b - xbar u|) + small (metalFoamSolver.C:120).
__global__ __launch_bounds__(1024) void k_initial_residual(LduOp<double> A,
                                                           const double* __restrict__ psi,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ u,
                                                           double* __restrict__ r, OState* os,
                                                           double* partials, int pstride) {
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  double v[3] = {0.0, 0.0, 0.0};
  for_rows(A, r0, r1, InitialResidualRow{psi, b, u, r, os->xbar, v});
  block_partials<3>(v, partials, pstride, lb);
}

// Inner operator of one restart: diag* = deflate ? diag - u : diag;
// diag32 = float(diag*), rD = float(1/diag*) (setCoeffs, metalFoamBackend.C:557-562);
// deflation sign-flip test (metalFoamSolver.C:199-207) and the Gershgorin bound
// max_i(1 + sum_f |upper_f| / |diag*_i|) (metalFoamSolver.C:229-249, |upper| only).
template <typename T>
__global__ __launch_bounds__(1024) void k_prepare_inner(LduOp<double> A,
                                                        const double* __restrict__ upperF,
                                                        const int32_t* __restrict__ losort,
                                                        const double* __restrict__ u,
                                                        int deflate, int wantBeta,
                                                        T* __restrict__ diagT,
                                                        T* __restrict__ rDT, OState* os,
                                                        double* partials, int pstride) {
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  double v[2] = {0.0, -1.0e308};
  prepare_rows<T>(A, r0, r1, upperF, losort, u, deflate_on(deflate, os, A.n), wantBeta, diagT,
                  rDT, v);
  block_partials<2, true>(v, partials, pstride, lb);
}

// Face staging, once per matrix: lowerL64[k] = lower[losort[k]] (upper when
// symmetric), upper32 = float(upper), lowerL32 = float(lowerL64).
__global__ __launch_bounds__(1024) void k_stage_faces(int nf, const int32_t* __restrict__ losort,
                                                      const double* __restrict__ upper,
                                                      const double* __restrict__ lower,
                                                      double* __restrict__ lowerL64,
                                                      float* __restrict__ upper32,
                                                      float* __restrict__ lowerL32,
                                                      float* __restrict__ lower32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const double* lo = lower ? lower : upper;
  const double l = lo[losort[k]];
  lowerL64[k] = l;
  upper32[k] = (float)upper[k];
  lowerL32[k] = (float)l;
  if (lower32) lower32[k] = (float)lo[k];
}

// The refinement kernels below (k_delta_mean, k_refine_residual,
// k_const_shift) return at entry when *guard -- the Krylov status of the
// inner solve they follow, NULL for none -- is still RUNNING: queued behind an
// inner loop that gave up part-way, they write nothing, and the host runs the
// inner solve and them again.

// sum(double(delta)) -> dmean
template <typename T>
__global__ __launch_bounds__(1024) void k_delta_mean(int n, int xcdRemap,
                                                     const T* __restrict__ delta, OState* os,
                                                     double* partials, int pstride,
                                                     const int32_t* guard) {
  if (guard && *guard == ST_RUNNING) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  double v[1] = {0.0};
  delta_mean_rows<T>(r0, r1, delta, v);
  block_partials<1>(v, partials, pstride, lb);
}

// psiNew = psiOld + (double(delta) - dmean); r = b - A psiNew with psiNew
// gathered on the fly; partials u.r, sum|r|, sum r (metalFoamSolver.C:318-365).
template <typename T>
__global__ __launch_bounds__(1024) void k_refine_residual(
    LduOp<double> A, const double* __restrict__ psiOld, const T* __restrict__ delta,
    double* __restrict__ psiNew, const double* __restrict__ b, const double* __restrict__ u,
    double* __restrict__ r, int useMean, OState* os, double* partials, int pstride,
    const int32_t* guard) {
  if (guard && *guard == ST_RUNNING) return;
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  const double dm = useMean ? os->dmean : 0.0;
  double v[3] = {0.0, 0.0, 0.0};
  for_rows(A, r0, r1, RefineRow<T>{psiOld, delta, psiNew, b, u, r, dm, v});
  block_partials<3>(v, partials, pstride, lb);
}

// Constant-mode correction psi += c, r -= c u (metalFoamSolver.C:342-361).
__global__ __launch_bounds__(1024) void k_const_shift(int n, int xcdRemap, double* __restrict__ psi,
                                                      double* __restrict__ r,
                                                      const double* __restrict__ u, OState* os,
                                                      double* partials, int pstride,
                                                      const int32_t* guard) {
  if (guard && *guard == ST_RUNNING) return;
  const int lb = logical_block(xcdRemap);
  int r0, r1;
  block_rows(n, lb, r0, r1);
  double v[2] = {0.0, 0.0};
  const_shift_rows(r0, r1, psi, r, u, os->c, v);
  block_partials<2>(v, partials, pstride, lb);
}

// Plain y = A x.
template <typename T>
__global__ __launch_bounds__(1024) void k_spmv(LduOp<T> A, const T* __restrict__ x,
                                               T* __restrict__ y) {
  const int lb = logical_block(A.xcdRemap);
  int r0, r1;
  block_rows(A.n, lb, r0, r1);
  for_rows(A, r0, r1, [&](int i, auto row) { y[i] = row(x[i], [&](int j) { return x[j]; }); });
}

// reciprocal diagonal
__global__ __launch_bounds__(1024) void k_recip(int n, const double* __restrict__ d,
                                                double* __restrict__ rD) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rD[i] = 1.0 / d[i];
}

// Build losortLow[k] = lowAddr[losort[k]] (once per mesh).
__global__ __launch_bounds__(1024) void k_gather_addr(int nf, const int32_t* __restrict__ losort,
                                                      const int32_t* __restrict__ low,
                                                      int32_t* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nf) out[k] = low[losort[k]];
}

// Row-slot coefficients (layout of LduOp::slotCoef) from the face-ordered
// upper and losort-ordered lower coefficients: a side's faces fill, in order,
// the row's non-empty slots; empty slots hold 0.  sym: the owner slots only.
// tileStart: the tiles' first rows (tile b = workgroup b, at most kSlotTile
// rows each), or nullptr for kSlotTile-row tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_build_slot_coeffs(
    int n, const int32_t* __restrict__ tileStart, const int32_t* __restrict__ ownStart,
    const int32_t* __restrict__ losortStart, const uint32_t* __restrict__ slotOff,
    const int32_t* __restrict__ slotTile, const uint8_t* __restrict__ slotRow, int sym,
    const T* __restrict__ upper, const T* __restrict__ lowerL, T* __restrict__ slotCoef) {
  int i, t;
  if (tileStart) {
    t = blockIdx.x;
    i = tileStart[t] + threadIdx.x;
    if (i >= tileStart[t + 1]) return;
  } else {
    i = blockIdx.x * blockDim.x + threadIdx.x;
    t = i / kSlotTile;
    if (i >= n) return;
  }
  const size_t nn = (size_t)n;
  const int32_t* d = slotTile + (size_t)t * kSlotDesc;
  // face tiles have no slots; symmetric tiles flagged kSlotNbrPlanes store
  // their neighbour coefficients too
  if (d[6] & kSlotFaces) return;
  const bool nbr = !sym || (d[6] & kSlotNbrPlanes);
  int o[2 * kShortRow];
  unsigned own;
  slot_decode(d, slotRow[i], slotOff, nn, i, o, own);
  int f = ownStart[i], k = losortStart[i];
#pragma unroll
  for (int q = 0; q < 2 * kShortRow; ++q) {
    T v = T(0);
    if (o[q] != 0) v = q < kShortRow ? upper[f++] : lowerL[k++];
    if (q < kShortRow || nbr) slotCoef[q * nn + i] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_copy(int n, const T* __restrict__ s, T* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// The best-iterate restore of an inner solve (metalFoamBackend.C:718-727) on
// the device: delta = deltaBest when the loop ended (status not RUNNING) and
// its best checked residual is below the final one in *st.
template <typename T>
__global__ __launch_bounds__(1024) void k_restore_best(int n, const KState* st,
                                                       T* __restrict__ delta,
                                                       const T* __restrict__ deltaBest) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && st->status != ST_RUNNING && st->bestRes < st->res) delta[i] = deltaBest[i];
}

// ---------------------------------------------------------------------------
// adjoint gradients (no reference counterpart)
// ---------------------------------------------------------------------------

// For A x = b and a loss L(x) with A^T lambda = dL/dx:
//   dL/dupper_f = -lambda[l] x[u]   (upper_f = A[l,u])
//   dL/dlower_f = -lambda[u] x[l]   (lower_f = A[u,l])
// A symmetric matrix stores one coefficient per face (upper == lower), whose
// gradient is the sum of both terms.
__global__ __launch_bounds__(256) void k_adjoint_faces(int nf, const int32_t* __restrict__ low,
                                                       const int32_t* __restrict__ upp,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ lam,
                                                       int symmetric, double* __restrict__ dUpper,
                                                       double* __restrict__ dLower) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int l = low[f], u = upp[f];
  const double gu = -lam[l] * x[u];
  const double gl = -lam[u] * x[l];
  if (symmetric) {
    if (dUpper) dUpper[f] = gu + gl;
  } else {
    if (dUpper) dUpper[f] = gu;
    if (dLower) dLower[f] = gl;
  }
}

// dL/ddiag_i = -lambda_i x_i, dL/db_i = lambda_i
__global__ __launch_bounds__(256) void k_adjoint_cells(int n, const double* __restrict__ x,
                                                       const double* __restrict__ lam,
                                                       double* __restrict__ dDiag,
                                                       double* __restrict__ dSource) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dDiag) dDiag[i] = -lam[i] * x[i];
  if (dSource) dSource[i] = lam[i];
}

__global__ __launch_bounds__(256) void k_fill(int n, double v, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

// ---------------------------------------------------------------------------
// finalisers (thread 0 of k_finalize; t = folded totals)
// ---------------------------------------------------------------------------

// rho = wA.rA, beta (reduce_scalars mode 0)
template <typename T>
struct FinBeta {
  KState* st;
  __device__ void operator()(const double* t) const { finalize_beta<T>(st, t[0]); }
};

// wApA, alpha (reduce_scalars mode 1); OpenFOAM mode: checkSingularity(|wApA|/normFactor)
template <typename T>
// Runs only after a beta (alphaDue): merged into k_cheb_first, it also sees
// the first iteration, where no p.Ap has been reduced.  Leaves the update
// pending (pend) for k_cheb_first.
struct FinAlpha {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (!st->alphaDue) return;
    st->alphaDue = 0;
    if (st->mode == MODE_INNER) {
      const T wApA = (T)t[0];
      st->wApA = wApA;
      st->alpha = (double)((T)st->rho / wApA);
    } else {
      st->wApA = t[0];
      if (fabs(t[0]) / st->normFactor < kVSmall) {
        st->status = ST_SINGULAR;
        return;
      }
      st->alpha = st->rho / t[0];
    }
    st->pend = 1;
  }
};

// sum|r| and the end-of-iteration checks
template <typename T>
struct FinPcgRes {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->mode == MODE_INNER) {
      st->res = (double)(T)t[0];
      inner_iteration_end(st, st->wApA);
    } else {
      st->res = t[0] / st->normFactor;
      openfoam_iteration_end(st);
    }
  }
};

// Chebyshev PCG: t[0] = rho (last Chebyshev step), t[1] = sum|r| of the
// update k_cheb_first applied.  The previous iteration's end-of-iteration
// checks (FinPcgRes) run first, then beta, exactly as the separate stages
// would in that order.
template <typename T>
struct FinResBeta {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->pend) {
      st->pend = 0;
      FinPcgRes<T>{st}(t + 1);
      if (st->status != ST_RUNNING) return;
    }
    finalize_beta<T>(st, t[0]);
  }
};

// rA0.rA, beta (reduce_scalars mode 3); OpenFOAM mode: singularity on rho and omega
template <typename T>
struct FinBicgRho {
  KState* st;
  __device__ void operator()(const double* t) const {
    const bool first = (st->it == 0);
    st->rA0rAold = st->rA0rA;
    if (st->mode == MODE_INNER) {
      const T rr = (T)t[0];
      st->rA0rA = rr;
      st->beta =
          first ? 0.0 : (double)((rr / (T)st->rA0rAold) * ((T)st->alpha / (T)st->omega));
    } else {
      st->rA0rA = t[0];
      if (fabs(t[0]) < kVSmall) {
        st->status = ST_SINGULAR;
      } else if (first) {
        st->beta = 0.0;
      } else if (fabs(st->omega) < kVSmall) {
        st->status = ST_SINGULAR;
      } else {
        st->beta = (t[0] / st->rA0rAold) * (st->alpha / st->omega);
      }
    }
  }
};

// rA0.AyA, alpha (reduce_scalars mode 4)
template <typename T>
struct FinBicgAlpha {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->mode == MODE_INNER) {
      const T a = (T)t[0];
      st->rA0AyA = a;
      st->alpha = (double)((T)st->rA0rA / a);
    } else {
      st->rA0AyA = t[0];
      st->alpha = st->rA0rA / t[0];
    }
  }
};

// sum|sA|: breakdown checks and the early-exit test (metalFoamBackend.C:828-866)
template <typename T>
struct FinBicgS {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->mode == MODE_INNER) {
      const double res = (double)(T)t[0];
      st->res = res;
      if (!isfinite(st->rA0rA) || !isfinite(st->alpha) || !isfinite(res))
        st->status = ST_SINGULAR;
      else if (res <= st->target)
        st->status = ST_EARLY;
    } else {
      st->res = t[0] / st->normFactor;
      if (of_converged(st, st->res)) st->status = ST_EARLY;
    }
  }
};

// tA.sA, tA.tA, omega (reduce_scalars mode 5)
template <typename T>
struct FinBicgOmega {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->mode == MODE_INNER) {
      const T tAsA = (T)t[0], tAtA = (T)t[1];
      st->tAsA = tAsA;
      st->tAtA = tAtA;
      st->omega = (double)(tAsA / tAtA);
    } else {
      st->tAsA = t[0];
      st->tAtA = t[1];
      st->omega = t[0] / t[1];
    }
  }
};

// sum|rA| and the end-of-iteration checks, or completion of an |sA| exit
template <typename T>
struct FinBicgUpdate {
  KState* st;
  __device__ void operator()(const double* t) const {
    if (st->status == ST_EARLY) {
      st->it += 1;
      st->status = ST_CONVERGED;
      return;
    }
    if (st->mode == MODE_INNER) {
      st->res = (double)(T)t[0];
      inner_iteration_end(st, st->omega);
    } else {
      st->res = t[0] / st->normFactor;
      openfoam_iteration_end(st);
    }
  }
};

// sum|sA| (breakdown and early-exit tests), then tA.sA, tA.tA and omega
// when the iteration goes on: FinBicgS followed by FinBicgOmega.
template <typename T>
struct FinBicgSOmega {
  KState* st;
  __device__ void operator()(const double* t) const {
    FinBicgS<T>{st}(t);
    if (st->status == ST_RUNNING) FinBicgOmega<T>{st}(t + 1);
  }
};

// End of an iteration (FinBicgUpdate), then rho = rA0.rA and beta of the next
// iteration (FinBicgRho) while the loop goes on.
template <typename T>
struct FinBicgUpdateRho {
  KState* st;
  __device__ void operator()(const double* t) const {
    FinBicgUpdate<T>{st}(t);
    if (st->status == ST_RUNNING) FinBicgRho<T>{st}(t + 1);
  }
};

struct FinRowsum {
  OState* os;
  int n;
  __device__ void operator()(const double* t) const {
    os->sumPsi = t[0];
    os->xbar = t[0] / (double)n;
    os->nPinned = t[1];
    os->uu = t[2];
    os->sumU = t[3];
  }
};

__device__ __forceinline__ double pinned_weight(const OState* os) {
  return os->sumU != 0.0 ? os->sumR / os->sumU : 0.0;
}

struct FinInitialResidual {
  OState* os;
  int n;
  __device__ void operator()(const double* t) const {
    os->normFactor = t[1] + kSmall;
    os->sumMagR = t[0];
    os->initRes = t[0] / os->normFactor;
    os->finalRes = os->initRes;
    os->sumR = t[2];
    os->rmean = t[2] / (double)n;
    os->rproj = pinned_weight(os);
  }
};

struct FinPrepare {
  OState* os;
  int deflate, wantBeta, n;  // deflate: the mode k_prepare_inner was given
  __device__ void operator()(const double* t) const {
    const int on = deflate_on(deflate, os, n);
    os->deflate = on;
    os->bail = t[0];
    if (wantBeta) {
      if (on)
        os->betaDefl = t[1];
      else
        os->betaExact = t[1];
    }
  }
};

struct FinDeltaMean {
  OState* os;
  int n;
  __device__ void operator()(const double* t) const { os->dmean = t[0] / (double)n; }
};

struct FinRefine {
  OState* os;
  int n;
  __device__ void operator()(const double* t) const {
    os->ur = t[0];
    os->sumMagR = t[1];
    os->finalRes = t[1] / os->normFactor;
    os->sumR = t[2];
    os->rmean = t[2] / (double)n;
    os->rproj = pinned_weight(os);
    os->c = (os->uu > kVSmall) ? t[0] / os->uu : 0.0;
  }
};

struct FinShift {
  OState* os;
  int n;
  __device__ void operator()(const double* t) const {
    os->sumMagR = t[0];
    os->finalRes = t[0] / os->normFactor;
    os->sumR = t[1];
    os->rmean = t[1] / (double)n;
    os->rproj = pinned_weight(os);
  }
};

}  // namespace hipfoam
