// hipfoam_persist.hpp -- the inner Chebyshev PCG loop of a small mesh as one
// launch on one XCD.
//
// On a small mesh (tens of 256-row tiles) each kernel of an iteration runs
// for about as long as its launch costs: an iteration is order + 2 launches
// of ~4.4 us each, nearly all of it fixed cost.  Here the whole inner loop
// runs in one launch whose K participating workgroups all sit on one XCD and
// share its L2, separated by a flag barrier instead of kernel boundaries:
//  - hardware dispatch deals workgroups to the 8 XCDs round-robin (from an
//    XCD that varies from launch to launch), so the workgroups b % 8 == 0 of
//    a grid of 8 K share one XCD; the others return at once.  Each
//    participant publishes its XCD id, and after a first barrier all check
//    that every participant reported the same one, from this launch;
//  - a store is complete in the XCD's L2 once its wave has waited on its
//    memory counter; a device-scope load bypasses the CU's L1 and reads that
//    L2, so the vectors other workgroups wrote are read with device-scope
//    loads (vld<true>), the read-only operator with plain ones;
//  - no device-scope release (which writes back the whole L2) and no
//    device-wide atomics are needed.
// Every workgroup keeps its own LDS copy of the Krylov state and folds all
// partials itself, as the merged second stages of the launched kernels do,
// so the copies stay identical and no state is exchanged.  Rows, tiles,
// partials (one per 256-row tile) and fold order are those of the per-tile
// launch shape, so the iterates are bit-identical to the launched loop.
//
// Every barrier wait is bounded (kPersistMaxPolls polls).  A participant
// that times out, or finds the participants on more than one XCD, sets the
// abort word; the others
// see it while polling, and every wave returns.  The host then resets the
// barrier words and runs the launched loop from the start.
#pragma once

#include "hipfoam_device.hpp"

namespace hipfoam {

constexpr int kPersistMaxGroups = 64;         // participants: at most 2 per CU of one XCD
constexpr long kPersistMaxPolls = 1L << 22;   // one barrier wait: about 1-2 s at most
// wave 0 of each participant polls the participants' flags one per lane
static_assert(kPersistMaxGroups <= 64, "the barrier polls one participant per lane");

// Barrier words of a persistent launch (device memory, zeroed at context
// creation and after an abandoned launch).
struct XcdSync {
  uint32_t* flags;   // [kPersistMaxGroups] arrival count of each participant
  uint32_t* xccs;    // [kPersistMaxGroups] each participant's XCD id, tagged
                     // with its launch: (first arrival count << 4) | XCD id
  uint32_t* abortw;  // nonzero: a participant gave up and the launch was abandoned;
                     // the first reason stays: kAbortOffXcd | wg << 4 | the XCD
                     // id it found elsewhere, or kAbortTimeout | the index of
                     // the barrier that timed out
  uint32_t* epoch;   // the last arrival count of the previous launch
};
constexpr int kXcdSyncWords = 2 * kPersistMaxGroups + 2;

constexpr uint32_t kAbortOffXcd = 0x10000u;
constexpr uint32_t kAbortTimeout = 0x20000u;

__device__ __forceinline__ void abandon(const XcdSync& S, uint32_t why) {
  uint32_t expected = 0u;
  __hip_atomic_compare_exchange_strong(S.abortw, &expected, why, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned xcc_id() {
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x & 0xfu;
}

// Barrier of the K participants (arrival counts only grow: target is one more
// than the previous one).  Every wave first waits until its stores are
// complete in the L2; then thread 0 publishes the arrival count and wave 0
// polls every participant's count with device-scope loads until all have
// reached target.  Returns false when the launch was abandoned.
__device__ __forceinline__ bool xcd_barrier(const XcdSync& S, int K, int w, uint32_t target,
                                            uint32_t index) {
  __shared__ int ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    if (lane == 0)
      __hip_atomic_store(S.flags + w, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    bool good = true;
    for (long polls = 1;; ++polls) {
      const uint32_t v = lane < K ? __hip_atomic_load(S.flags + lane, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : target;
      if (__all((int)(v - target) >= 0)) break;
      if (polls > kPersistMaxPolls) {
        if (lane == 0) abandon(S, kAbortTimeout | (index & 0xffffu));
        good = false;
        break;
      }
      if ((polls & 31) == 0 &&
          __hip_atomic_load(S.abortw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        good = false;
        break;
      }
    }
    if (lane == 0) ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// The state a persistent kernel keeps in LDS, and the end of the kernel.
constexpr int kStateWords = (int)(sizeof(KState) / sizeof(uint32_t));
static_assert(sizeof(KState) % sizeof(uint32_t) == 0 && kStateWords <= 256, "KState word copy");
constexpr uint32_t kNoStart = 0xffffffffu;

// Prologue of a persistent kernel (participant wg of K): copies *start into
// the LDS state, then checks that every participant runs on the same XCD --
// each reports its XCD id tagged with this launch, then one barrier, then
// every participant reads all reports.  Returns the epoch the launch's
// arrival counts start from (its first barrier used epoch + 1), or kNoStart
// when the launch was abandoned.
__device__ __forceinline__ uint32_t persistent_start(const XcdSync& S, int K, int wg,
                                                     const KState* start, KState* ks) {
  if ((int)threadIdx.x < kStateWords)
    reinterpret_cast<uint32_t*>(ks)[threadIdx.x] =
        reinterpret_cast<const uint32_t*>(start)[threadIdx.x];
  const uint32_t tgt0 = __hip_atomic_load(S.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t mine = ((tgt0 + 1) << 4) | xcc_id();
  if (threadIdx.x == 0) S.xccs[wg] = mine;
  if (!xcd_barrier(S, K, wg, tgt0 + 1, 0)) return kNoStart;
  __shared__ int same;
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const uint32_t v = lane < K ? vld<true>(S.xccs, lane) : mine;
    const bool ok = __all(v == mine);
    if (!ok) {
      const unsigned long long diff = __ballot(v != mine);
      const uint32_t other = __shfl(v, __ffsll((long long)diff) - 1);
      if (lane == 0) abandon(S, kAbortOffXcd | ((uint32_t)wg << 4) | (other & 0xfu));
    }
    if (lane == 0) same = ok;
  }
  __syncthreads();
  return same ? tgt0 : kNoStart;
}

// Every participant left the loop at the same point with the same state:
// participant 0 stores it in *kst and *kst2 and the last arrival count.
__device__ __forceinline__ void persistent_end(const XcdSync& S, int wg, const KState* ks,
                                               KState* kst, KState* kst2, uint32_t tgt) {
  if (wg != 0) return;
  if ((int)threadIdx.x < kStateWords) {
    const uint32_t word = reinterpret_cast<const uint32_t*>(ks)[threadIdx.x];
    reinterpret_cast<uint32_t*>(kst)[threadIdx.x] = word;
    reinterpret_cast<uint32_t*>(kst2)[threadIdx.x] = word;
  }
  if (threadIdx.x == 0)
    __hip_atomic_store(S.epoch, tgt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// What an abandoned launch leaves behind, written after a launch that ran to
// its end (hipfoam_simulate_abandon, for tests of the recovery): the state
// back at RUNNING and the abort word set to kAbortSimulated.
constexpr uint32_t kAbortSimulated = 0x40000u;

__global__ void k_simulate_abandon(KState* kst, uint32_t* abortw) {
  if (threadIdx.x == 0) {
    kst->status = ST_RUNNING;
    *abortw = kAbortSimulated;
  }
}

// The vectors of one inner PCG solve (the context's Work<T>).
template <typename T>
struct PcgVectors {
  T* r;
  T* r1;
  T* p0;
  T* p1;
  T* wA;
  T* Ap;
  T* delta;
  T* deltaBest;
  T* z0;
  T* z1;
  T* w;
};

// The Chebyshev PCG loop (order >= 2) of the launched kernels, iteration by
// iteration: k_cheb_first with FinAlpha merged, k_cheb_step for steps
// 2 .. order-1, k_pcg_p_spmv with FinResBeta merged, with r and p ping-
// ponging on the iteration parity, until the state leaves RUNNING or
// maxLoops iterations.  One 256-thread workgroup per participant; K
// participants take the tiles t = w, w + K, ... of every phase.  The state
// starts from *kst2 and ends in *kst and *kst2.
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_persistent(LduOp<T> A, const T* __restrict__ rD,
                                                        PcgVectors<T> V, KState* kst,
                                                        KState* kst2, double* P, double* P2,
                                                        int pstride, int order, int K,
                                                        int maxLoops, XcdSync S) {
  if (blockIdx.x & 7) return;
  const int wg = (int)(blockIdx.x >> 3);
  if (wg >= K) return;
  __shared__ KState ks;
  const uint32_t tgt0 = persistent_start(S, K, wg, kst2, &ks);
  if (tgt0 == kNoStart) return;
  uint32_t tgt = tgt0 + 1;
  const int n = A.n;
  const int ntiles = (n + kSlotTile - 1) / kSlotTile;
  for (int it = 0; it < maxLoops; ++it) {
    const bool odd = (it & 1) != 0;
    const T* rin = odd ? V.r1 : V.r;
    T* rout = odd ? V.r : V.r1;
    const T* pold = odd ? V.p1 : V.p0;
    T* pnew = odd ? V.p0 : V.p1;
    // k_cheb_first: alpha's second stage, the pending update, the first step
    {
      double t[1];
      fold_partials<1, false, true>(P2, pstride, ntiles, t);
      if (threadIdx.x == 0 && ks.status == ST_RUNNING) FinAlpha<T>{&ks}(t);
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T c1 = (T)ks.chebC1, c2 = (T)ks.chebC2[0], c3 = (T)ks.chebC3[0];
      const bool pend = ks.pend != 0;
      const T alpha = pend ? (T)ks.alpha : (T)0;
      const bool copy = pend && ks.copyBest != 0;
      T* zout = order == 2 ? V.wA : V.z0;
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        double v[2] = {0.0, 0.0};
        if (order == 2) {
          if (pend)
            for_rows(A, r0, r1,
                     ChebFirstRow<T, true, true, true>{rin, rout, rD, pold, V.Ap, V.delta,
                                                       V.deltaBest, V.w, zout, c1, c2, c3,
                                                       alpha, copy, v});
          else
            for_rows(A, r0, r1,
                     ChebFirstRow<T, true, true, false>{rin, rout, rD, pold, V.Ap, V.delta,
                                                        V.deltaBest, V.w, zout, c1, c2, c3,
                                                        alpha, copy, v});
        } else {
          if (pend)
            for_rows(A, r0, r1,
                     ChebFirstRow<T, false, true, true>{rin, rout, rD, pold, V.Ap, V.delta,
                                                        V.deltaBest, V.w, zout, c1, c2, c3,
                                                        alpha, copy, v});
          else
            for_rows(A, r0, r1,
                     ChebFirstRow<T, false, true, false>{rin, rout, rD, pold, V.Ap, V.delta,
                                                         V.deltaBest, V.w, zout, c1, c2, c3,
                                                         alpha, copy, v});
        }
        block_partials<2>(v, P, pstride, tile);
        __syncthreads();  // thread 0 has read the block sums before the next tile's
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
    }
    // k_cheb_step: steps 2 .. order-1; the last forms wA and reduces rho
    const T* zin = order == 2 ? V.wA : V.z0;
    for (int s = 2; s < order; ++s) {
      const bool last = s == order - 1;
      T* zo = last ? V.wA : ((s & 1) ? V.z0 : V.z1);
      const T c2 = (T)ks.chebC2[s - 1], c3 = (T)ks.chebC3[s - 1];
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        if (last) {
          double dot = 0.0;
          for_rows(A, r0, r1, ChebStepRow<T, true, true>{rout, rD, zin, V.w, zo, c2, c3, &dot});
          double v[1] = {dot};
          block_partials<1>(v, P, pstride, tile);
          __syncthreads();
        } else {
          for_rows(A, r0, r1,
                   ChebStepRow<T, false, true>{rout, rD, zin, V.w, zo, c2, c3, nullptr});
        }
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
      zin = zo;
    }
    // k_pcg_p_spmv: the checks and beta, then p, A p and p . Ap
    {
      double t[2];
      fold_partials<2, false, true>(P, pstride, ntiles, t);
      if (threadIdx.x == 0 && ks.status == ST_RUNNING) FinResBeta<T>{&ks}(t);
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T beta = (T)ks.beta;
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        double dot = 0.0;
        for_rows(A, r0, r1, PcgPRow<T, true>{V.wA, pold, pnew, V.Ap, beta, &dot});
        double v[1] = {dot};
        block_partials<1>(v, P2, pstride, tile);
        __syncthreads();
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
    }
  }
  persistent_end(S, wg, &ks, kst, kst2, tgt);
}


// The vectors of one inner BiCGStab solve (the context's Work<T>).
template <typename T>
struct BicgVectors {
  T* r;
  T* p0;
  T* p1;
  T* AyA;
  T* AyA1;
  T* yA;
  T* rA0;
  T* sA;
  T* zA;
  T* tA;
  T* delta;
  T* deltaBest;
};

// The BiCGStab loop of the launched kernels after its start (k_bicg_rho and
// FinBicgRho, launched before this kernel), iteration by iteration:
// k_bicg_py_spmv, k_bicg_sz_spmv with FinBicgAlpha merged, k_bicg_update_rho
// with FinBicgSOmega merged, then the end-of-iteration stage FinBicgUpdateRho
// (a separate finalise when launched, on RUNNING or EARLY states), here at
// the start of the next iteration; p and AyA ping-pong on the iteration
// parity.  The state starts from *kst and ends in *kst and *kst2.
template <typename T>
__global__ __launch_bounds__(256) void k_bicg_persistent(LduOp<T> A, const T* __restrict__ rD,
                                                         BicgVectors<T> V, KState* kst,
                                                         KState* kst2, double* P, double* P2,
                                                         int pstride, int K, int maxLoops,
                                                         XcdSync S) {
  if (blockIdx.x & 7) return;
  const int wg = (int)(blockIdx.x >> 3);
  if (wg >= K) return;
  __shared__ KState ks;
  const uint32_t tgt0 = persistent_start(S, K, wg, kst, &ks);
  if (tgt0 == kNoStart) return;
  uint32_t tgt = tgt0 + 1;
  const int n = A.n;
  const int ntiles = (n + kSlotTile - 1) / kSlotTile;
  for (int it = 0; it < maxLoops; ++it) {
    const bool odd = (it & 1) != 0;
    const T* pold = odd ? V.p1 : V.p0;
    T* pnew = odd ? V.p0 : V.p1;
    const T* AyAold = odd ? V.AyA1 : V.AyA;
    T* AyAnew = odd ? V.AyA : V.AyA1;
    // the previous iteration's end and the next rho, beta
    if (it > 0) {
      double t[2];
      fold_partials<2, false, true>(P, pstride, ntiles, t);
      if (threadIdx.x == 0 && (ks.status == ST_RUNNING || ks.status == ST_EARLY))
        FinBicgUpdateRho<T>{&ks}(t);
      __syncthreads();
    }
    if (ks.status != ST_RUNNING) break;
    // k_bicg_py_spmv
    {
      const T beta = (T)ks.beta, omega = (T)ks.omega;
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        double d = 0.0;
        for_rows(A, r0, r1,
                 BicgPyRow<T, true>{V.r, rD, pold, pnew, AyAold, AyAnew, V.yA, V.rA0, beta,
                                    omega, &d});
        double v[1] = {d};
        block_partials<1>(v, P, pstride, tile);
        __syncthreads();
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
    }
    // k_bicg_sz_spmv: alpha's second stage, then sA, zA, tA
    {
      double t[1];
      fold_partials<1, false, true>(P, pstride, ntiles, t);
      if (threadIdx.x == 0 && ks.status == ST_RUNNING) FinBicgAlpha<T>{&ks}(t);
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T alpha = (T)ks.alpha;
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        double d[3] = {0.0, 0.0, 0.0};
        for_rows(A, r0, r1, BicgSzRow<T, true>{V.r, rD, AyAnew, V.sA, V.zA, V.tA, alpha, d});
        block_partials<3>(d, P2, pstride, tile);
        __syncthreads();
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
    }
    // k_bicg_update_rho: the |sA| test and omega, then delta and rA
    {
      double t[3];
      fold_partials<3, false, true>(P2, pstride, ntiles, t);
      if (threadIdx.x == 0 && ks.status == ST_RUNNING) FinBicgSOmega<T>{&ks}(t);
      __syncthreads();
      const int status = ks.status;
      if (status != ST_RUNNING && status != ST_EARLY) break;
      const T alpha = (T)ks.alpha;
      for (int tile = wg; tile < ntiles; tile += K) {
        const int r0 = tile * kSlotTile, r1 = min(r0 + kSlotTile, n);
        if (status == ST_EARLY) {
          bicg_early_rows<T, true>(r0, r1, V.delta, V.yA, alpha);
        } else {
          double v[2] = {0.0, 0.0};
          bicg_update_rows<T, true>(r0, r1, V.delta, V.deltaBest, V.yA, V.zA, V.r, V.sA, V.tA,
                                    V.rA0, alpha, (T)ks.omega, ks.copyBest != 0, v);
          block_partials<2>(v, P, pstride, tile);
          __syncthreads();
        }
      }
      ++tgt;
      if (!xcd_barrier(S, K, wg, tgt, tgt - tgt0 - 1)) return;
    }
  }
  persistent_end(S, wg, &ks, kst, kst2, tgt);
}

}  // namespace hipfoam
