// hipfoam_local.hpp -- the inner loop of a mesh of at most 2,816 rows in one
// workgroup: every vector of the solve in LDS, every thread's matrix rows in
// registers.
//
// Below a dozen 256-row tiles the one-XCD persistent loop still spends each
// phase on a flag barrier across workgroups and a round of L2 loads of the
// rows and the vectors (about 1.8 us a phase).  Here one workgroup of 512
// threads runs the whole inner loop on one CU:
//  - the twelve vectors of the solve sit in LDS (at most 3,072 FP32 rows,
//    144 KiB, or 1,536 FP64 rows, 144 KiB, of the CU's 160 KiB), so the
//    gathers of a phase are LDS loads;
//  - each thread holds the rows it owns (row p * 512 + thread in pass p) in
//    registers from the first phase to the last, so no phase reads the
//    operator;
//  - phases are separated by workgroup barriers, and the reductions end in
//    wave 0: it forms the tile partials from the per-wave sums, keeps them
//    in its lanes, folds them and applies the finaliser, one barrier per
//    reduction instead of three.
// Row p * 512 + t is row t % 256 of tile 2 p + t / 256.  Each tile's partial
// sum is formed as a 256-thread workgroup of the per-tile launch shape forms
// it (block_partials: wave sums, then the four waves in order), and the fold
// over at most 64 tile partials is fold_partials' tree.  Rows, partials and
// fold order are those of the launched loop, so the iterates are
// bit-identical to it.  Only meshes whose tiles are all row-slot tiles
// (patterned or explicit, no face tiles) qualify.  Measured against the
// one-XCD loop (2D Poisson, refined solve, same box): 5.7 against 15.8 us
// per iteration at 400 rows, 9.7 against 15.8 at 1,536, 11.8 against 15.8 at
// 2,048, 13.4 against 15.9 at 2,304 (six passes); plain FP64 4.9 against
// 16.9 at 400 and 9.7 against 17.3 at 1,536.
#pragma once

#include "hipfoam_persist.hpp"

namespace hipfoam {

constexpr int kLocalThreads = 512;
constexpr int kLocalTilesPerPass = kLocalThreads / kSlotTile;
constexpr int kLocalVectors = 12;  // the 11 of PcgVectors and rD
constexpr int kLocalMaxTiles = 12;

// Rows of the one-workgroup PCG loop: 2,816 FP32 rows or 1,536 FP64 rows.
// k_pcg_local<T, NP> holds NP passes of kLocalThreads rows (NP * 24 KiB of
// FP32 vectors in LDS): FP32 meshes up to 2,048 rows run the four-pass
// build, larger ones the six-pass build, which is 0.7-0.9 us per iteration
// slower on the small meshes (more unrolled code, more registers) but beats
// the one-XCD loop up to about 2,800 rows (2,500 rows: 14.2 against 15.8 us
// per iteration; 3,072 rows: 15.7 against 15.5).
template <typename T>
constexpr int local_max_rows() {
  return sizeof(T) == sizeof(float) ? 2816 : 1536;
}
// passes of the build that runs a mesh of n rows
template <typename T>
inline int local_pcg_passes(int n) {
  return sizeof(T) == sizeof(float) ? (n <= 4 * kLocalThreads ? 4 : 6) : 3;
}
static_assert(6 * kLocalThreads <= kLocalMaxTiles * kSlotTile, "tile partials of one workgroup");
static_assert(kSlotTile % 64 == 0 && kLocalThreads % kSlotTile == 0, "whole tiles per pass");
static_assert(kLocalMaxTiles <= 64, "wave 0 holds one tile partial per lane");
static_assert(kLocalVectors * 6 * kLocalThreads * sizeof(float) <= 144 * 1024 &&
                  kLocalVectors * 3 * kLocalThreads * sizeof(double) <= 144 * 1024,
              "vectors within 144 KiB of LDS");

// This thread's row of each pass, in registers for the whole launch.
template <typename T, int NP>
struct LocalRows {
  RowSlots<T> rs[NP];
  T dg[NP];
  __device__ __forceinline__ void load(const LduOp<T>& A, int passes) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int i = p * kLocalThreads + (int)threadIdx.x;
      if (p < passes && i < A.n) {
        rs[p].load(A, i, i / kSlotTile);
        dg[p] = A.diag[i];
      }
    }
  }
};

// The partial sums of the tiles of every pass, left in wave 0: lane t of
// wave 0 gets tp[j] = tile t's sum of v[.][j] as block_partials forms it in
// a 256-thread workgroup (wave sums, then the tile's waves in order), and
// stores it in part[j * kLocalMaxTiles + t] (lanes past the last tile: 0).
// One barrier, after the wave sums.
template <int NV, int NP>
__device__ __forceinline__ void local_tile_partials(const double (&v)[NP][NV], double (&tp)[NV],
                                                    double* part, int passes, int ntiles) {
  constexpr int W = kLocalThreads / 64;  // waves of the workgroup
  constexpr int WPT = kSlotTile / 64;    // waves per tile
  __shared__ double sh[NP][NV][W];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p)
    if (p < passes)
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const double t = wave_sum(v[p][j]);
        if (lane == 0) sh[p][j][wid] = t;
      }
  __syncthreads();
  if (wid == 0) {
    const int p = lane / kLocalTilesPerPass, q = lane % kLocalTilesPerPass;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double t = 0.0;
      if (lane < ntiles) {
        t = sh[p][j][WPT * q];
        for (int w = 1; w < WPT; ++w) t = t + sh[p][j][WPT * q + w];
        part[j * kLocalMaxTiles + lane] = t;
      }
      tp[j] = t;
    }
  }
}

// fold_partials over the tile partials tp that wave 0's lanes hold (lane t:
// tile t, zero past the last tile), in wave 0: the same tree.  With at most
// 64 tiles fold_partials' per-thread sums are 0 + the tile's partial (its
// further terms are zeros) and its other three waves sum to +0, which the
// last three additions keep.
template <int NV>
__device__ __forceinline__ void local_fold(const double (&tp)[NV], double (&out)[NV]) {
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double v = 0.0 + tp[j];
    double t = wave_sum(v);
    t = t + 0.0;
    t = t + 0.0;
    t = t + 0.0;
    out[j] = t;
  }
}

// One phase: fn(i, row) for this thread's row of every pass (make(v) builds
// the row functor over the partial sums v of that pass), then either the
// tile partials into wave 0 (NV > 0) or a barrier.
template <int NV, typename T, int NP, typename Make>
__device__ __forceinline__ void local_phase(const LocalRows<T, NP>& R, int n, int passes,
                                            int ntiles, double (&tp)[NV > 0 ? NV : 1],
                                            double* part, Make make) {
  constexpr int NVX = NV > 0 ? NV : 1;
  double v[NP][NVX];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
#pragma unroll
    for (int j = 0; j < NVX; ++j) v[p][j] = 0.0;
    const int i = p * kLocalThreads + (int)threadIdx.x;
    if (p < passes && i < n) {
      const auto fn = make(v[p]);
      fn(i, [&](T xi, auto xf) { return R.rs[p].apply_d(R.dg[p], xi, xf); });
    }
  }
  if constexpr (NV > 0)
    local_tile_partials<NV, NP>(v, tp, part, passes, ntiles);
  else
    __syncthreads();
}

// The Chebyshev PCG loop (order >= 2) of k_pcg_persistent -- k_cheb_first
// with FinAlpha merged, k_cheb_step for steps 2 .. order-1, k_pcg_p_spmv with
// FinResBeta merged, r and p ping-ponging on the iteration parity -- in one
// workgroup of kLocalThreads threads, for a mesh of at most NP * kLocalThreads
// rows in row-slot tiles only.  The vectors G and rD, the
// partials P and P2 and the state *kst2 are loaded at the start (vectors and
// state into LDS, the partials also into wave 0's lanes); the vectors and
// partials are stored back at the end and the state into *kst and *kst2.
template <typename T, int NP>
__global__ __launch_bounds__(kLocalThreads) void k_pcg_local(LduOp<T> A,
                                                             const T* __restrict__ rDg,
                                                             PcgVectors<T> G, KState* kst,
                                                             KState* kst2, double* P, double* P2,
                                                             int pstride, int order,
                                                             int maxLoops) {
  constexpr int CAP = NP * kLocalThreads;
  __shared__ T lv[kLocalVectors * CAP];
  __shared__ double lp[2 * kLocalMaxTiles];
  __shared__ double lp2[kLocalMaxTiles];
  __shared__ KState ks;
  const int n = A.n, tid = (int)threadIdx.x;
  const int ntiles = (n + kSlotTile - 1) / kSlotTile;
  const int passes = (n + kLocalThreads - 1) / kLocalThreads;
  T* const gv[kLocalVectors] = {G.r,     G.r1,        G.p0, G.p1, G.wA, G.Ap,
                                G.delta, G.deltaBest, G.z0, G.z1, G.w,  const_cast<T*>(rDg)};
#pragma unroll
  for (int k = 0; k < kLocalVectors; ++k)
    for (int i = tid; i < n; i += kLocalThreads) lv[k * CAP + i] = gv[k][i];
  // wave 0 keeps the tile partials of P (two values) and P2 in its lanes
  double tpP[2] = {0.0, 0.0}, tpP2[1] = {0.0}, tpX[1] = {0.0};
  if (tid < ntiles) {
    tpP[0] = lp[tid] = P[tid];
    tpP[1] = lp[kLocalMaxTiles + tid] = P[pstride + tid];
    tpP2[0] = lp2[tid] = P2[tid];
  }
  if (tid < kStateWords)
    reinterpret_cast<uint32_t*>(&ks)[tid] = reinterpret_cast<const uint32_t*>(kst2)[tid];
  LocalRows<T, NP> R;
  R.load(A, passes);
  __syncthreads();
  const PcgVectors<T> V{lv,           lv + CAP,     lv + 2 * CAP, lv + 3 * CAP,
                        lv + 4 * CAP, lv + 5 * CAP, lv + 6 * CAP, lv + 7 * CAP,
                        lv + 8 * CAP, lv + 9 * CAP, lv + 10 * CAP};
  const T* rD = lv + 11 * CAP;
  for (int it = 0; it < maxLoops; ++it) {
    const bool odd = (it & 1) != 0;
    const T* rin = odd ? V.r1 : V.r;
    T* rout = odd ? V.r : V.r1;
    const T* pold = odd ? V.p1 : V.p0;
    T* pnew = odd ? V.p0 : V.p1;
    // k_cheb_first: alpha's second stage, the pending update, the first step
    {
      if (tid < 64) {
        double t[1];
        local_fold<1>(tpP2, t);
        if (tid == 0 && ks.status == ST_RUNNING) FinAlpha<T>{&ks}(t);
      }
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T c1 = (T)ks.chebC1, c2 = (T)ks.chebC2[0], c3 = (T)ks.chebC3[0];
      const bool pend = ks.pend != 0;
      const T alpha = pend ? (T)ks.alpha : (T)0;
      const bool copy = pend && ks.copyBest != 0;
      T* zout = order == 2 ? V.wA : V.z0;
      if (order == 2) {
        if (pend)
          local_phase<2>(R, n, passes, ntiles, tpP, lp, [&](double* v) {
            return ChebFirstRow<T, true, false, true>{rin,     rout,        rD,  pold, V.Ap,
                                                      V.delta, V.deltaBest, V.w, zout, c1,
                                                      c2,      c3,          alpha, copy, v};
          });
        else
          local_phase<2>(R, n, passes, ntiles, tpP, lp, [&](double* v) {
            return ChebFirstRow<T, true, false, false>{rin,     rout,        rD,  pold, V.Ap,
                                                       V.delta, V.deltaBest, V.w, zout, c1,
                                                       c2,      c3,          alpha, copy, v};
          });
      } else {
        if (pend)
          local_phase<2>(R, n, passes, ntiles, tpP, lp, [&](double* v) {
            return ChebFirstRow<T, false, false, true>{rin,     rout,        rD,  pold, V.Ap,
                                                       V.delta, V.deltaBest, V.w, zout, c1,
                                                       c2,      c3,          alpha, copy, v};
          });
        else
          local_phase<2>(R, n, passes, ntiles, tpP, lp, [&](double* v) {
            return ChebFirstRow<T, false, false, false>{rin,     rout,        rD,  pold, V.Ap,
                                                        V.delta, V.deltaBest, V.w, zout, c1,
                                                        c2,      c3,          alpha, copy, v};
          });
      }
    }
    // k_cheb_step: steps 2 .. order-1; the last forms wA and reduces rho
    const T* zin = order == 2 ? V.wA : V.z0;
    for (int s = 2; s < order; ++s) {
      const bool last = s == order - 1;
      T* zo = last ? V.wA : ((s & 1) ? V.z0 : V.z1);
      const T c2 = (T)ks.chebC2[s - 1], c3 = (T)ks.chebC3[s - 1];
      if (last) {
        double tr[1] = {0.0};
        local_phase<1>(R, n, passes, ntiles, tr, lp, [&](double* v) {
          return ChebStepRow<T, true, false>{rout, rD, zin, V.w, zo, c2, c3, v};
        });
        tpP[0] = tr[0];
      } else {
        local_phase<0>(R, n, passes, ntiles, tpX, lp, [&](double*) {
          return ChebStepRow<T, false, false>{rout, rD, zin, V.w, zo, c2, c3, nullptr};
        });
      }
      zin = zo;
    }
    // k_pcg_p_spmv: the checks and beta, then p, A p and p . Ap
    {
      if (tid < 64) {
        double t[2];
        local_fold<2>(tpP, t);
        if (tid == 0 && ks.status == ST_RUNNING) FinResBeta<T>{&ks}(t);
      }
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T beta = (T)ks.beta;
      local_phase<1>(R, n, passes, ntiles, tpP2, lp2, [&](double* v) {
        return PcgPRow<T, false>{V.wA, pold, pnew, V.Ap, beta, v};
      });
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kLocalVectors - 1; ++k)
    for (int i = tid; i < n; i += kLocalThreads) gv[k][i] = lv[k * CAP + i];
  if (tid < ntiles) {
    P[tid] = lp[tid];
    P[pstride + tid] = lp[kLocalMaxTiles + tid];
    P2[tid] = lp2[tid];
  }
  if (tid < kStateWords) {
    const uint32_t word = reinterpret_cast<const uint32_t*>(&ks)[tid];
    reinterpret_cast<uint32_t*>(kst)[tid] = word;
    reinterpret_cast<uint32_t*>(kst2)[tid] = word;
  }
}

}  // namespace hipfoam

namespace hipfoam {

// Rows of the one-workgroup BiCGStab loop: 1,536 FP32 rows and 1,280 FP64
// rows, three passes of 512 (thirteen vectors, the twelve of BicgVectors and
// rD: 78 KiB FP32, 130 KiB FP64).  A one-XCD BiCGStab iteration is only three phases (9.1-9.4 us at
// 400 to 2,048 rows), so one CU stops paying earlier than for PCG: measured
// against it (refined solve, 2D convection-diffusion, same box), 4.2 against
// 9.1 us per iteration at 400 rows and 7.2 against 9.4 at 1,280; plain FP64
// 4.5 against 9.8 and 8.0 against 10.1; a version holding 2,048 FP32 rows
// took 11.2 against 9.3 us at 2,025 and 2,048 rows.
constexpr int kLocalBicgVectors = 13;
template <typename T>
constexpr int local_bicg_max_rows() {
  return sizeof(T) == sizeof(float) ? 1536 : 1280;
}
template <typename T>
constexpr int local_bicg_passes() {
  return (local_bicg_max_rows<T>() + kLocalThreads - 1) / kLocalThreads;
}
static_assert(kLocalBicgVectors * 1280 * sizeof(double) <= 144 * 1024,
              "BiCGStab vectors within 144 KiB of LDS");

// The BiCGStab loop of k_bicg_persistent after its start (k_bicg_rho and
// FinBicgRho, launched before this kernel) in one workgroup of
// kLocalThreads threads, for a mesh of at most local_bicg_max_rows<T>() rows
// in row-slot tiles only: k_bicg_py_spmv, k_bicg_sz_spmv with FinBicgAlpha,
// k_bicg_update_rho with FinBicgSOmega, and FinBicgUpdateRho at the start of
// the next iteration; p and AyA ping-pong on the iteration parity.  Wave 0's
// lanes mirror the partial buffers P (two values a tile) and P2 (three), so
// every fold sees what k_bicg_persistent's folds see.  The vectors, the
// partials and the state *kst are loaded at the start; the vectors and
// partials are stored back at the end and the state into *kst and *kst2.
template <typename T>
__global__ __launch_bounds__(kLocalThreads) void k_bicg_local(LduOp<T> A,
                                                              const T* __restrict__ rDg,
                                                              BicgVectors<T> G, KState* kst,
                                                              KState* kst2, double* P,
                                                              double* P2, int pstride,
                                                              int maxLoops) {
  constexpr int CAP = local_bicg_max_rows<T>(), NP = local_bicg_passes<T>();
  __shared__ T lv[kLocalBicgVectors * CAP];
  __shared__ double lp[2 * kLocalMaxTiles];
  __shared__ double lp2[3 * kLocalMaxTiles];
  __shared__ KState ks;
  const int n = A.n, tid = (int)threadIdx.x;
  const int ntiles = (n + kSlotTile - 1) / kSlotTile;
  const int passes = (n + kLocalThreads - 1) / kLocalThreads;
  T* const gv[kLocalBicgVectors] = {G.r,  G.p0, G.p1, G.AyA,   G.AyA1,      G.yA,
                                    G.rA0, G.sA, G.zA, G.tA, G.delta, G.deltaBest,
                                    const_cast<T*>(rDg)};
#pragma unroll
  for (int k = 0; k < kLocalBicgVectors; ++k)
    for (int i = tid; i < n; i += kLocalThreads) lv[k * CAP + i] = gv[k][i];
  double tpP[2] = {0.0, 0.0}, tpP2[3] = {0.0, 0.0, 0.0};
  if (tid < ntiles) {
#pragma unroll
    for (int j = 0; j < 2; ++j) tpP[j] = lp[j * kLocalMaxTiles + tid] = P[j * pstride + tid];
#pragma unroll
    for (int j = 0; j < 3; ++j)
      tpP2[j] = lp2[j * kLocalMaxTiles + tid] = P2[j * pstride + tid];
  }
  if (tid < kStateWords)
    reinterpret_cast<uint32_t*>(&ks)[tid] = reinterpret_cast<const uint32_t*>(kst)[tid];
  LocalRows<T, NP> R;
  R.load(A, passes);
  __syncthreads();
  const BicgVectors<T> V{lv,           lv + CAP,     lv + 2 * CAP, lv + 3 * CAP,
                         lv + 4 * CAP, lv + 5 * CAP, lv + 6 * CAP, lv + 7 * CAP,
                         lv + 8 * CAP, lv + 9 * CAP, lv + 10 * CAP, lv + 11 * CAP};
  const T* rD = lv + 12 * CAP;
  for (int it = 0; it < maxLoops; ++it) {
    const bool odd = (it & 1) != 0;
    const T* pold = odd ? V.p1 : V.p0;
    T* pnew = odd ? V.p0 : V.p1;
    const T* AyAold = odd ? V.AyA1 : V.AyA;
    T* AyAnew = odd ? V.AyA : V.AyA1;
    // the previous iteration's end and the next rho, beta
    if (it > 0) {
      if (tid < 64) {
        double t[2];
        local_fold<2>(tpP, t);
        if (tid == 0 && (ks.status == ST_RUNNING || ks.status == ST_EARLY))
          FinBicgUpdateRho<T>{&ks}(t);
      }
      __syncthreads();
    }
    if (ks.status != ST_RUNNING) break;
    // k_bicg_py_spmv
    {
      const T beta = (T)ks.beta, omega = (T)ks.omega;
      double tp[1] = {0.0};
      local_phase<1>(R, n, passes, ntiles, tp, lp, [&](double* v) {
        return BicgPyRow<T, false>{V.r, rD, pold, pnew, AyAold, AyAnew, V.yA, V.rA0, beta,
                                   omega, v};
      });
      tpP[0] = tp[0];
    }
    // k_bicg_sz_spmv: alpha's second stage, then sA, zA, tA
    {
      if (tid < 64) {
        double t[1];
        const double tp[1] = {tpP[0]};
        local_fold<1>(tp, t);
        if (tid == 0 && ks.status == ST_RUNNING) FinBicgAlpha<T>{&ks}(t);
      }
      __syncthreads();
      if (ks.status != ST_RUNNING) break;
      const T alpha = (T)ks.alpha;
      local_phase<3>(R, n, passes, ntiles, tpP2, lp2, [&](double* v) {
        return BicgSzRow<T, false>{V.r, rD, AyAnew, V.sA, V.zA, V.tA, alpha, v};
      });
    }
    // k_bicg_update_rho: the |sA| test and omega, then delta and rA
    {
      if (tid < 64) {
        double t[3];
        local_fold<3>(tpP2, t);
        if (tid == 0 && ks.status == ST_RUNNING) FinBicgSOmega<T>{&ks}(t);
      }
      __syncthreads();
      const int status = ks.status;
      if (status != ST_RUNNING && status != ST_EARLY) break;
      const T alpha = (T)ks.alpha;
      if (status == ST_EARLY) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const int r0 = p * kLocalThreads;
          if (p < passes) bicg_early_rows<T, false>(r0, min(r0 + kLocalThreads, n), V.delta,
                                                    V.yA, alpha);
        }
        __syncthreads();
      } else {
        const T omega = (T)ks.omega;
        const bool copy = ks.copyBest != 0;
        double v[NP][2];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          v[p][0] = v[p][1] = 0.0;
          const int r0 = p * kLocalThreads;
          if (p < passes)
            bicg_update_rows<T, false>(r0, min(r0 + kLocalThreads, n), V.delta, V.deltaBest,
                                       V.yA, V.zA, V.r, V.sA, V.tA, V.rA0, alpha, omega, copy,
                                       v[p]);
        }
        local_tile_partials<2, NP>(v, tpP, lp, passes, ntiles);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kLocalBicgVectors - 1; ++k)
    for (int i = tid; i < n; i += kLocalThreads) gv[k][i] = lv[k * CAP + i];
  if (tid < ntiles) {
#pragma unroll
    for (int j = 0; j < 2; ++j) P[j * pstride + tid] = lp[j * kLocalMaxTiles + tid];
#pragma unroll
    for (int j = 0; j < 3; ++j) P2[j * pstride + tid] = lp2[j * kLocalMaxTiles + tid];
  }
  if (tid < kStateWords) {
    const uint32_t word = reinterpret_cast<const uint32_t*>(&ks)[tid];
    reinterpret_cast<uint32_t*>(kst)[tid] = word;
    reinterpret_cast<uint32_t*>(kst2)[tid] = word;
  }
}

}  // namespace hipfoam
