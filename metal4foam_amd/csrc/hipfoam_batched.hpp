// hipfoam_batched.hpp -- batched solves: many independent LDU systems, each
// with several right-hand sides, advanced together by one launch per step.
//
// Layout.  The systems are concatenated into one block-diagonal "super-mesh"
// (global row and face numbering, addressing rebased once at set-up), so the
// LduOp of the single-system path applies unchanged.  Right-hand sides are
// stored [rhs][row] (stride N = total rows); the off-diagonal coefficients
// are shared by every RHS of a system.  The FP32 diagonal and rD come in two
// planes over the super-mesh, exact and deflated (diag - A 1); each segment
// selects one (segPlane), so a segment may switch to the exact operator on
// its own while the RHS of a system still read one diagonal.
//
// Launch shape.  Rows are cut into 256-row tiles that never straddle a
// system.  The Krylov iteration kernels run one workgroup per (tile, group of
// RHS) and apply each row's coefficients, loaded once, to every RHS of the
// group; the FP64 refinement kernels run one workgroup per (tile, rhs),
// logical index lb = tile * nRhs + rhs, so the workgroups reading the same
// matrix rows are logically adjacent and, with the XCD remap, share an XCD's
// L2.
//
// Segments.  seg = system * nRhs + rhs owns a KState / OState.  A workgroup of
// a segment that has stopped returns at entry.  Reductions write one partial
// per workgroup; k_seg_finalize runs one workgroup per segment and folds that
// segment's partials (workgroups tile*nRhs + rhs over the system's tiles) in
// a fixed order.
#pragma once

#include "hipfoam_device.hpp"

namespace hipfoam {

struct BatchMap {
  const int32_t* __restrict__ tileRow0;  // [ntiles + 1] first global row of each tile
  const int32_t* __restrict__ tileSys;   // [ntiles] system of each tile
  const int32_t* __restrict__ sysTile0;  // [nsys + 1] first tile of each system
  const int32_t* __restrict__ segPlane;  // [nseg] diagonal plane: 0 exact, 1 deflated
  int32_t ntiles, nrhs, N, xcdRemap;
  int32_t group;  // RHS per workgroup of the Krylov iteration kernels, 1..kRhsGroup
};

struct BlockSeg {
  int t, r0, r1, rhs, seg;
  long voff;  // rhs * N: offset of this RHS in the vectors
};

__device__ __forceinline__ BlockSeg block_seg(const BatchMap& B, int lb) {
  BlockSeg s;
  const int t = lb / B.nrhs;
  s.t = t;
  s.rhs = lb - t * B.nrhs;
  s.r0 = B.tileRow0[t];
  s.r1 = B.tileRow0[t + 1];
  s.seg = B.tileSys[t] * B.nrhs + s.rhs;
  s.voff = (long)s.rhs * B.N;
  return s;
}

// Offset of a segment's diagonal plane in the FP32 diag / rD arrays
__device__ __forceinline__ long diag_off(const BatchMap& B, int seg) {
  return (long)B.segPlane[seg] * B.N;
}

// Per-segment FP32 diagonal view: the operator with the segment's diag plane.
template <typename T>
__device__ __forceinline__ LduOp<T> seg_op(LduOp<T> A, long doff) {
  A.diag += doff;
  return A;
}

// ---------------------------------------------------------------------------
// Krylov iteration kernels: one workgroup per (tile, group of up to
// B.group <= kRhsGroup right-hand sides).  Each thread owns one row: it loads the row's
// matrix slots once and applies them to every RHS of the group whose segment
// is still running, one RHS after another.  Each RHS's values are reduced
// across the wavefront at once into LDS, so no per-RHS accumulator stays
// live; the partials keep the one-per-(tile, rhs) layout
// partials[j*pstride + tile*nRhs + rhs] that k_seg_finalize folds, summed in
// the order of block_partials (wavefront butterflies, then the waves).
// ---------------------------------------------------------------------------

constexpr int kRhsGroup = 8;

struct TileGroup {
  int t, sys, r0, r1, rb, nr;  // tile, system, rows [r0, r1), RHS [rb, rb + nr)
};

__device__ __forceinline__ TileGroup tile_group(const BatchMap& B, int lb) {
  const int ng = (B.nrhs + B.group - 1) / B.group;
  TileGroup u;
  u.t = lb / ng;
  u.rb = (lb - u.t * ng) * B.group;
  u.nr = B.nrhs - u.rb < B.group ? B.nrhs - u.rb : B.group;
  u.r0 = B.tileRow0[u.t];
  u.r1 = B.tileRow0[u.t + 1];
  u.sys = B.tileSys[u.t];
  return u;
}

// bit q set: the segment of RHS rb + q has a status in `mask`
__device__ __forceinline__ unsigned status_bits(const KState* stBase, const TileGroup& u, int nrhs,
                                                int mask) {
  unsigned run = 0;
  for (int q = 0; q < u.nr; ++q)
    if ((mask >> stBase[u.sys * nrhs + u.rb + q].status) & 1) run |= 1u << q;
  return run;
}

// Per-RHS partial sums of NV values: add() reduces one RHS's values across
// the wavefront into LDS; write() (after every add, called by all threads)
// folds the waves in order and stores the partials of the RHS in `run`.
template <int NV>
struct RhsPartials {
  double (*sh)[16];  // [kRhsGroup * NV][16] in LDS
  __device__ __forceinline__ void add(int q, const double (&v)[NV]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const double t = wave_sum(v[j]);
      if (lane == 0) sh[q * NV + j][wid] = t;
    }
  }
  __device__ __forceinline__ void write(double* partials, int pstride, int base, unsigned run) {
    __syncthreads();
    const int k = threadIdx.x;
    if (k < kRhsGroup * NV) {
      const int q = k / NV, j = k - q * NV;
      if ((run >> q) & 1) {
        const int nw = (blockDim.x + 63) >> 6;
        double t = sh[k][0];
        for (int w = 1; w < nw; ++w) t += sh[k][w];
        partials[j * pstride + base + q] = t;
      }
    }
  }
};

#define HIPFOAM_RHS_PARTIALS(NV, name)              \
  __shared__ double name##_sh[kRhsGroup * NV][16]; \
  RhsPartials<NV> name{name##_sh}

// Row i of A x for one RHS: the slots loaded once per thread when the
// operator uses them, the face loops otherwise.
template <typename XF>
__device__ __forceinline__ float batch_row(const LduOp<float>& A, const RowSlots<float>& rs,
                                           long doff, int i, float xi, XF xf) {
  if (A.slots && !rs.faces) return rs.apply(A.diag + doff, i, xi, xf);
  return ldu_row_faces(seg_op(A, doff), i, xi, xf);
}

// y = A x for every RHS, with each segment's diagonal plane: the timing
// kernel of the batch autotuner (same rows, slots and RHS groups as the
// iteration kernels)
__global__ __launch_bounds__(256) void kb_spmv(BatchMap B, LduOp<float> A,
                                               const float* __restrict__ x,
                                               float* __restrict__ y) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const int i = u.r0 + threadIdx.x;
  if (i >= u.r1) return;
  RowSlots<float> rs;
  if (A.slots) rs.load(A, i, u.t);
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    const int seg = u.sys * B.nrhs + u.rb + q;
    const long voff = (long)(u.rb + q) * B.N, doff = diag_off(B, seg);
    const float* xx = x + voff;
    y[voff + i] = batch_row(A, rs, doff, i, xx[i], [&](int j) { return xx[j]; });
  }
}

// ---------------------------------------------------------------------------
// PCG (metalFoamKernels.metal:112-276 per segment)
// ---------------------------------------------------------------------------

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void kb_cheb_step(BatchMap B, LduOp<float> A,
                                                    const float* __restrict__ r,
                                                    const float* __restrict__ rD,
                                                    const float* __restrict__ zin,
                                                    float* __restrict__ w, float* __restrict__ zout,
                                                    KState* stBase, double* partials, int pstride,
                                                    int step) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(1, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
  RowSlots<float> rs;
  if (in && A.slots) rs.load(A, i, u.t);
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    double v[1] = {0.0};
    if (in) {
      const int seg = u.sys * B.nrhs + u.rb + q;
      const KState* st = stBase + seg;
      const long voff = (long)(u.rb + q) * B.N, doff = diag_off(B, seg);
      const float* rr = r + voff;
      const float* dd = rD + doff;
      const float c2 = (float)st->chebC2[step - 1];
      const float c3 = (float)st->chebC3[step - 1];
      const float c1 = (float)st->chebC1;
      // the row's own operands load before the gathers, in one batch
      const float ri = rr[i], di = dd[i];
      float zi_new, wi;
      if (FIRST) {
        const float z0 = c1 * di * ri;
        const float y = batch_row(A, rs, doff, i, z0, [&](int j) { return c1 * dd[j] * rr[j]; });
        wi = c2 * z0 + c3 * di * (ri - y);
        zi_new = z0 + wi;
      } else {
        const float* zi_ = zin + voff;
        const float zi = zi_[i], wold = w[voff + i];
        const float y = batch_row(A, rs, doff, i, zi, [&](int j) { return zi_[j]; });
        wi = c2 * wold + c3 * di * (ri - y);
        zi_new = zi + wi;
      }
      if (!LAST) w[voff + i] = wi;
      zout[voff + i] = zi_new;
      if (LAST) v[0] = (double)(zi_new * ri);
    }
    if (LAST) part.add(q, v);
  }
  if (LAST) part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

__global__ __launch_bounds__(256) void kb_pcg_p_spmv(BatchMap B, LduOp<float> A,
                                                     const float* __restrict__ wA,
                                                     const float* __restrict__ pold,
                                                     float* __restrict__ pnew,
                                                     float* __restrict__ Ap, KState* stBase,
                                                     double* partials, int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(1, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
  RowSlots<float> rs;
  if (in && A.slots) rs.load(A, i, u.t);
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    double v[1] = {0.0};
    if (in) {
      const int seg = u.sys * B.nrhs + u.rb + q;
      const long voff = (long)(u.rb + q) * B.N, doff = diag_off(B, seg);
      const float beta = (float)stBase[seg].beta;
      const float* w_ = wA + voff;
      const float* po = pold + voff;
      const float pi = w_[i] + beta * po[i];
      const float y = batch_row(A, rs, doff, i, pi, [&](int j) { return w_[j] + beta * po[j]; });
      pnew[voff + i] = pi;
      Ap[voff + i] = y;
      v[0] = (double)(pi * y);
    }
    part.add(q, v);
  }
  part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

__global__ __launch_bounds__(256) void kb_pcg_update(BatchMap B, float* __restrict__ delta,
                                                     float* __restrict__ deltaBest,
                                                     float* __restrict__ r,
                                                     const float* __restrict__ p,
                                                     const float* __restrict__ Ap,
                                                     KState* stBase, double* partials,
                                                     int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(1, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    double v[1] = {0.0};
    if (in) {
      const KState* st = stBase + u.sys * B.nrhs + u.rb + q;
      const long k = (long)(u.rb + q) * B.N + i;
      const float alpha = (float)st->alpha;
      const float d = delta[k];
      if (st->copyBest != 0) deltaBest[k] = d;
      delta[k] = d + alpha * p[k];
      const float ri = r[k] - alpha * Ap[k];
      r[k] = ri;
      v[0] = (double)fabsf(ri);
    }
    part.add(q, v);
  }
  part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

// ---------------------------------------------------------------------------
// BiCGStab (metalFoamKernels.metal:279-458 per segment): p and y, A y,
// s and z, A z, then the update, which also reduces the next iteration's rho
// (p and AyA ping-pong)
// ---------------------------------------------------------------------------

// rho = rA0.rA of the first iteration
__global__ __launch_bounds__(256) void kb_bicg_rho(BatchMap B, const float* __restrict__ rA0,
                                                   const float* __restrict__ r, KState* stBase,
                                                   double* partials, int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(1, part);
  const int i = u.r0 + threadIdx.x;
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    const long k = (long)(u.rb + q) * B.N + i;
    double v[1] = {i < u.r1 ? (double)(rA0[k] * r[k]) : 0.0};
    part.add(q, v);
  }
  part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

// The BiCGStab SpMVs of a batch are limited by gathered loads rather than
// HBM bytes: forming p_j and y_j = rD_j p_j at each neighbour costs four
// gathers per neighbour and RHS (three for s and z).  So a streaming pass
// forms the vector and its preconditioned image at every row first, and the
// SpMV gathers that one vector.  Every value is the same float expression as
// when formed at the gathers.

// p = r + beta (p - omega AyA), y = rD p, for every running RHS
__global__ __launch_bounds__(256) void kb_bicg_py(BatchMap B, const float* __restrict__ r,
                                                  const float* __restrict__ rD,
                                                  const float* __restrict__ pold,
                                                  float* __restrict__ pnew,
                                                  const float* __restrict__ AyAold,
                                                  float* __restrict__ yA,
                                                  const KState* stBase) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  const int i = u.r0 + threadIdx.x;
  if (!run || i >= u.r1) return;
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    const int seg = u.sys * B.nrhs + u.rb + q;
    const KState* st = stBase + seg;
    const long k = (long)(u.rb + q) * B.N + i;
    const float beta = (float)st->beta, omega = (float)st->omega;
    const float di = rD[diag_off(B, seg) + i];
    const float pi = r[k] + beta * (pold[k] - omega * AyAold[k]);
    pnew[k] = pi;
    yA[k] = di * pi;
  }
}

// AyA = A y and rA0.AyA
__global__ __launch_bounds__(256) void kb_bicg_y_spmv(BatchMap B, LduOp<float> A,
                                                      const float* __restrict__ yA,
                                                      float* __restrict__ AyAnew,
                                                      const float* __restrict__ rA0,
                                                      KState* stBase, double* partials,
                                                      int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(1, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
  RowSlots<float> rs;
  if (in && A.slots) rs.load(A, i, u.t);
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    double v[1] = {0.0};
    if (in) {
      const int seg = u.sys * B.nrhs + u.rb + q;
      const long voff = (long)(u.rb + q) * B.N, doff = diag_off(B, seg);
      const float* y_ = yA + voff;
      const float a0 = rA0[voff + i];
      const float y = batch_row(A, rs, doff, i, y_[i], [&](int j) { return y_[j]; });
      AyAnew[voff + i] = y;
      v[0] = (double)(a0 * y);
    }
    part.add(q, v);
  }
  part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

// s = r - alpha AyA, z = rD s, for every running RHS
__global__ __launch_bounds__(256) void kb_bicg_sz(BatchMap B, const float* __restrict__ r,
                                                  const float* __restrict__ rD,
                                                  const float* __restrict__ AyA,
                                                  float* __restrict__ sA, float* __restrict__ zA,
                                                  const KState* stBase) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  const int i = u.r0 + threadIdx.x;
  if (!run || i >= u.r1) return;
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    const int seg = u.sys * B.nrhs + u.rb + q;
    const long k = (long)(u.rb + q) * B.N + i;
    const float alpha = (float)stBase[seg].alpha;
    const float di = rD[diag_off(B, seg) + i];
    const float si = r[k] - alpha * AyA[k];
    sA[k] = si;
    zA[k] = di * si;
  }
}

// t = A z, and |s|, s.t, t.t
__global__ __launch_bounds__(256) void kb_bicg_z_spmv(BatchMap B, LduOp<float> A,
                                                      const float* __restrict__ sA,
                                                      const float* __restrict__ zA,
                                                      float* __restrict__ tA, KState* stBase,
                                                      double* partials, int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  if (!run) return;
  HIPFOAM_RHS_PARTIALS(3, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
  RowSlots<float> rs;
  if (in && A.slots) rs.load(A, i, u.t);
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    if (!((run >> q) & 1)) continue;
    double v[3] = {0.0, 0.0, 0.0};
    if (in) {
      const int seg = u.sys * B.nrhs + u.rb + q;
      const long voff = (long)(u.rb + q) * B.N, doff = diag_off(B, seg);
      const float* z_ = zA + voff;
      const float si = sA[voff + i];
      const float t = batch_row(A, rs, doff, i, z_[i], [&](int j) { return z_[j]; });
      tA[voff + i] = t;
      v[0] = (double)fabsf(si);
      v[1] = (double)(si * t);
      v[2] = (double)(t * t);
    }
    part.add(q, v);
  }
  part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

__global__ __launch_bounds__(256) void kb_bicg_update_rho(BatchMap B, float* __restrict__ delta,
                                                          float* __restrict__ deltaBest,
                                                          const float* __restrict__ yA,
                                                          const float* __restrict__ zA,
                                                          float* __restrict__ r,
                                                          const float* __restrict__ sA,
                                                          const float* __restrict__ tA,
                                                          const float* __restrict__ rA0,
                                                          KState* stBase, double* partials,
                                                          int pstride) {
  const TileGroup u = tile_group(B, logical_block(B.xcdRemap));
  const unsigned run = status_bits(stBase, u, B.nrhs, 1 << ST_RUNNING);
  const unsigned early = status_bits(stBase, u, B.nrhs, 1 << ST_EARLY);
  if (!(run | early)) return;
  HIPFOAM_RHS_PARTIALS(2, part);
  const int i = u.r0 + threadIdx.x;
  const bool in = i < u.r1;
#pragma unroll 1
  for (int q = 0; q < u.nr; ++q) {
    const KState* st = stBase + u.sys * B.nrhs + u.rb + q;
    const long k = (long)(u.rb + q) * B.N + i;
    if ((early >> q) & 1) {
      if (in) delta[k] += (float)st->alpha * yA[k];  // FinBicgUpdateRho completes the exit
      continue;
    }
    if (!((run >> q) & 1)) continue;
    double v[2] = {0.0, 0.0};
    if (in) {
      const float alpha = (float)st->alpha, omega = (float)st->omega;
      const float d = delta[k];
      if (st->copyBest != 0) deltaBest[k] = d;
      delta[k] = d + (alpha * yA[k] + omega * zA[k]);
      const float ri = sA[k] - omega * tA[k];
      r[k] = ri;
      v[0] = (double)fabsf(ri);
      v[1] = (double)(rA0[k] * ri);
    }
    part.add(q, v);
  }
  if (run) part.write(partials, pstride, u.t * B.nrhs + u.rb, run);
}

#undef HIPFOAM_RHS_PARTIALS

// ---------------------------------------------------------------------------
// segmented second stage
// ---------------------------------------------------------------------------

// One workgroup per segment: fold partials of workgroups tile*nRhs + rhs over
// the system's tiles (fixed order), then apply the finaliser to the segment.
template <int NV, bool LASTMAX, typename F>
__global__ __launch_bounds__(256) void k_seg_finalize(BatchMap B, const double* __restrict__ partials,
                                                      int pstride, int checkStatus,
                                                      int runMask, F f) {
  const int seg = blockIdx.x;
  const int sys = seg / B.nrhs, rhs = seg - sys * B.nrhs;
  const F fs = f.at(seg);
  if (checkStatus && !((runMask >> fs.status()) & 1)) return;
  __shared__ double sh[NV][4];
  const int t0 = B.sysTile0[sys], t1 = B.sysTile0[sys + 1];
  double v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = (LASTMAX && j == NV - 1) ? -1.0e308 : 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x) {
    const int lb = t * B.nrhs + rhs;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const double p = partials[j * pstride + lb];
      v[j] = (LASTMAX && j == NV - 1) ? fmax(v[j], p) : v[j] + p;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const bool mx = LASTMAX && j == NV - 1;
    const double t = mx ? wave_max(v[j]) : wave_sum(v[j]);
    if (lane == 0) sh[j][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double out[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const bool mx = LASTMAX && j == NV - 1;
      double t = sh[j][0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = mx ? fmax(t, sh[j][w]) : t + sh[j][w];
      out[j] = t;
    }
    fs(out);
  }
}

// Adapter giving a single-system finaliser per-segment addressing.
template <typename F>
struct SegK {
  F f;
  __device__ SegK at(int seg) const {
    SegK g = *this;
    g.f.st += seg;
    return g;
  }
  __device__ int status() const { return f.st->status; }
  __device__ void operator()(const double* t) const { f(t); }
};

// FP64 finalisers per segment; segments whose mask entry is 0 are left alone.
template <typename F>
struct SegO {
  F f;
  const int32_t* segRows;  // rows of each segment (the FP64 finalisers divide by n)
  const int32_t* mask;     // per-segment run flag
  int run;
  __device__ SegO at(int seg) const {
    SegO g = *this;
    g.f.os += seg;
    g.f.n = segRows[seg];
    g.run = mask[seg];
    return g;
  }
  __device__ int status() const { return run ? ST_RUNNING : ST_CONVERGED; }
  __device__ void operator()(const double* t) const { f(t); }
};

// Per-segment prepare: bail count and the Gershgorin bound of the mode the
// segment asked for (segMode: 1 exact, 2 deflated, +4 want beta).
struct FinPrepareB {
  OState* os;
  const int32_t* segMode;
  int seg;
  __device__ FinPrepareB at(int s) const {
    FinPrepareB g = *this;
    g.os += s;
    g.seg = s;
    return g;
  }
  __device__ int status() const { return ST_RUNNING; }
  __device__ void operator()(const double* t) const {
    const int mode = segMode[seg];
    if (!(mode & 3)) return;
    os->bail = t[0];
    if (mode & 4) {
      if ((mode & 3) == 2)
        os->betaDefl = t[1];
      else
        os->betaExact = t[1];
    }
  }
};

// flag = 1 while any segment is still iterating (status RUNNING or EARLY)
__global__ __launch_bounds__(256) void kb_any_running(const KState* stBase, int nseg,
                                                      int32_t* flag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
    const int st = stBase[s].status;
    if (st == ST_RUNNING || st == ST_EARLY) any = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) *flag = any;
}

// ---------------------------------------------------------------------------
// FP64 refinement steps, per segment
// ---------------------------------------------------------------------------

// u = A 1 (once per system, written for rhs 0 of the segment only), Σ psi, pinned count,
// u.u, Σ u
__global__ __launch_bounds__(256) void kb_rowsum(BatchMap B, LduOp<double> A,
                                                 const double* __restrict__ psi,
                                                 double* __restrict__ u, const int32_t* active,
                                                 double* partials, int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && active[g.seg]) {
    const double ui = ldu_row_t(A, i, g.t, 1.0, [](int) { return 1.0; });
    if (g.rhs == 0) u[i] = ui;
    v[0] = psi[g.voff + i];
    v[1] = (fabs(ui) > 1e-8 * fabs(A.diag[i])) ? 1.0 : 0.0;
    v[2] = ui * ui;
    v[3] = ui;
  }
  block_partials<4>(v, partials, pstride, lb);
}

__global__ __launch_bounds__(256) void kb_initial_residual(BatchMap B, LduOp<double> A,
                                                           const double* __restrict__ psi,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ u,
                                                           double* __restrict__ r,
                                                           const OState* osBase,
                                                           const int32_t* active,
                                                           double* partials, int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  double v[3] = {0.0, 0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && active[g.seg]) {
    const double xbar = osBase[g.seg].xbar;
    const double* ps = psi + g.voff;
    const double wA = ldu_row_t(A, i, g.t, ps[i], [&](int j) { return ps[j]; });
    const double ri = b[g.voff + i] - wA;
    r[g.voff + i] = ri;
    const double t = u[i] * xbar;
    v[0] = fabs(ri);
    v[1] = fabs(wA - t) + fabs(b[g.voff + i] - t);
    v[2] = ri;
  }
  block_partials<3>(v, partials, pstride, lb);
}

// FP32 inner diagonal of one segment: deflate / exact per segment flags.
__global__ __launch_bounds__(256) void kb_prepare_inner(BatchMap B, LduOp<double> A,
                                                        const double* __restrict__ upperF,
                                                        const int32_t* __restrict__ losort,
                                                        const double* __restrict__ u,
                                                        const int32_t* __restrict__ segMode,
                                                        float* __restrict__ diag32,
                                                        float* __restrict__ rD32,
                                                        int32_t* __restrict__ segPlane,
                                                        double* partials, int pstride) {
  // segMode: 0 = leave, 1 = exact diag, 2 = deflated diag (diag - u), +4 = want beta.
  // The segment's plane (0 exact, 1 deflated) is written by every segment
  // that selects it, with the same values: both depend on the matrix only.
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const int mode = segMode[g.seg];
  const int plane = (mode & 3) == 2 ? 1 : 0;
  if ((mode & 3) && threadIdx.x == 0) segPlane[g.seg] = plane;
  double v[2] = {0.0, -1.0e308};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && (mode & 3)) {
    const double d = A.diag[i];
    double ds = d;
    if ((mode & 3) == 2) {
      ds = d - u[i];
      if (ds * d <= kVSmall * fabs(ds)) v[0] = 1.0;
    }
    const long doff = (long)plane * B.N;
    diag32[doff + i] = (float)ds;
    rD32[doff + i] = (float)(1.0 / ds);
    if (mode & 4) {
      double so = 0.0;
      for (int f = A.ownStart[i]; f < A.ownStart[i + 1]; ++f) so += fabs(upperF[f]);
      for (int k = A.losortStart[i]; k < A.losortStart[i + 1]; ++k) so += fabs(upperF[losort[k]]);
      v[1] = 1.0 + so / fabs(ds);
    }
  }
  block_partials<2, true>(v, partials, pstride, lb);
}

// rA = float(r64 - mean) (deflated segments; with pinned, r64 - rproj u);
// delta = p0 = p1 = AyA = 0; rA0 = rA, for active segments
__global__ __launch_bounds__(256) void kb_inner_init(BatchMap B, const double* __restrict__ r64,
                                                     const double* __restrict__ u, int pinned,
                                                     const OState* osBase, const int32_t* deflated,
                                                     const int32_t* active, float* __restrict__ rA,
                                                     float* __restrict__ rA0,
                                                     float* __restrict__ delta,
                                                     float* __restrict__ p0,
                                                     float* __restrict__ p1,
                                                     float* __restrict__ AyA) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const int i = g.r0 + threadIdx.x;
  if (i >= g.r1 || !active[g.seg]) return;
  const long k = g.voff + i;
  const bool defl = deflated[g.seg] != 0;
  const double m = defl ? (pinned ? osBase[g.seg].rproj : osBase[g.seg].rmean) : 0.0;
  const float v = (float)(defl ? (pinned ? r64[k] - m * u[i] : r64[k] - m) : r64[k]);
  rA[k] = v;
  if (rA0) rA0[k] = v;
  delta[k] = 0.f;
  p0[k] = 0.f;
  p1[k] = 0.f;
  if (AyA) AyA[k] = 0.f;
}

__global__ __launch_bounds__(256) void kb_delta_mean(BatchMap B, const float* __restrict__ delta,
                                                     const int32_t* active, double* partials,
                                                     int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const int i = g.r0 + threadIdx.x;
  double v[1] = {(i < g.r1 && active[g.seg]) ? (double)delta[g.voff + i] : 0.0};
  block_partials<1>(v, partials, pstride, lb);
}

// psi += double(delta) - dmean (psi ping-pong), r = b - A psi, partials u.r, |r|, r
__global__ __launch_bounds__(256) void kb_refine_residual(
    BatchMap B, LduOp<double> A, const double* __restrict__ psiOld,
    const float* __restrict__ delta, double* __restrict__ psiNew, const double* __restrict__ b,
    const double* __restrict__ u, double* __restrict__ r, const OState* osBase,
    const int32_t* deflated, const int32_t* active, double* partials, int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  double v[3] = {0.0, 0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1) {
    const long k = g.voff + i;
    if (!active[g.seg]) {
      psiNew[k] = psiOld[k];
    } else {
      const double dm = deflated[g.seg] ? osBase[g.seg].dmean : 0.0;
      const double* po = psiOld + g.voff;
      const float* dl = delta + g.voff;
      auto pf = [&](int j) { return po[j] + ((double)dl[j] - dm); };
      const double pi = pf(i);
      const double wA = ldu_row_t(A, i, g.t, pi, pf);
      psiNew[k] = pi;
      const double ri = b[k] - wA;
      r[k] = ri;
      v[0] = u[i] * ri;
      v[1] = fabs(ri);
      v[2] = ri;
    }
  }
  block_partials<3>(v, partials, pstride, lb);
}

__global__ __launch_bounds__(256) void kb_const_shift(BatchMap B, double* __restrict__ psi,
                                                      double* __restrict__ r,
                                                      const double* __restrict__ u,
                                                      const OState* osBase,
                                                      const int32_t* shift, double* partials,
                                                      int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  double v[2] = {0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && shift[g.seg]) {
    const long k = g.voff + i;
    const double c = osBase[g.seg].c;
    psi[k] += c;
    const double ri = r[k] - c * u[i];
    r[k] = ri;
    v[0] = fabs(ri);
    v[1] = ri;
  }
  block_partials<2>(v, partials, pstride, lb);
}

}  // namespace hipfoam
