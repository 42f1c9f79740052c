// hipfoam_batched.hpp -- batched solves: many independent LDU systems, each
// with several right-hand sides, advanced together by one launch per step.
//
// Layout.  The systems are concatenated into one block-diagonal "super-mesh"
// (global row and face numbering, addressing rebased once at set-up), so the
// LduOp of the single-system path applies unchanged.  Right-hand sides are
// stored [rhs][row] (stride N = total rows); the off-diagonal coefficients
// are shared by every RHS of a system, the FP32 diagonal / rD are per RHS
// (a segment may switch to the exact operator on its own).
//
// Launch shape.  Rows are cut into tiles that never straddle a system.  One
// workgroup handles one (tile, rhs): logical index lb = tile * nRhs + rhs, so
// the nRhs workgroups reading the same matrix rows are logically adjacent and,
// with the XCD remap, run on the same XCD and share its L2.
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
  int32_t ntiles, nrhs, N, xcdRemap;
};

struct BlockSeg {
  int r0, r1, rhs, seg;
  long voff;  // rhs * N: offset of this RHS in the vectors
};

__device__ __forceinline__ BlockSeg block_seg(const BatchMap& B, int lb) {
  BlockSeg s;
  const int t = lb / B.nrhs;
  s.rhs = lb - t * B.nrhs;
  s.r0 = B.tileRow0[t];
  s.r1 = B.tileRow0[t + 1];
  s.seg = B.tileSys[t] * B.nrhs + s.rhs;
  s.voff = (long)s.rhs * B.N;
  return s;
}

// Per-segment FP32 diagonal view: the operator with diag/rD of RHS `rhs`.
template <typename T>
__device__ __forceinline__ LduOp<T> seg_op(LduOp<T> A, long voff) {
  A.diag += voff;
  return A;
}

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void kb_cheb_step(BatchMap B, LduOp<float> A0,
                                                    const float* __restrict__ r,
                                                    const float* __restrict__ rD,
                                                    const float* __restrict__ zin,
                                                    float* __restrict__ w, float* __restrict__ zout,
                                                    KState* stBase, double* partials, int step) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  KState* st = stBase + g.seg;
  if (st->status != ST_RUNNING) return;
  const LduOp<float> A = seg_op(A0, g.voff);
  const float* rr = r + g.voff;
  const float* dd = rD + g.voff;
  const float* zi_ = zin + g.voff;
  const float c2 = (float)st->chebC2[step - 1];
  const float c3 = (float)st->chebC3[step - 1];
  const float c1 = (float)st->chebC1;
  double dot = 0.0;
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1) {
    float zi_new, wi;
    if (FIRST) {
      const float z0 = c1 * dd[i] * rr[i];
      const float v = ldu_row(A, i, z0, [&](int j) { return c1 * dd[j] * rr[j]; });
      wi = c2 * z0 + c3 * dd[i] * (rr[i] - v);
      zi_new = z0 + wi;
    } else {
      const float zi = zi_[i];
      const float v = ldu_row(A, i, zi, [&](int j) { return zi_[j]; });
      wi = c2 * w[g.voff + i] + c3 * dd[i] * (rr[i] - v);
      zi_new = zi + wi;
    }
    if (!LAST) w[g.voff + i] = wi;
    zout[g.voff + i] = zi_new;
    if (LAST) dot = (double)(zi_new * rr[i]);
  }
  if (LAST) {
    double v[1] = {dot};
    block_partials<1>(v, partials, 0, lb);
  }
}

__global__ __launch_bounds__(256) void kb_pcg_p_spmv(BatchMap B, LduOp<float> A0,
                                                     const float* __restrict__ wA,
                                                     const float* __restrict__ pold,
                                                     float* __restrict__ pnew,
                                                     float* __restrict__ Ap, KState* stBase,
                                                     double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  KState* st = stBase + g.seg;
  if (st->status != ST_RUNNING) return;
  const LduOp<float> A = seg_op(A0, g.voff);
  const float beta = (float)st->beta;
  const float* w_ = wA + g.voff;
  const float* po = pold + g.voff;
  double dot = 0.0;
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1) {
    const float pi = w_[i] + beta * po[i];
    const float y = ldu_row(A, i, pi, [&](int j) { return w_[j] + beta * po[j]; });
    pnew[g.voff + i] = pi;
    Ap[g.voff + i] = y;
    dot = (double)(pi * y);
  }
  double v[1] = {dot};
  block_partials<1>(v, partials, 0, lb);
}

__global__ __launch_bounds__(256) void kb_pcg_update(BatchMap B, float* __restrict__ delta,
                                                     float* __restrict__ deltaBest,
                                                     float* __restrict__ r,
                                                     const float* __restrict__ p,
                                                     const float* __restrict__ Ap,
                                                     KState* stBase, double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  KState* st = stBase + g.seg;
  if (st->status != ST_RUNNING) return;
  const float alpha = (float)st->alpha;
  const bool copy = st->copyBest != 0;
  double m = 0.0;
  const long i = g.voff + g.r0 + threadIdx.x;
  if (g.r0 + (int)threadIdx.x < g.r1) {
    const float d = delta[i];
    if (copy) deltaBest[i] = d;
    delta[i] = d + alpha * p[i];
    const float ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    m = (double)fabsf(ri);
  }
  double v[1] = {m};
  block_partials<1>(v, partials, 0, lb);
}

// ---------------------------------------------------------------------------
// BiCGStab
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void kb_bicg_rho(BatchMap B, const float* __restrict__ rA0,
                                                   const float* __restrict__ r, KState* stBase,
                                                   double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  if (stBase[g.seg].status != ST_RUNNING) return;
  double d = 0.0;
  const long i = g.voff + g.r0 + threadIdx.x;
  if (g.r0 + (int)threadIdx.x < g.r1) d = (double)(rA0[i] * r[i]);
  double v[1] = {d};
  block_partials<1>(v, partials, 0, lb);
}

__global__ __launch_bounds__(256) void kb_bicg_p_y(BatchMap B, float* __restrict__ p,
                                                   const float* __restrict__ r,
                                                   const float* __restrict__ AyA,
                                                   const float* __restrict__ rD,
                                                   float* __restrict__ yA, KState* stBase) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const KState* st = stBase + g.seg;
  if (st->status != ST_RUNNING) return;
  const float beta = (float)st->beta, omega = (float)st->omega;
  const long i = g.voff + g.r0 + threadIdx.x;
  if (g.r0 + (int)threadIdx.x < g.r1) {
    const float pi = r[i] + beta * (p[i] - omega * AyA[i]);
    p[i] = pi;
    yA[i] = rD[i] * pi;
  }
}

__global__ __launch_bounds__(256) void kb_bicg_spmv_alpha(BatchMap B, LduOp<float> A0,
                                                          const float* __restrict__ yA,
                                                          const float* __restrict__ rA0,
                                                          float* __restrict__ AyA,
                                                          KState* stBase, double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  if (stBase[g.seg].status != ST_RUNNING) return;
  const LduOp<float> A = seg_op(A0, g.voff);
  const float* y_ = yA + g.voff;
  double d = 0.0;
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1) {
    const float y = ldu_row(A, i, y_[i], [&](int j) { return y_[j]; });
    AyA[g.voff + i] = y;
    d = (double)(rA0[g.voff + i] * y);
  }
  double v[1] = {d};
  block_partials<1>(v, partials, 0, lb);
}

__global__ __launch_bounds__(256) void kb_bicg_s(BatchMap B, float* __restrict__ sA,
                                                 const float* __restrict__ r,
                                                 const float* __restrict__ AyA, KState* stBase,
                                                 double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const KState* st = stBase + g.seg;
  if (st->status != ST_RUNNING) return;
  const float alpha = (float)st->alpha;
  double m = 0.0;
  const long i = g.voff + g.r0 + threadIdx.x;
  if (g.r0 + (int)threadIdx.x < g.r1) {
    const float s = r[i] - alpha * AyA[i];
    sA[i] = s;
    m = (double)fabsf(s);
  }
  double v[1] = {m};
  block_partials<1>(v, partials, 0, lb);
}

__global__ __launch_bounds__(256) void kb_bicg_z_spmv(BatchMap B, LduOp<float> A0,
                                                      const float* __restrict__ rD,
                                                      const float* __restrict__ sA,
                                                      float* __restrict__ zA,
                                                      float* __restrict__ tA, KState* stBase,
                                                      double* partials, int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  if (stBase[g.seg].status != ST_RUNNING) return;
  const LduOp<float> A = seg_op(A0, g.voff);
  const float* d_ = rD + g.voff;
  const float* s_ = sA + g.voff;
  double d[2] = {0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1) {
    const float zi = d_[i] * s_[i];
    const float t = ldu_row(A, i, zi, [&](int j) { return d_[j] * s_[j]; });
    zA[g.voff + i] = zi;
    tA[g.voff + i] = t;
    d[0] = (double)(s_[i] * t);
    d[1] = (double)(t * t);
  }
  block_partials<2>(d, partials, pstride, lb);
}

__global__ __launch_bounds__(256) void kb_bicg_update(BatchMap B, float* __restrict__ delta,
                                                      float* __restrict__ deltaBest,
                                                      const float* __restrict__ yA,
                                                      const float* __restrict__ zA,
                                                      float* __restrict__ r,
                                                      const float* __restrict__ sA,
                                                      const float* __restrict__ tA,
                                                      KState* stBase, double* partials) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const KState* st = stBase + g.seg;
  const int status = st->status;
  if (status != ST_RUNNING && status != ST_EARLY) return;
  const float alpha = (float)st->alpha;
  const long i = g.voff + g.r0 + threadIdx.x;
  const bool in = g.r0 + (int)threadIdx.x < g.r1;
  if (status == ST_EARLY) {
    if (in) delta[i] += alpha * yA[i];
    return;
  }
  const float omega = (float)st->omega;
  const bool copy = st->copyBest != 0;
  double m = 0.0;
  if (in) {
    const float d = delta[i];
    if (copy) deltaBest[i] = d;
    delta[i] = d + (alpha * yA[i] + omega * zA[i]);
    const float ri = sA[i] - omega * tA[i];
    r[i] = ri;
    m = (double)fabsf(ri);
  }
  double v[1] = {m};
  block_partials<1>(v, partials, 0, lb);
}

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

// u = A 1 (once per system, written for rhs 0 of the segment only), Σ psi, pinned count, u.u
__global__ __launch_bounds__(256) void kb_rowsum(BatchMap B, LduOp<double> A,
                                                 const double* __restrict__ psi,
                                                 double* __restrict__ u, const int32_t* active,
                                                 double* partials, int pstride) {
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  double v[3] = {0.0, 0.0, 0.0};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && active[g.seg]) {
    const double ui = ldu_row(A, i, 1.0, [](int) { return 1.0; });
    if (g.rhs == 0) u[i] = ui;
    v[0] = psi[g.voff + i];
    v[1] = (fabs(ui) > 1e-8 * fabs(A.diag[i])) ? 1.0 : 0.0;
    v[2] = ui * ui;
  }
  block_partials<3>(v, partials, pstride, lb);
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
    const double wA = ldu_row(A, i, ps[i], [&](int j) { return ps[j]; });
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
                                                        double* partials, int pstride) {
  // segMode: 0 = leave, 1 = exact diag, 2 = deflated diag (diag - u), +4 = want beta
  const int lb = logical_block(B.xcdRemap);
  const BlockSeg g = block_seg(B, lb);
  const int mode = segMode[g.seg];
  double v[2] = {0.0, -1.0e308};
  const int i = g.r0 + threadIdx.x;
  if (i < g.r1 && (mode & 3)) {
    const double d = A.diag[i];
    double ds = d;
    if ((mode & 3) == 2) {
      ds = d - u[i];
      if (ds * d <= kVSmall * fabs(ds)) v[0] = 1.0;
    }
    diag32[g.voff + i] = (float)ds;
    rD32[g.voff + i] = (float)(1.0 / ds);
    if (mode & 4) {
      double so = 0.0;
      for (int f = A.ownStart[i]; f < A.ownStart[i + 1]; ++f) so += fabs(upperF[f]);
      for (int k = A.losortStart[i]; k < A.losortStart[i + 1]; ++k) so += fabs(upperF[losort[k]]);
      v[1] = 1.0 + so / fabs(ds);
    }
  }
  block_partials<2, true>(v, partials, pstride, lb);
}

// rA = float(r64 - mean); delta = p0 = p1 = AyA = 0; rA0 = rA, for active segments
__global__ __launch_bounds__(256) void kb_inner_init(BatchMap B, const double* __restrict__ r64,
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
  const double m = deflated[g.seg] ? osBase[g.seg].rmean : 0.0;
  const float v = (float)(deflated[g.seg] ? r64[k] - m : r64[k]);
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
      const double wA = ldu_row(A, i, pi, pf);
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
