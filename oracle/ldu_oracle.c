/*
 * ldu_oracle.c -- CPU restatement of the metal4Foam solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline: it is loaded by tests/, by the smoke check and by the CPU leg of
 * the benchmark, and never by the product library.  The product path (the
 * HIP library behind include/hipfoam.h) must not link, load or call it.
 *
 * What it restates (reference = scotCW/metal4Foam, paths relative to its
 * root; OpenFOAM v2506 built-ins restated from their public semantics):
 *
 *   orc_amul / orc_tmul / orc_sumA    lduMatrix::Amul/Tmul/sumA, called at
 *                                     src/metal4Foam/solvers/metalFoamSolver.C:109,333,168
 *   orc_norm_factor                   lduMatrix::solver::normFactor, called at
 *                                     metalFoamSolver.C:120
 *   orc_spmv_f                        FP32 row-gather LDU SpMV,
 *                                     src/metal4Foam/shaders/metalFoamKernels.metal:87-104
 *   orc_cheb_constants / _apply_f     encCheb, src/metal4Foam/backend/metalFoamBackend.C:339-385
 *                                     + cheb_init/cheb_step, metalFoamKernels.metal:153-194
 *   orc_inner_pcg_f                   metalFoamBackend::pcg, metalFoamBackend.C:568-732
 *   orc_inner_bicgstab_f              metalFoamBackend::bicgstab, metalFoamBackend.C:735-959
 *   orc_refined_solve                 metalFoamSolver::solveImpl, metalFoamSolver.C:77-437
 *   orc_pcg / orc_pbicgstab           OpenFOAM PCG / PBiCGStab (DIC, DILU, diagonal, or the
 *                                     Chebyshev-Jacobi preconditioner in FP64)
 *   orc_adjoint_grads                 dL/dupper, dL/dlower, dL/ddiag of L(x(A)), A x = b
 *   orc_pcg_diag_par                  OpenFOAM PCG(diagonal) over OpenMP threads (CPU
 *                                     baseline on all host cores)
 *
 * Reductions are accumulated in double and rounded once (the reference's
 * Metal simd_sum order cannot be reproduced bit-for-bit; parity with the FP32
 * path is by tolerance).  Build with -ffp-contract=off so every float
 * expression rounds exactly as written.
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_VSMALL 1.0e-300 /* OpenFOAM doubleScalarVSMALL */
#define ORC_SMALL 1.0e-20   /* solverPerformance small_ */

typedef struct
{
    int32_t n, nf;
    const int32_t* low;         /* lowerAddr (owner) [nf] */
    const int32_t* upp;         /* upperAddr (neighbour) [nf] */
    const int32_t* ownStart;    /* [n+1] */
    const int32_t* losort;      /* [nf] */
    const int32_t* losortStart; /* [n+1] */
} orc_addr;

typedef struct
{
    int nIterations;
    double finalResSum;
    int converged, singular, stagnated;
} orc_inner_stats;

typedef struct
{
    double tolerance, relTol;
    int maxIter, minIter;
    int deflate; /* 0 off, 1 on, 2 auto, 3 auto + projection along A 1 */
    int chebyshevOrder;
    double chebyshevRatio;
    int maxRestarts;
    int iterationsPerSync;
} orc_controls;

typedef struct
{
    double initialResidual, finalResidual, normFactor;
    int nIterations, restarts, converged, singular, deflated, polished;
    double chebBeta;
} orc_perf;

/* ------------------------------------------------------------------------ */
/* FP64 operator primitives (OpenFOAM lduMatrix semantics)                   */
/* ------------------------------------------------------------------------ */

/* y = A x, face loop as lduMatrix::Amul (metalFoamSolver.C:109, 333). */
void orc_amul(const orc_addr* a, const double* diag, const double* upper,
              const double* lower, const double* x, double* y)
{
    if (!lower) lower = upper;
    for (int32_t i = 0; i < a->n; ++i) y[i] = diag[i] * x[i];
    for (int32_t f = 0; f < a->nf; ++f)
    {
        y[a->upp[f]] += lower[f] * x[a->low[f]];
        y[a->low[f]] += upper[f] * x[a->upp[f]];
    }
}

/* y = A^T x (lduMatrix::Tmul: upper and lower swap roles). */
void orc_tmul(const orc_addr* a, const double* diag, const double* upper,
              const double* lower, const double* x, double* y)
{
    orc_amul(a, diag, lower ? lower : upper, upper, x, y);
}

/* u = A 1 (lduMatrix::sumA, metalFoamSolver.C:160-168; no coupled patches). */
void orc_sumA(const orc_addr* a, const double* diag, const double* upper,
              const double* lower, double* u)
{
    if (!lower) lower = upper;
    for (int32_t i = 0; i < a->n; ++i) u[i] = diag[i];
    for (int32_t f = 0; f < a->nf; ++f)
    {
        u[a->low[f]] += upper[f];
        u[a->upp[f]] += lower[f];
    }
}

/* Default L1 normFactor (metalFoamSolver.C:120):
 * sum(|A psi - xbar A1| + |b - xbar A1|) + small, xbar = average(psi). */
double orc_norm_factor(const orc_addr* a, const double* diag,
                       const double* upper, const double* lower,
                       const double* psi, const double* b, const double* Apsi)
{
    const int32_t n = a->n;
    double* u = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_sumA(a, diag, upper, lower, u);
    double avg = 0;
    for (int32_t i = 0; i < n; ++i) avg += psi[i];
    avg /= (n > 0 ? n : 1);
    double s = 0;
    for (int32_t i = 0; i < n; ++i)
    {
        const double t = u[i] * avg;
        s += fabs(Apsi[i] - t) + fabs(b[i] - t);
    }
    free(u);
    return s + ORC_SMALL;
}

static double sum_mag(int32_t n, const double* r)
{
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += fabs(r[i]);
    return s;
}

static int check_convergence(double res, double res0, double tol, double relTol)
{
    return (res < tol) || (relTol > ORC_SMALL && res < relTol * res0);
}

/* ------------------------------------------------------------------------ */
/* FP32 kernels of the reference GPU path                                   */
/* ------------------------------------------------------------------------ */

/* Row-gather SpMV in FP32, term order diag, owner faces ascending, losort
 * faces ascending (metalFoamKernels.metal:87-104). */
void orc_spmv_f(const orc_addr* a, const float* diag, const float* upper,
                const float* lower, const float* x, float* y)
{
    if (!lower) lower = upper;
    for (int32_t i = 0; i < a->n; ++i)
    {
        float s = diag[i] * x[i];
        for (int32_t f = a->ownStart[i]; f < a->ownStart[i + 1]; ++f)
            s += upper[f] * x[a->upp[f]];
        for (int32_t k = a->losortStart[i]; k < a->losortStart[i + 1]; ++k)
        {
            const int32_t f = a->losort[k];
            s += lower[f] * x[a->low[f]];
        }
        y[i] = s;
    }
}

static float dotf(int32_t n, const float* x, const float* y)
{
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += (double)(x[i] * y[i]);
    return (float)s;
}

static float summagf(int32_t n, const float* x)
{
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += (double)fabsf(x[i]);
    return (float)s;
}

/* Chebyshev constants of encCheb (metalFoamBackend.C:349-372): c1, then
 * (c2, c3) for each of the order-1 steps.  Returns the count written. */
int orc_cheb_constants(int order, double beta, double ratio, float* c1,
                       float* c2, float* c3)
{
    const double alpha = beta / ratio;
    const double theta = 0.5 * (beta + alpha);
    const double delta = 0.5 * (beta - alpha);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    *c1 = (float)(1.0 / theta);
    for (int i = 1; i < order; ++i)
    {
        const double rhoNew = 1.0 / (2.0 * sigma - rho);
        c2[i - 1] = (float)(rhoNew * rho);
        c3[i - 1] = (float)(2.0 * rhoNew / delta);
        rho = rhoNew;
    }
    return order - 1;
}

/* z = Cheb_order(rD A) applied to r, FP32 (cheb_init / ldu_spmv / cheb_step,
 * metalFoamKernels.metal:153-194; recurrence metalFoamBackend.C:342-385).
 * w, v: scratch [n]. */
void orc_cheb_apply_f(const orc_addr* a, const float* diag, const float* upper,
                      const float* lower, const float* rD, int order,
                      double beta, double ratio, const float* r, float* z,
                      float* w, float* v)
{
    float c1, c2[64], c3[64];
    if (order > 65) order = 65;
    orc_cheb_constants(order, beta, ratio, &c1, c2, c3);
    for (int32_t i = 0; i < a->n; ++i)
    {
        const float wi = c1 * rD[i] * r[i];
        w[i] = wi;
        z[i] = wi;
    }
    for (int s = 1; s < order; ++s)
    {
        orc_spmv_f(a, diag, upper, lower, z, v);
        const float cc2 = c2[s - 1], cc3 = c3[s - 1];
        for (int32_t i = 0; i < a->n; ++i)
        {
            const float wi = cc2 * w[i] + cc3 * rD[i] * (r[i] - v[i]);
            w[i] = wi;
            z[i] += wi;
        }
    }
}

/* FP64 Chebyshev-Jacobi apply (same recurrence, all in double): the
 * preconditioner of the plain-FP64 comparator. */
void orc_cheb_apply_d(const orc_addr* a, const double* diag,
                      const double* upper, const double* lower,
                      const double* rD, int order, double beta, double ratio,
                      const double* r, double* z, double* w, double* v)
{
    const double alpha = beta / ratio;
    const double theta = 0.5 * (beta + alpha);
    const double delta = 0.5 * (beta - alpha);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    const double c1 = 1.0 / theta;
    for (int32_t i = 0; i < a->n; ++i)
    {
        const double wi = c1 * rD[i] * r[i];
        w[i] = wi;
        z[i] = wi;
    }
    for (int s = 1; s < order; ++s)
    {
        orc_amul(a, diag, upper, lower, z, v);
        const double rhoNew = 1.0 / (2.0 * sigma - rho);
        const double c2 = rhoNew * rho;
        const double c3 = 2.0 * rhoNew / delta;
        rho = rhoNew;
        for (int32_t i = 0; i < a->n; ++i)
        {
            const double wi = c2 * w[i] + c3 * rD[i] * (r[i] - v[i]);
            w[i] = wi;
            z[i] += wi;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FP32 inner solvers (the reference backend's host loops)                  */
/* ------------------------------------------------------------------------ */

typedef struct
{
    const orc_addr* a;
    float *diag, *upper, *lower, *rD; /* lower == upper when symmetric */
} orc_coeffs_f;

/* setCoeffs (metalFoamBackend.C:522-565): round to FP32, rD = float(1/diag). */
static void set_coeffs_f(orc_coeffs_f* c, const double* diag,
                         const double* upper, const double* lower)
{
    const orc_addr* a = c->a;
    for (int32_t i = 0; i < a->n; ++i)
    {
        c->diag[i] = (float)diag[i];
        c->rD[i] = (float)(1.0 / diag[i]);
    }
    for (int32_t f = 0; f < a->nf; ++f) c->upper[f] = (float)upper[f];
    if (lower)
    {
        for (int32_t f = 0; f < a->nf; ++f) c->lower[f] = (float)lower[f];
    }
    else
    {
        memcpy(c->lower, c->upper, sizeof(float) * (size_t)a->nf);
    }
}

/* metalFoamBackend::pcg (metalFoamBackend.C:568-732).  delta [n] out. */
static orc_inner_stats inner_pcg_f(const orc_coeffs_f* c, const double* r,
                                   double target, int maxIter, int ips,
                                   int chebOrder, double chebBeta,
                                   double chebRatio, float* delta)
{
    const orc_addr* a = c->a;
    const int32_t n = a->n;
    orc_inner_stats st;
    memset(&st, 0, sizeof(st));
    if (n == 0)
    {
        st.singular = 1;
        return st;
    }
    const size_t bytes = sizeof(float) * (size_t)n;
    float* rA = (float*)malloc(bytes);
    float* p = (float*)calloc((size_t)n, sizeof(float));
    float* wA = (float*)malloc(bytes);
    float* best = (float*)calloc((size_t)n, sizeof(float));
    float* cw = (float*)malloc(bytes);
    float* cv = (float*)malloc(bytes);
    for (int32_t i = 0; i < n; ++i)
    {
        rA[i] = (float)r[i];
        delta[i] = 0.f;
    }
    float wArA = 0.f, wArAold = 0.f, wApA = 0.f, alpha = 0.f, beta = 0.f,
          res = 0.f;
    if (ips < 1) ips = 1;
    int it = 0, noProgress = 0;
    double bestRes = 1.0e300;
    const int stagLimit = 50;
    while (it < maxIter)
    {
        const int batch = ips < maxIter - it ? ips : maxIter - it;
        for (int k = 0; k < batch; ++k)
        {
            const int first = (it + k == 0);
            if (chebOrder > 1)
                orc_cheb_apply_f(a, c->diag, c->upper, c->lower, c->rD,
                                 chebOrder, chebBeta, chebRatio, rA, wA, cw,
                                 cv);
            else
                for (int32_t i = 0; i < n; ++i) wA[i] = rA[i] * c->rD[i];
            const float t0 = dotf(n, wA, rA);
            wArAold = wArA;
            wArA = t0;
            beta = first ? 0.f : t0 / wArAold;
            for (int32_t i = 0; i < n; ++i) p[i] = wA[i] + beta * p[i];
            orc_spmv_f(a, c->diag, c->upper, c->lower, p, wA);
            wApA = dotf(n, p, wA);
            alpha = wArA / wApA;
            for (int32_t i = 0; i < n; ++i)
            {
                delta[i] += alpha * p[i];
                rA[i] = rA[i] - alpha * wA[i];
            }
            res = summagf(n, rA);
        }
        it += batch;
        st.nIterations = it;
        st.finalResSum = (double)res;
        if (!isfinite((double)res) || !isfinite((double)wApA))
        {
            st.singular = 1;
            break;
        }
        if ((double)res < bestRes)
        {
            bestRes = res;
            noProgress = 0;
            memcpy(best, delta, bytes);
        }
        else
        {
            noProgress += batch;
        }
        if ((double)res <= target)
        {
            st.converged = 1;
            break;
        }
        if (noProgress >= stagLimit)
        {
            st.stagnated = 1;
            break;
        }
    }
    if (bestRes < st.finalResSum)
    {
        memcpy(delta, best, bytes);
        st.finalResSum = bestRes;
    }
    free(rA);
    free(p);
    free(wA);
    free(best);
    free(cw);
    free(cv);
    return st;
}

/* metalFoamBackend::bicgstab (metalFoamBackend.C:735-959). */
static orc_inner_stats inner_bicgstab_f(const orc_coeffs_f* c, const double* r,
                                        double target, int maxIter,
                                        float* delta)
{
    const orc_addr* a = c->a;
    const int32_t n = a->n;
    orc_inner_stats st;
    memset(&st, 0, sizeof(st));
    if (n == 0)
    {
        st.singular = 1;
        return st;
    }
    const size_t bytes = sizeof(float) * (size_t)n;
    float* rA = (float*)malloc(bytes);
    float* rA0 = (float*)malloc(bytes);
    float* p = (float*)calloc((size_t)n, sizeof(float));
    float* yA = (float*)calloc((size_t)n, sizeof(float));
    float* AyA = (float*)calloc((size_t)n, sizeof(float));
    float* sA = (float*)malloc(bytes);
    float* zA = (float*)malloc(bytes);
    float* tA = (float*)malloc(bytes);
    float* best = (float*)calloc((size_t)n, sizeof(float));
    for (int32_t i = 0; i < n; ++i)
    {
        rA[i] = (float)r[i];
        rA0[i] = rA[i];
        delta[i] = 0.f;
    }
    float rA0rA = 0.f, rA0rAold = 0.f, alpha = 0.f, beta = 0.f, omega = 0.f;
    int it = 0, noProgress = 0;
    double bestRes = 1.0e300;
    const int stagLimit = 50;
    while (it < maxIter)
    {
        const int first = (it == 0);
        const float t3 = dotf(n, rA0, rA);
        rA0rAold = rA0rA;
        rA0rA = t3;
        beta = first ? 0.f : (t3 / rA0rAold) * (alpha / omega);
        for (int32_t i = 0; i < n; ++i)
            p[i] = rA[i] + beta * (p[i] - omega * AyA[i]);
        for (int32_t i = 0; i < n; ++i) yA[i] = c->rD[i] * p[i];
        orc_spmv_f(a, c->diag, c->upper, c->lower, yA, AyA);
        const float t4 = dotf(n, AyA, rA0);
        alpha = rA0rA / t4;
        for (int32_t i = 0; i < n; ++i) sA[i] = rA[i] - alpha * AyA[i];
        float res = summagf(n, sA);
        if (!isfinite((double)rA0rA) || !isfinite((double)alpha) ||
            !isfinite((double)res))
        {
            st.singular = 1;
            break;
        }
        if ((double)res <= target)
        {
            for (int32_t i = 0; i < n; ++i) delta[i] += alpha * yA[i];
            st.nIterations = ++it;
            st.finalResSum = (double)res;
            st.converged = 1;
            goto done_early;
        }
        for (int32_t i = 0; i < n; ++i) zA[i] = c->rD[i] * sA[i];
        orc_spmv_f(a, c->diag, c->upper, c->lower, zA, tA);
        {
            const float tAsA = dotf(n, tA, sA);
            const float tAtA = dotf(n, tA, tA);
            omega = tAsA / tAtA;
        }
        for (int32_t i = 0; i < n; ++i)
        {
            delta[i] += alpha * yA[i] + omega * zA[i];
            rA[i] = sA[i] - omega * tA[i];
        }
        res = summagf(n, rA);
        ++it;
        st.nIterations = it;
        st.finalResSum = (double)res;
        if (!isfinite((double)res) || !isfinite((double)omega))
        {
            st.singular = 1;
            break;
        }
        if ((double)res < bestRes)
        {
            bestRes = res;
            noProgress = 0;
            memcpy(best, delta, bytes);
        }
        else
        {
            noProgress += 1;
        }
        if ((double)res <= target)
        {
            st.converged = 1;
            break;
        }
        if (noProgress >= stagLimit)
        {
            st.stagnated = 1;
            break;
        }
    }
    if (bestRes < st.finalResSum)
    {
        memcpy(delta, best, bytes);
        st.finalResSum = bestRes;
    }
done_early:
    free(rA);
    free(rA0);
    free(p);
    free(yA);
    free(AyA);
    free(sA);
    free(zA);
    free(tA);
    free(best);
    return st;
}

/* Public wrappers of the inner solvers (coefficients given in FP64 and
 * rounded as setCoeffs does). */
orc_inner_stats orc_inner_pcg_f(const orc_addr* a, const double* diag,
                                const double* upper, const double* r,
                                double target, int maxIter, int ips,
                                int chebOrder, double chebBeta,
                                double chebRatio, float* delta)
{
    orc_coeffs_f c;
    c.a = a;
    c.diag = (float*)malloc(sizeof(float) * (size_t)a->n);
    c.rD = (float*)malloc(sizeof(float) * (size_t)a->n);
    c.upper = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));
    c.lower = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));
    set_coeffs_f(&c, diag, upper, NULL);
    orc_inner_stats st = inner_pcg_f(&c, r, target, maxIter, ips, chebOrder,
                                     chebBeta, chebRatio, delta);
    free(c.diag);
    free(c.rD);
    free(c.upper);
    free(c.lower);
    return st;
}

orc_inner_stats orc_inner_bicgstab_f(const orc_addr* a, const double* diag,
                                     const double* upper, const double* lower,
                                     const double* r, double target,
                                     int maxIter, float* delta)
{
    orc_coeffs_f c;
    c.a = a;
    c.diag = (float*)malloc(sizeof(float) * (size_t)a->n);
    c.rD = (float*)malloc(sizeof(float) * (size_t)a->n);
    c.upper = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));
    c.lower = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));
    set_coeffs_f(&c, diag, upper, lower);
    orc_inner_stats st = inner_bicgstab_f(&c, r, target, maxIter, delta);
    free(c.diag);
    free(c.rD);
    free(c.upper);
    free(c.lower);
    return st;
}

/* Gershgorin bound of rD*A (metalFoamSolver.C:229-249): uses |upper| only. */
double orc_gershgorin(const orc_addr* a, const double* diagStar,
                      const double* upper)
{
    double* s = (double*)calloc((size_t)(a->n > 0 ? a->n : 1), sizeof(double));
    for (int32_t f = 0; f < a->nf; ++f)
    {
        s[a->low[f]] += fabs(upper[f]);
        s[a->upp[f]] += fabs(upper[f]);
    }
    double b = 0;
    for (int32_t i = 0; i < a->n; ++i)
    {
        const double t = 1.0 + s[i] / fabs(diagStar[i]);
        if (t > b) b = t;
    }
    free(s);
    return b;
}

/* ------------------------------------------------------------------------ */
/* metalFoamSolver::solveImpl (metalFoamSolver.C:77-437)                     */
/* ------------------------------------------------------------------------ */

int orc_refined_solve(const orc_addr* a, const double* diag,
                      const double* upper, const double* lower, int useBiCG,
                      const orc_controls* ctl, double* psi,
                      const double* source, orc_perf* perf)
{
    const int32_t n = a->n;
    const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    memset(perf, 0, sizeof(*perf));
    double* wA = (double*)malloc(nb);
    double* rA = (double*)malloc(nb);
    double* u = (double*)malloc(nb);
    double* diag0 = (double*)malloc(nb);
    double* rin = (double*)malloc(nb);
    float* delta = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));

    /* metalFoamSolver.C:108-110 */
    orc_amul(a, diag, upper, lower, psi, wA);
    for (int32_t i = 0; i < n; ++i) rA[i] = source[i] - wA[i];
    /* :120 */
    const double normFactor =
        orc_norm_factor(a, diag, upper, lower, psi, source, wA);
    perf->normFactor = normFactor;
    /* :127-129 */
    perf->initialResidual = sum_mag(n, rA) / normFactor;
    perf->finalResidual = perf->initialResidual;

    int converged = check_convergence(perf->finalResidual,
                                      perf->initialResidual, ctl->tolerance,
                                      ctl->relTol);
    if (ctl->minIter > 0 || !converged)
    {
        orc_coeffs_f c;
        c.a = a;
        c.diag = (float*)malloc(sizeof(float) * (size_t)n);
        c.rD = (float*)malloc(sizeof(float) * (size_t)n);
        c.upper = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));
        c.lower = (float*)malloc(sizeof(float) * (size_t)(a->nf + 1));

        /* :167-168 */
        orc_sumA(a, diag, upper, lower, u);
        /* :170-190; deflate 3 detects like auto (2) */
        const int pinnedProj = (ctl->deflate == 3);
        int deflate = 0;
        if (!useBiCG && ctl->deflate != 0)
        {
            if (ctl->deflate == 1)
            {
                deflate = 1;
            }
            else
            {
                int32_t nPinned = 0;
                for (int32_t i = 0; i < n; ++i)
                    if (fabs(u[i]) > 1e-8 * fabs(diag[i])) ++nPinned;
                deflate = (10 * (int64_t)nPinned <= n);
            }
        }
        /* :192-209 */
        const double* diagPtr = diag;
        if (deflate)
        {
            for (int32_t i = 0; i < n; ++i)
            {
                diag0[i] = diag[i] - u[i];
                if (diag0[i] * diag[i] <= ORC_VSMALL * fabs(diag0[i]))
                {
                    deflate = 0;
                    break;
                }
            }
            if (deflate) diagPtr = diag0;
        }
        perf->deflated = deflate;
        /* :211-220 (PCG and symmetric matrices pass lower == NULL) */
        set_coeffs_f(&c, diagPtr, upper, useBiCG ? lower : NULL);

        /* :222-249 */
        const int chebOrder = useBiCG ? 1 : ctl->chebyshevOrder;
        const double chebRatio = ctl->chebyshevRatio;
        double chebBeta = 2.0;
        if (chebOrder > 1) chebBeta = orc_gershgorin(a, diagPtr, upper);
        perf->chebBeta = chebBeta;

        /* :251-257 */
        const double tgt =
            fmax(ctl->tolerance, ctl->relTol * perf->initialResidual) *
            normFactor;

        int totalIter = 0, outer, singular = 0;
        double prevRes = perf->initialResidual;
        for (outer = 0; outer <= ctl->maxRestarts; ++outer)
        {
            /* :278-287; deflate 3 (no reference counterpart) projects the
             * inner right-hand side along u = A 1 instead of the constant:
             * r - (sum r / sum u) u has no component the deflated operator
             * cannot produce, whatever sum(b) is */
            const double* rPtr = rA;
            if (deflate)
            {
                double mean = 0;
                for (int32_t i = 0; i < n; ++i) mean += rA[i];
                if (pinnedProj)
                {
                    double su = 0;
                    for (int32_t i = 0; i < n; ++i) su += u[i];
                    const double w = su != 0.0 ? mean / su : 0.0;
                    for (int32_t i = 0; i < n; ++i) rin[i] = rA[i] - w * u[i];
                }
                else
                {
                    mean /= n;
                    for (int32_t i = 0; i < n; ++i) rin[i] = rA[i] - mean;
                }
                rPtr = rin;
            }
            /* :289-311 */
            orc_inner_stats st =
                useBiCG ? inner_bicgstab_f(&c, rPtr, tgt,
                                           ctl->maxIter - totalIter, delta)
                        : inner_pcg_f(&c, rPtr, tgt, ctl->maxIter - totalIter,
                                      ctl->iterationsPerSync, chebOrder,
                                      chebBeta, chebRatio, delta);
            totalIter += st.nIterations;
            singular = st.singular;
            /* :318-330 */
            double delMean = 0;
            if (deflate)
            {
                for (int32_t i = 0; i < n; ++i) delMean += (double)delta[i];
                delMean /= n;
            }
            for (int32_t i = 0; i < n; ++i) psi[i] += (double)delta[i] - delMean;
            /* :332-340 */
            orc_amul(a, diag, upper, lower, psi, wA);
            for (int32_t i = 0; i < n; ++i) rA[i] = source[i] - wA[i];
            /* :342-361 */
            if (deflate)
            {
                double ur = 0, uu = 0;
                for (int32_t i = 0; i < n; ++i)
                {
                    ur += u[i] * rA[i];
                    uu += u[i] * u[i];
                }
                if (uu > ORC_VSMALL)
                {
                    const double cc = ur / uu;
                    for (int32_t i = 0; i < n; ++i)
                    {
                        psi[i] += cc;
                        rA[i] -= cc * u[i];
                    }
                }
            }
            /* :363-365 */
            perf->finalResidual = sum_mag(n, rA) / normFactor;
            perf->nIterations = totalIter;
            /* :371-391 */
            if (deflate && perf->finalResidual > prevRes / 1.5 && !st.singular)
            {
                deflate = 0;
                perf->polished = 1;
                set_coeffs_f(&c, diag, upper, NULL);
            }
            prevRes = perf->finalResidual;
            /* :404-414 */
            if (singular) break;
            if (totalIter >= ctl->maxIter) break;
            if (perf->nIterations >= ctl->minIter &&
                check_convergence(perf->finalResidual, perf->initialResidual,
                                  ctl->tolerance, ctl->relTol))
                break;
        }
        perf->restarts = outer;
        perf->singular = singular;
        free(c.diag);
        free(c.rD);
        free(c.upper);
        free(c.lower);
    }
    perf->converged = check_convergence(perf->finalResidual,
                                        perf->initialResidual, ctl->tolerance,
                                        ctl->relTol);
    free(wA);
    free(rA);
    free(u);
    free(diag0);
    free(rin);
    free(delta);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* OpenFOAM built-in solvers in FP64 (CPU baseline; plain-FP64 comparator)   */
/* ------------------------------------------------------------------------ */

enum
{
    ORC_PRE_DIAGONAL = 0,
    ORC_PRE_DIC = 1, /* DILU for PBiCGStab */
    ORC_PRE_CHEBYSHEV = 2
};

typedef struct
{
    const orc_addr* a;
    const double *diag, *upper, *lower;
    int kind;
    double* rD;
    int order;
    double beta, ratio;
    double *w, *v;
} orc_precon;

static void precon_init(orc_precon* P)
{
    const orc_addr* a = P->a;
    const int32_t n = a->n;
    P->rD = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    P->w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    P->v = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) P->rD[i] = P->diag[i];
    if (P->kind == ORC_PRE_DIC)
    {
        const double* lo = P->lower ? P->lower : P->upper;
        /* DIC/DILU calcReciprocalD: rD[u] -= upper*lower/rD[l] */
        for (int32_t f = 0; f < a->nf; ++f)
            P->rD[a->upp[f]] -= P->upper[f] * lo[f] / P->rD[a->low[f]];
    }
    for (int32_t i = 0; i < n; ++i) P->rD[i] = 1.0 / P->rD[i];
}

static void precon_free(orc_precon* P)
{
    free(P->rD);
    free(P->w);
    free(P->v);
}

static void precon_apply(const orc_precon* P, double* wA, const double* rA)
{
    const orc_addr* a = P->a;
    const int32_t n = a->n;
    if (P->kind == ORC_PRE_CHEBYSHEV)
    {
        orc_cheb_apply_d(a, P->diag, P->upper, P->lower, P->rD, P->order,
                         P->beta, P->ratio, rA, wA, P->w, P->v);
        return;
    }
    for (int32_t i = 0; i < n; ++i) wA[i] = P->rD[i] * rA[i];
    if (P->kind != ORC_PRE_DIC) return;
    if (!P->lower)
    {
        /* DIC: forward over faces, backward over faces reversed */
        for (int32_t f = 0; f < a->nf; ++f)
            wA[a->upp[f]] -= P->rD[a->upp[f]] * P->upper[f] * wA[a->low[f]];
        for (int32_t f = a->nf - 1; f >= 0; --f)
            wA[a->low[f]] -= P->rD[a->low[f]] * P->upper[f] * wA[a->upp[f]];
    }
    else
    {
        /* DILU: forward in losort order with lower, backward with upper */
        for (int32_t k = 0; k < a->nf; ++k)
        {
            const int32_t f = a->losort[k];
            wA[a->upp[f]] -= P->rD[a->upp[f]] * P->lower[f] * wA[a->low[f]];
        }
        for (int32_t f = a->nf - 1; f >= 0; --f)
            wA[a->low[f]] -= P->rD[a->low[f]] * P->upper[f] * wA[a->upp[f]];
    }
}

static double dotd(int32_t n, const double* x, const double* y)
{
    double s = 0;
    for (int32_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

/* OpenFOAM PCG::scalarSolve.  precon: 0 diagonal, 1 DIC, 2 Chebyshev-Jacobi
 * (order, beta, ratio as the GPU comparator uses). */
int orc_pcg(const orc_addr* a, const double* diag, const double* upper,
            int precon, int order, double beta, double ratio, double tol,
            double relTol, int maxIter, int minIter, double* psi,
            const double* source, orc_perf* perf)
{
    const int32_t n = a->n;
    const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    memset(perf, 0, sizeof(*perf));
    double* wA = (double*)malloc(nb);
    double* rA = (double*)malloc(nb);
    double* pA = (double*)malloc(nb);
    orc_amul(a, diag, upper, NULL, psi, wA);
    for (int32_t i = 0; i < n; ++i) rA[i] = source[i] - wA[i];
    const double nf = orc_norm_factor(a, diag, upper, NULL, psi, source, wA);
    perf->normFactor = nf;
    perf->initialResidual = sum_mag(n, rA) / nf;
    perf->finalResidual = perf->initialResidual;
    int nIter = 0;
    if (minIter > 0 || !check_convergence(perf->finalResidual,
                                          perf->initialResidual, tol, relTol))
    {
        orc_precon P;
        memset(&P, 0, sizeof(P));
        P.a = a;
        P.diag = diag;
        P.upper = upper;
        P.lower = NULL;
        P.kind = precon;
        P.order = order;
        P.beta = beta;
        P.ratio = ratio;
        precon_init(&P);
        double wArA = 1e300, wArAold;
        do
        {
            wArAold = wArA;
            precon_apply(&P, wA, rA);
            wArA = dotd(n, wA, rA);
            if (nIter == 0)
            {
                for (int32_t i = 0; i < n; ++i) pA[i] = wA[i];
            }
            else
            {
                const double b = wArA / wArAold;
                for (int32_t i = 0; i < n; ++i) pA[i] = wA[i] + b * pA[i];
            }
            orc_amul(a, diag, upper, NULL, pA, wA);
            const double wApA = dotd(n, wA, pA);
            if (fabs(wApA) / nf < ORC_VSMALL)
            {
                perf->singular = 1;
                break;
            }
            const double alpha = wArA / wApA;
            for (int32_t i = 0; i < n; ++i)
            {
                psi[i] += alpha * pA[i];
                rA[i] -= alpha * wA[i];
            }
            perf->finalResidual = sum_mag(n, rA) / nf;
        } while (((++nIter < maxIter) &&
                  !check_convergence(perf->finalResidual,
                                     perf->initialResidual, tol, relTol)) ||
                 nIter < minIter);
        precon_free(&P);
    }
    perf->nIterations = nIter;
    perf->converged = check_convergence(perf->finalResidual,
                                        perf->initialResidual, tol, relTol);
    free(wA);
    free(rA);
    free(pA);
    return 0;
}

/* OpenFOAM PBiCGStab::scalarSolve.  precon: 0 diagonal, 1 DILU. */
int orc_pbicgstab(const orc_addr* a, const double* diag, const double* upper,
                  const double* lower, int precon, double tol, double relTol,
                  int maxIter, int minIter, double* psi, const double* source,
                  orc_perf* perf)
{
    const int32_t n = a->n;
    const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    memset(perf, 0, sizeof(*perf));
    double* yA = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* rA = (double*)malloc(nb);
    double* pA = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* AyA = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* rA0 = (double*)malloc(nb);
    double* sA = (double*)malloc(nb);
    double* zA = (double*)malloc(nb);
    double* tA = (double*)malloc(nb);
    orc_amul(a, diag, upper, lower, psi, yA);
    for (int32_t i = 0; i < n; ++i) rA[i] = source[i] - yA[i];
    const double nf = orc_norm_factor(a, diag, upper, lower, psi, source, yA);
    perf->normFactor = nf;
    perf->initialResidual = sum_mag(n, rA) / nf;
    perf->finalResidual = perf->initialResidual;
    int nIter = 0;
    if (minIter > 0 || !check_convergence(perf->finalResidual,
                                          perf->initialResidual, tol, relTol))
    {
        orc_precon P;
        memset(&P, 0, sizeof(P));
        P.a = a;
        P.diag = diag;
        P.upper = upper;
        P.lower = lower;
        P.kind = precon;
        precon_init(&P);
        double rA0rA = 0, alpha = 0, omega = 0;
        for (int32_t i = 0; i < n; ++i) rA0[i] = rA[i];
        do
        {
            const double rA0rAold = rA0rA;
            rA0rA = dotd(n, rA0, rA);
            if (fabs(rA0rA) < ORC_VSMALL)
            {
                perf->singular = 1;
                break;
            }
            if (nIter == 0)
            {
                for (int32_t i = 0; i < n; ++i) pA[i] = rA[i];
            }
            else
            {
                if (fabs(omega) < ORC_VSMALL)
                {
                    perf->singular = 1;
                    break;
                }
                const double beta = (rA0rA / rA0rAold) * (alpha / omega);
                for (int32_t i = 0; i < n; ++i)
                    pA[i] = rA[i] + beta * (pA[i] - omega * AyA[i]);
            }
            precon_apply(&P, yA, pA);
            orc_amul(a, diag, upper, lower, yA, AyA);
            const double rA0AyA = dotd(n, rA0, AyA);
            alpha = rA0rA / rA0AyA;
            for (int32_t i = 0; i < n; ++i) sA[i] = rA[i] - alpha * AyA[i];
            perf->finalResidual = sum_mag(n, sA) / nf;
            if (check_convergence(perf->finalResidual, perf->initialResidual,
                                  tol, relTol))
            {
                for (int32_t i = 0; i < n; ++i) psi[i] += alpha * yA[i];
                nIter++;
                break;
            }
            precon_apply(&P, zA, sA);
            orc_amul(a, diag, upper, lower, zA, tA);
            const double tAtA = dotd(n, tA, tA);
            omega = dotd(n, tA, sA) / tAtA;
            for (int32_t i = 0; i < n; ++i)
            {
                psi[i] += alpha * yA[i] + omega * zA[i];
                rA[i] = sA[i] - omega * tA[i];
            }
            perf->finalResidual = sum_mag(n, rA) / nf;
        } while (((++nIter < maxIter) &&
                  !check_convergence(perf->finalResidual,
                                     perf->initialResidual, tol, relTol)) ||
                 nIter < minIter);
        precon_free(&P);
    }
    perf->nIterations = nIter;
    perf->converged = check_convergence(perf->finalResidual,
                                        perf->initialResidual, tol, relTol);
    free(yA);
    free(rA);
    free(pA);
    free(AyA);
    free(rA0);
    free(sA);
    free(zA);
    free(tA);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoint gradients (no reference counterpart: new in this build)           */
/* ------------------------------------------------------------------------ */

/* For A x = b and a loss L(x) with adjoint A^T lambda = dL/dx:
 *   dL/db = lambda, dL/ddiag_i = -lambda_i x_i,
 *   dL/dupper_f = -lambda[l] x[u]   (upper_f = A[l,u])
 *   dL/dlower_f = -lambda[u] x[l]   (lower_f = A[u,l]). */
void orc_adjoint_grads(const orc_addr* a, const double* x, const double* lam,
                       double* dUpper, double* dLower, double* dDiag)
{
    for (int32_t i = 0; i < a->n; ++i) dDiag[i] = -lam[i] * x[i];
    for (int32_t f = 0; f < a->nf; ++f)
    {
        const int32_t l = a->low[f], u = a->upp[f];
        dUpper[f] = -lam[l] * x[u];
        dLower[f] = -lam[u] * x[l];
    }
}

/* ------------------------------------------------------------------------ */
/* CPU baseline on all host cores: OpenFOAM PCG(diagonal) in FP64, OpenMP    */
/* ------------------------------------------------------------------------ */

/* y = A x by rows (each row gathers its owner faces through ownStart and its
 * neighbour faces through losortStart/losort), so the rows are independent
 * and split over threads without atomics.  Same terms as orc_amul; the sum
 * order within a row differs from the face loop's. */
static void amul_rows_par(const orc_addr* a, const double* diag,
                          const double* upper, const double* lower,
                          const double* x, double* y)
{
    if (!lower) lower = upper;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < a->n; ++i)
    {
        double s = diag[i] * x[i];
        for (int32_t f = a->ownStart[i]; f < a->ownStart[i + 1]; ++f)
            s += upper[f] * x[a->upp[f]];
        for (int32_t k = a->losortStart[i]; k < a->losortStart[i + 1]; ++k)
        {
            const int32_t f = a->losort[k];
            s += lower[f] * x[a->low[f]];
        }
        y[i] = s;
    }
}

static double dot_par(int32_t n, const double* x, const double* y)
{
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int32_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

static double summag_par(int32_t n, const double* x)
{
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int32_t i = 0; i < n; ++i) s += fabs(x[i]);
    return s;
}

/* OpenFOAM PCG::scalarSolve with the diagonal preconditioner (the loop of
 * orc_pcg, precon 0), every O(n) step split over `nthreads` OpenMP threads
 * (0: the runtime's default).  Built without OpenMP it runs on one thread. */
int orc_pcg_diag_par(const orc_addr* a, const double* diag,
                     const double* upper, double tol, double relTol,
                     int maxIter, int minIter, int nthreads, double* psi,
                     const double* source, orc_perf* perf)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int32_t n = a->n;
    const size_t nb = sizeof(double) * (size_t)(n > 0 ? n : 1);
    memset(perf, 0, sizeof(*perf));
    double* wA = (double*)malloc(nb);
    double* rA = (double*)malloc(nb);
    double* pA = (double*)malloc(nb);
    double* rD = (double*)malloc(nb);
    double* u = (double*)malloc(nb);
    double* ones = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    /* wA = A psi, rA = b - wA; normFactor = sum(|A psi - xbar u| + |b - xbar u|)
     * + small with u = A 1 and xbar = mean(psi) */
    amul_rows_par(a, diag, upper, NULL, psi, wA);
    double xbar = 0;
#pragma omp parallel for schedule(static) reduction(+ : xbar)
    for (int32_t i = 0; i < n; ++i)
    {
        rA[i] = source[i] - wA[i];
        ones[i] = 1.0;
        xbar += psi[i];
    }
    xbar = n > 0 ? xbar / n : 0.0;
    amul_rows_par(a, diag, upper, NULL, ones, u);
    double nfs = 0;
#pragma omp parallel for schedule(static) reduction(+ : nfs)
    for (int32_t i = 0; i < n; ++i)
        nfs += fabs(wA[i] - xbar * u[i]) + fabs(source[i] - xbar * u[i]);
    const double nf = nfs + ORC_SMALL;
    perf->normFactor = nf;
    perf->initialResidual = summag_par(n, rA) / nf;
    perf->finalResidual = perf->initialResidual;
    int nIter = 0;
    if (minIter > 0 || !check_convergence(perf->finalResidual,
                                          perf->initialResidual, tol, relTol))
    {
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) rD[i] = 1.0 / diag[i];
        double wArA = 1e300, wArAold;
        do
        {
            wArAold = wArA;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) wA[i] = rD[i] * rA[i];
            wArA = dot_par(n, wA, rA);
            if (nIter == 0)
            {
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < n; ++i) pA[i] = wA[i];
            }
            else
            {
                const double b = wArA / wArAold;
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < n; ++i) pA[i] = wA[i] + b * pA[i];
            }
            amul_rows_par(a, diag, upper, NULL, pA, wA);
            const double wApA = dot_par(n, wA, pA);
            if (fabs(wApA) / nf < ORC_VSMALL)
            {
                perf->singular = 1;
                break;
            }
            const double alpha = wArA / wApA;
            double rs = 0;
#pragma omp parallel for schedule(static) reduction(+ : rs)
            for (int32_t i = 0; i < n; ++i)
            {
                psi[i] += alpha * pA[i];
                rA[i] -= alpha * wA[i];
                rs += fabs(rA[i]);
            }
            perf->finalResidual = rs / nf;
        } while (((++nIter < maxIter) &&
                  !check_convergence(perf->finalResidual,
                                     perf->initialResidual, tol, relTol)) ||
                 nIter < minIter);
    }
    perf->nIterations = nIter;
    perf->converged = check_convergence(perf->finalResidual,
                                        perf->initialResidual, tol, relTol);
    free(wA);
    free(rA);
    free(pA);
    free(rD);
    free(u);
    free(ones);
    return 0;
}
