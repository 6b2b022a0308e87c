/*
 * oracle.c -- CPU restatement of OPLib's operator-method pricing path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, by smoke() and by
 * bench.py's cpu_baseline leg.  The shipped libraries never link or load it.
 * See oracle.h for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp, x86-64-v3).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* lattice (A1)                                                         */
/* ------------------------------------------------------------------ */

/* Uniform grid on [xlo,xhi] with nx nodes, then the reference's spot
 * rescaling: find the node nearest to xval0 and scale the whole grid so that
 * node lands exactly on xval0 (SGrid.cs:540-557; first-minimum tie rule of
 * SGrid.cs:541-551).  For a uniform pivot set grid_iterate (SGrid.cs:561-592)
 * reproduces the uniform spacing, so only the rescale matters here. */
void or_xgrid_uniform(int nx, double xlo, double xhi, double xval0,
                      double *xval, int *x0)
{
    double h = (xhi - xlo) / (double)(nx - 1);
    for (int x = 0; x < nx; x++) xval[x] = xlo + h * (double)x;
    int best = 0;
    double minerr = fabs(xval[0] - xval0);
    for (int x = 1; x < nx; x++) {
        double e = fabs(xval[x] - xval0);
        if (e < minerr) { best = x; minerr = e; }
    }
    double scale = xval0 / xval[best];
    for (int x = 0; x < nx; x++) xval[x] *= scale;
    *x0 = best;
}

/* Per-node inverse of the 2x2 stencil matrix (S1DGrid.cs:78-147):
 *   mat = [ D-   D+  ]    D- = x[i-1]-x[i],  D+ = x[i+1]-x[i]
 *         [ D-^2 D+^2]
 * with the boundary substitutions of S1DGrid.cs:91-125, then invm = mat^-1
 * (the reference calls DGESV with B = I).  Stored column-major per node:
 * invm[4*x + r + 2*c] = (mat^-1)(r,c). */
void or_stencil_invm(int nx, const double *xval, double *invm)
{
    for (int x0 = 0; x0 < nx; x0++) {
        double m00, m01, m10, m11;
        m00 = (x0 > 0) ? xval[x0 - 1] - xval[x0] : 1.0;
        m01 = (x0 < nx - 1) ? xval[x0 + 1] - xval[x0] : 1.0;
        if (x0 > 0 && x0 < nx - 1) {
            m10 = m00 * m00;
            m11 = m01 * m01;
        } else if (x0 == 0) {
            m10 = 0.0; m11 = 1.0;
        } else { /* x0 == nx-1 */
            m10 = m00 * m00;  /* x0>0 && x0<d branch of S1DGrid.cs:115 */
            m11 = m01 * m01;
        }
        /* S1DGrid.cs:115 tests (x0 > 0 && x0 < d), which is true for the
         * last node too, so the last node keeps the squared entries; its
         * generator row is zero anyway (absorbing). */
        double det = m00 * m11 - m01 * m10;
        invm[4 * x0 + 0] = m11 / det;
        invm[4 * x0 + 1] = -m10 / det;
        invm[4 * x0 + 2] = -m01 / det;
        invm[4 * x0 + 3] = m00 / det;
    }
}

/* LV generator (CLVModel.cs:86-136; device twin sglvg.cu:42-99).  Rows 0 and
 * d-1 are zero (absorbing).  Interior row y0: rhs = (drift, vol^2),
 * rate(y0->y0-1) = invm(0,:).rhs, rate(y0->y0+1) = invm(1,:).rhs,
 * diagonal = -(sum). */
void or_lv_generator(int d, const double *invm, const double *drift,
                     const double *vol, double *gen)
{
    memset(gen, 0, sizeof(double) * (size_t)d * d);
    for (int x0 = 0; x0 < d; x0++) {
        double r0 = 0.0, r1 = 0.0;
        if (x0 > 0 && x0 < d - 1) { r0 = drift[x0]; r1 = vol[x0] * vol[x0]; }
        double sum = 0.0, xi;
        if (x0 > 0) {
            xi = invm[4 * x0 + 0] * r0 + invm[4 * x0 + 2] * r1;
            gen[x0 + (size_t)d * (x0 - 1)] = xi;
            sum += xi;
        }
        if (x0 < d - 1) {
            xi = invm[4 * x0 + 1] * r0 + invm[4 * x0 + 3] * r1;
            gen[x0 + (size_t)d * (x0 + 1)] = xi;
            sum += xi;
        }
        gen[x0 + (size_t)d * x0] = -sum;
    }
}

/* Regime-switching generator on y = x + nx*r (state order of
 * S2DGrid.cs:89 / SGrid.cs:471-474):
 *   L = blockdiag_r( L_x(sigma_r) ) + Q (x) I_nx,
 * L_x(sigma_r) is the LV generator with drift rate*x and absolute vol
 * sigma_r*x; Q has off-diagonal q_off and a diagonal making rows sum to 0.
 * Not the reference's SV dynamics (CSVModel.cs:187-379); it keeps only the
 * reference's state ordering and its "diagonal = -sum" convention. */
void or_rs_generator(int nx, int nr, const double *invm, const double *xval,
                     const double *sigma_r, double rate, double q_off,
                     double *gen)
{
    int d = nx * nr;
    memset(gen, 0, sizeof(double) * (size_t)d * d);
    for (int r = 0; r < nr; r++) {
        for (int x0 = 0; x0 < nx; x0++) {
            int y0 = x0 + nx * r;
            double r0 = 0.0, r1 = 0.0;
            if (x0 > 0 && x0 < nx - 1) {
                double v = sigma_r[r] * xval[x0];
                r0 = rate * xval[x0];
                r1 = v * v;
            }
            double sum = 0.0, xi;
            if (x0 > 0) {
                xi = invm[4 * x0 + 0] * r0 + invm[4 * x0 + 2] * r1;
                gen[y0 + (size_t)d * (y0 - 1)] = xi;
                sum += xi;
            }
            if (x0 < nx - 1) {
                xi = invm[4 * x0 + 1] * r0 + invm[4 * x0 + 3] * r1;
                gen[y0 + (size_t)d * (y0 + 1)] = xi;
                sum += xi;
            }
            for (int r2 = 0; r2 < nr; r2++) {
                if (r2 == r) continue;
                gen[y0 + (size_t)d * (x0 + nx * r2)] = q_off;
                sum += q_off;
            }
            gen[y0 + (size_t)d * y0] = -sum;
        }
    }
}

/* Config-1 local vol: sigma(x) = clip(0.25*(x/100)^-0.5, 0.05, 1), absolute
 * vol sigma(x)*x (SVol is in price units, CLVModel.cs:109), drift rate*x. */
void or_lv_params_c1(int nx, const double *xval, double rate, double *drift,
                     double *vol)
{
    for (int x = 0; x < nx; x++) {
        double s = (xval[x] > 0.0) ? 0.25 * pow(xval[x] / 100.0, -0.5) : 1.0;
        if (s < 0.05) s = 0.05;
        if (s > 1.0) s = 1.0;
        drift[x] = rate * xval[x];
        vol[x] = s * xval[x];
    }
}

/* CEV-type local vol with skew/curvature (calibration sweep and surrogate
 * configurations): sigma(x) = sigma0*(x/100)^(beta-1)
 *   * (1 + skew*ln(x/100) + curv*ln^2(x/100)), floored at 1e-4 so that the
 * variance stays positive; x = 0 uses sigma = 0 (absorbing row anyway). */
void or_lv_params_cev(int nx, const double *xval, double rate, double sigma0,
                      double beta, double skew, double curv, double *drift,
                      double *vol)
{
    for (int x = 0; x < nx; x++) {
        double s = 0.0;
        if (xval[x] > 0.0) {
            double lx = log(xval[x] / 100.0);
            s = sigma0 * exp((beta - 1.0) * lx) * (1.0 + skew * lx + curv * lx * lx);
            if (s < 1e-4) s = 1e-4;
        }
        drift[x] = rate * xval[x];
        vol[x] = s * xval[x];
    }
}

/* ------------------------------------------------------------------ */
/* planner (A4)                                                         */
/* ------------------------------------------------------------------ */

/* CModel Common.cs:289-307 in years: dt = min(0.5/maxdiag, dt0),
 * niter = ceil(log2(DeltaT/dt)), dt_j = DeltaT * 2^-niter. */
int or_plan_niter(double DeltaT, double maxdiag, double dt0, double *dt_out)
{
    double dt = 0.5 / maxdiag;
    if (dt > dt0) dt = dt0;
    int niter = (int)ceil(log(DeltaT / dt) / log(2.0));
    *dt_out = DeltaT * pow(2.0, -(double)niter);
    return niter;
}

double or_max_abs_diag(int d, const double *gen)
{
    double m = 0.0;
    for (int y = 0; y < d; y++) {
        double a = fabs(gen[y + (size_t)d * y]);
        if (a > m) m = a;
    }
    return m;
}

/* ------------------------------------------------------------------ */
/* first step (A5)                                                      */
/* ------------------------------------------------------------------ */

/* U0 = I + dt*L in FP64 (host_d_mc_ek, CModel Host Double.cs:89-121). */
void or_first_step(int d, const double *gen, double dt, double *U)
{
    size_t n = (size_t)d * d;
    for (size_t e = 0; e < n; e++) U[e] = dt * gen[e];
    for (int y = 0; y < d; y++) U[y + (size_t)d * y] += 1.0;
}

/* FP32 first step exactly as global_sek (sek.cu:40-59): id + dt*gen, with
 * the multiply-add fused (nvcc and hipcc both contract it into one FMA). */
void or_sek_f32(int d, const float *gen, float dt, float *U)
{
    for (int y2 = 0; y2 < d; y2++)
        for (int y1 = 0; y1 < d; y1++) {
            size_t yy = (size_t)y1 + (size_t)d * y2;
            U[yy] = fmaf(dt, gen[yy], (y1 == y2) ? 1.0f : 0.0f);
        }
}

/* global_sekd / global_sekd_lambda (sek.cu:62-103): (id + dt gen) /
 * (1 + dt r[y1]) and / (1 + dt r[y1] lambda), FP32, with the contraction
 * of a product into the add that follows it (one FMA each, as sek_f32)
 * and an IEEE division.  lambda = NULL: the sekd form. */
void or_sekd_f32(int d, const float *gen, float dt, const float *r, const float *lambda, float *U)
{
    for (int y2 = 0; y2 < d; y2++)
        for (int y1 = 0; y1 < d; y1++) {
            size_t yy = (size_t)y1 + (size_t)d * y2;
            float num = fmaf(dt, gen[yy], (y1 == y2) ? 1.0f : 0.0f);
            float den = lambda ? fmaf(dt * r[y1], lambda[0], 1.0f) : fmaf(dt, r[y1], 1.0f);
            U[yy] = num / den;
        }
}

/* global_sglvg (sglvg.cu:42-99), FP32: ni generators gen[i d^2 + y0 + d y1]
 * zeroed, then for interior rows y0 (0 < y0 < d-1) with rhs = (drift,
 * vol^2): rate down xi = 0 + invm[4 y0] drift, then + invm[4 y0 + 2] vol^2
 * (each product contracted into its add: the second step is one FMA), rate
 * up likewise with invm[4 y0 + 1] and invm[4 y0 + 3], diagonal -(down + up).
 * pars: drift (ni d) then vol (ni d). */
void or_sglvg_f32(int ni, int d, const float *invm, const float *pars, float *gen)
{
    memset(gen, 0, sizeof(float) * (size_t)ni * d * d);
    for (int i = 0; i < ni; i++)
        for (int y0 = 1; y0 < d - 1; y0++) {
            const float r0 = pars[y0 + (size_t)d * i];
            const float v = pars[(size_t)ni * d + y0 + (size_t)d * i];
            const float r1 = v * v;
            const float lo = fmaf(invm[4 * y0 + 2], r1, invm[4 * y0 + 0] * r0);
            const float hi = fmaf(invm[4 * y0 + 3], r1, invm[4 * y0 + 1] * r0);
            float *g = gen + (size_t)i * d * d;
            g[y0 + (size_t)d * (y0 - 1)] = lo;
            g[y0 + (size_t)d * (y0 + 1)] = hi;
            g[y0 + (size_t)d * y0] = -(lo + hi);
        }
}

/* ------------------------------------------------------------------ */
/* GEMM (A6-A9)                                                         */
/* ------------------------------------------------------------------ */

static int g_nth = 0;
int or_set_threads(int nth)
{
    int old = g_nth;
    g_nth = nth;
    return old;
}
static int nthreads(void)
{
#ifdef _OPENMP
    return g_nth > 0 ? g_nth : omp_get_max_threads();
#else
    return 1;
#endif
}

typedef double v4d __attribute__((vector_size(32)));

/* C = A*B, column-major FP64, the contraction host_d_mc_sq/host_d_mc_mk hand
 * to MKL DGEMM (linalg.c:93-101, CModel Host Double.cs:132-277).  Register
 * tile 8x4 over a packed A panel, OpenMP over column blocks of C. */
void or_dgemm(int m, int n, int k, const double *A, int lda, const double *B,
              int ldb, double *C, int ldc)
{
    const int MB = 128, KB = 256;
#pragma omp parallel num_threads(nthreads())
    {
        double *Ap = (double *)aligned_alloc(64, sizeof(double) * MB * KB);
#pragma omp for schedule(static)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) C[i + (size_t)ldc * j] = 0.0;
        for (int kb = 0; kb < k; kb += KB) {
            int kc = (k - kb < KB) ? k - kb : KB;
            for (int ib = 0; ib < m; ib += MB) {
                int mc = (m - ib < MB) ? m - ib : MB;
                int mcp = (mc + 7) & ~7;
                /* every thread packs its own copy: cheap and barrier-free */
                for (int kk = 0; kk < kc; kk++) {
                    const double *src = A + ib + (size_t)lda * (kb + kk);
                    double *dst = Ap + (size_t)mcp * kk;
                    int i = 0;
                    for (; i < mc; i++) dst[i] = src[i];
                    for (; i < mcp; i++) dst[i] = 0.0;
                }
#pragma omp for schedule(static)
                for (int jb = 0; jb < n; jb += 4) {
                    int nc = (n - jb < 4) ? n - jb : 4;
                    for (int i = 0; i < mcp; i += 8) {
                        v4d c00 = {0}, c01 = {0}, c10 = {0}, c11 = {0};
                        v4d c20 = {0}, c21 = {0}, c30 = {0}, c31 = {0};
                        const double *b0 = B + (kb) + (size_t)ldb * (jb + 0);
                        const double *b1 = B + (kb) + (size_t)ldb * (jb + (nc > 1 ? 1 : 0));
                        const double *b2 = B + (kb) + (size_t)ldb * (jb + (nc > 2 ? 2 : 0));
                        const double *b3 = B + (kb) + (size_t)ldb * (jb + (nc > 3 ? 3 : 0));
                        for (int kk = 0; kk < kc; kk++) {
                            v4d a0, a1;
                            memcpy(&a0, Ap + (size_t)mcp * kk + i, 32);
                            memcpy(&a1, Ap + (size_t)mcp * kk + i + 4, 32);
                            double s0 = b0[kk], s1 = b1[kk], s2 = b2[kk], s3 = b3[kk];
                            c00 += a0 * s0; c01 += a1 * s0;
                            c10 += a0 * s1; c11 += a1 * s1;
                            c20 += a0 * s2; c21 += a1 * s2;
                            c30 += a0 * s3; c31 += a1 * s3;
                        }
                        v4d cc[4][2] = {{c00, c01}, {c10, c11}, {c20, c21}, {c30, c31}};
                        for (int jj = 0; jj < nc; jj++) {
                            double *cp = C + ib + i + (size_t)ldc * (jb + jj);
                            int lim = mc - i < 8 ? mc - i : 8;
                            double t[8];
                            memcpy(t, &cc[jj][0], 32);
                            memcpy(t + 4, &cc[jj][1], 32);
                            for (int q = 0; q < lim; q++) cp[q] += t[q];
                        }
                    }
                }
            }
        }
        free(Ap);
    }
}

/* FP32 operands, FP64 accumulation: the tolerance oracle for the FP32 GEMM
 * entry points (cublasSgemm semantics of sgemm.cu:41-44 with alpha/beta). */
void or_sgemm_f64acc(int m, int n, int k, float alpha, const float *A, int lda,
                     const float *B, int ldb, float beta, float *C, int ldc)
{
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int j = 0; j < n; j++) {
        double *acc = (double *)calloc((size_t)m, sizeof(double));
        for (int kk = 0; kk < k; kk++) {
            double b = B[kk + (size_t)ldb * j];
            const float *a = A + (size_t)lda * kk;
            for (int i = 0; i < m; i++) acc[i] += (double)a[i] * b;
        }
        for (int i = 0; i < m; i++) {
            double c = (beta != 0.0f) ? (double)beta * C[i + (size_t)ldc * j] : 0.0;
            C[i + (size_t)ldc * j] = (float)((double)alpha * acc[i] + c);
        }
        free(acc);
    }
}

/* nsq repeated squarings U <- U*U (device_mc_sfex, CModel Device.cs:216-228;
 * host_d_mc_sq, CModel Host Double.cs:132-172).  work is d*d scratch. */
void or_square_chain(int d, int nsq, double *U, double *work)
{
    for (int r = 0; r < nsq; r++) {
        or_dgemm(d, d, d, U, d, U, d, work, d);
        memcpy(U, work, sizeof(double) * (size_t)d * d);
    }
}

/* Divide every row by its sum (probability-mass renormalisation). */
void or_row_renorm(int d, double *U)
{
    double *s = (double *)calloc((size_t)d, sizeof(double));
    for (int c = 0; c < d; c++)
        for (int r = 0; r < d; r++) s[r] += U[r + (size_t)d * c];
    for (int c = 0; c < d; c++)
        for (int r = 0; r < d; r++) U[r + (size_t)d * c] /= s[r];
    free(s);
}

/* ------------------------------------------------------------------ */
/* gather GEMV (A10)                                                    */
/* ------------------------------------------------------------------ */

/* global_sgemv3 / global_sgemv4 semantics (sgemv3.cu:92-159,
 * sgemv4.cu:90-170): for every listed column c,
 * Z[col1[c] + i] = sum_k A[i,k] * Z[col0[c] + k], i < d.  Column offsets are
 * element offsets into Z.  The list is read in groups of 16 (one thread
 * block's columns); inside a group the stores stop at the first
 * col1 >= sentinel (sgemv3.cu:141-157, sgemv4.cu:155-170: `return` at the
 * first sentinel), the next group starts afresh.  All reads happen before
 * any write.  FP64 accumulation (the checker of an FP32 kernel). */
void or_sgemv_gather(int d, const float *A, float *Z, int ncol,
                     const int *col0, const int *col1, int sentinel)
{
    double *out = (double *)calloc((size_t)d * (ncol > 0 ? ncol : 1), sizeof(double));
    char *live = (char *)calloc((size_t)(ncol > 0 ? ncol : 1), 1);
    for (int g0 = 0; g0 < ncol; g0 += 16) {
        for (int c = g0; c < ncol && c < g0 + 16; c++) {
            if (col1[c] >= sentinel) break;
            live[c] = 1;
            for (int kk = 0; kk < d; kk++) {
                double z = Z[col0[c] + kk];
                const float *a = A + (size_t)d * kk;
                for (int i = 0; i < d; i++) out[(size_t)d * c + i] += (double)a[i] * z;
            }
        }
    }
    for (int c = 0; c < ncol; c++)
        if (live[c])
            for (int i = 0; i < d; i++) Z[col1[c] + i] = (float)out[(size_t)d * c + i];
    free(out);
    free(live);
}

/* Backward induction in FP64: V <- df * U * V, nsteps times, V is d x ncol. */
void or_backward_induction(int d, const double *U, int nsteps, double df,
                           int ncol, double *V)
{
    double *W = (double *)malloc(sizeof(double) * (size_t)d * ncol);
    for (int s = 0; s < nsteps; s++) {
        or_dgemm(d, ncol, d, U, d, V, d, W, d);
        for (size_t e = 0; e < (size_t)d * ncol; e++) V[e] = df * W[e];
    }
    free(W);
}

/* ------------------------------------------------------------------ */
/* cumulative tables (A11)                                              */
/* ------------------------------------------------------------------ */

/* global_sck (sck.cu:43-84): per (m,row) a sequential FP32 running sum of
 * the row's raw entries (stride d; a negative entry lowers the sum, as in
 * the reference), then every entry divided by the row total and capped at
 * 1 (sck.cu:74-80). */
void or_sck_f32(int d, int nm, const float *tpk, float *ctpk)
{
    size_t d2 = (size_t)d * d;
    for (int m = 0; m < nm; m++)
        for (int row = 0; row < d; row++) {
            const float *t = tpk + m * d2 + row;
            float *c = ctpk + m * d2 + row;
            float s = 0.0f;
            for (int y = 0; y < d; y++) {
                s += t[(size_t)d * y];
                c[(size_t)d * y] = s;
            }
            for (int y = 0; y < d; y++) {
                float v = c[(size_t)d * y] / s;
                c[(size_t)d * y] = (v < 1.0f) ? v : 1.0f;
            }
        }
}

/* The same table stored row-contiguous, crow[m*d*d + row*d + y], with
 * entries below zero taken as zero (a rounding-level negative would make
 * the row non-monotone and the sampler's bisection ill-defined): the
 * sequential order of or_sck_f32, so on nonnegative tables this is that
 * table transposed, bit for bit. */
void or_sck_rows_f32(int d, int nm, const float *tpk, float *crow)
{
    size_t d2 = (size_t)d * d;
    for (int m = 0; m < nm; m++)
        for (int row = 0; row < d; row++) {
            const float *t = tpk + m * d2 + row;
            float *c = crow + m * d2 + (size_t)row * d;
            float s = 0.0f;
            for (int y = 0; y < d; y++) {
                float v = t[(size_t)d * y];
                s += (v > 0.0f) ? v : 0.0f;
                c[y] = s;
            }
            for (int y = 0; y < d; y++) {
                float v = c[y] / s;
                c[y] = (v < 1.0f) ? v : 1.0f;
            }
        }
}

/* The row-contiguous table in a fixed two-level order (the GPU's
 * ophip_sck_rows): each row is cut into chunks of `chunk` columns; chunk q
 * holds a sequential FP32 running sum l_q(y) of its clamped entries from 0
 * (plus_identity: 1 added to the diagonal entry first); the chunk totals
 * t_q = l_q(last) are chained in order, carry_0 = 0, carry_{q+1} =
 * carry_q + t_q; the running sum is S(y) = carry_q + l_q(y) and the row
 * total T = S(d-1); C(y) = min(1, S(y) / T), so C(d-1) = 1 exactly and rows
 * are non-decreasing.  Where every partial sum is exact (e.g. dyadic
 * entries) this equals the sequential order bit for bit. */
void or_sck_rows2_f32(int d, int nm, int chunk, int plus_identity, const float *tpk, float *crow)
{
    size_t d2 = (size_t)d * d;
    for (int m = 0; m < nm; m++)
        for (int row = 0; row < d; row++) {
            const float *t = tpk + m * d2 + row;
            float *c = crow + m * d2 + (size_t)row * d;
            float carry = 0.0f;
            for (int y0 = 0; y0 < d; y0 += chunk) {
                float l = 0.0f;
                for (int y = y0; y < d && y < y0 + chunk; y++) {
                    float v = t[(size_t)d * y];
                    if (plus_identity && y == row) v += 1.0f;
                    l += (v > 0.0f) ? v : 0.0f;
                    c[y] = carry + l;
                }
                carry = carry + l;
            }
            const float T = c[d - 1];
            for (int y = 0; y < d; y++) {
                float v = c[y] / T;
                c[y] = (v < 1.0f) ? v : 1.0f;
            }
        }
}

/* ------------------------------------------------------------------ */
/* MT19937 (A12, A14)                                                   */
/* ------------------------------------------------------------------ */
/* Matsumoto & Nishimura (1998), init_genrand / genrand_int32.  state[624]
 * holds the index. */

void or_mt19937_init(uint32_t seed, uint32_t *st)
{
    st[0] = seed;
    for (uint32_t i = 1; i < 624; i++)
        st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + i;
    st[624] = 624;
}

static void mt_twist(uint32_t *st)
{
    for (int i = 0; i < 624; i++) {
        uint32_t y = (st[i] & 0x80000000u) | (st[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = st[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        st[i] = v;
    }
    st[624] = 0;
}

uint32_t or_mt19937_next(uint32_t *st)
{
    if (st[624] >= 624) mt_twist(st);
    uint32_t y = st[st[624]++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* The build's uniform mapping: the top 24 bits scaled to [0,1), exact in
 * FP32.  (The reference GPU sampler maps ((float)x+1)/2^32 onto (0,1],
 * smc.cu:217, which can round to exactly 1.0; the CPU sampler uses
 * x/2^32 in [0,1), rangen.cpp:125.  [0,1) with strict "u < C" keeps the
 * inverse CDF exact.) */
float or_u01_from_bits(uint32_t x)
{
    return (float)(x >> 8) * 0x1p-24f;
}

void or_mt_uniforms(uint32_t seed, int n, float *out)
{
    uint32_t st[625];
    or_mt19937_init(seed, st);
    for (int i = 0; i < n; i++) out[i] = or_u01_from_bits(or_mt19937_next(st));
}

/* ------------------------------------------------------------------ */
/* chain sampler (A13)                                                  */
/* ------------------------------------------------------------------ */

/* Inverse CDF by definition: the smallest j with u < crow[j], clamped to
 * d-1 (the CPU sampler's result, rangen.cpp:128-149). */
int or_sample_linear(const float *crow, int d, float u)
{
    for (int j = 0; j < d; j++)
        if (u < crow[j]) return j;
    return d - 1;
}

int or_sample_colmajor(const float *ctpk_m, int d, int yprev, float u)
{
    for (int j = 0; j < d; j++)
        if (u < ctpk_m[yprev + (size_t)d * j]) return j;
    return d - 1;
}

/* Bisection with the same result as the linear scan on a non-decreasing
 * row: invariant crow[lo] <= u (lo=-1 virtual), u < crow[hi] or hi=d-1. */
static int sample_bisect(const float *base, size_t stride, int d, float u)
{
    int lo = -1, hi = d - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (u < base[stride * (size_t)mid]) hi = mid; else lo = mid;
    }
    return hi;
}

/* The reference's GPU search (smc.cu:224-238, 336-347, 447-460,
 * 560-572): lb = 0, ub = d-1, a fixed number of halvings (10 for
 * 512 < d <= 1024, 9 for 256 < d <= 512, 8 for 128 < d <= 256, 7 for
 * d <= 128), mid = (ub+lb)/2, ub = mid when u < C[mid] else lb = mid,
 * result ub.  Column-major table (C[yprev + d*j]).  The count of halvings
 * is the deepest path to ub = 0 for that size bucket, so the result is the
 * inverse CDF min{j : u < C[j]} clamped to d-1 (state 0 included: with
 * ub = 1, lb = 0 the last halving tests C[0]). */
int or_sample_ref_bisect(const float *ctpk_m, int d, int yprev, float u)
{
    int niter = d > 512 ? 10 : d > 256 ? 9 : d > 128 ? 8 : 7;
    int ub = d - 1, lb = 0;
    for (int it = 0; it < niter; it++) {
        int mid = (ub + lb) / 2;
        if (u < ctpk_m[yprev + (size_t)d * mid]) ub = mid;
        else lb = mid;
    }
    return ub;
}

/* The build's Monte Carlo mapping (one MT19937 stream per wavefront):
 * stream s (seed seed0+s) serves `per` paths, scenario index s*per + j,
 * in groups of 64 consecutive paths (the last group holds per mod 64 when
 * that is not 0).  For every group the chains run step by step; at step k
 * the group's n paths take the stream's next n outputs in path order (so
 * a stream consumes exactly per*nk outputs per call).  Chain:
 * y_{-1} = y0, y_k = inverse CDF of row y_{k-1} of table m_k[k]
 * (smc.cu:195-255 restarts at y0 per scenario).  Output time-major
 * y_sk[k*nscen + scen] (smc.cu:222, rangen.cpp:122), nscen = nstreams*per.
 * layout_rows = 0: reference layout ctpk[m*d*d + yprev + d*y];
 * layout_rows = 1: row-contiguous ctpk[m*d*d + yprev*d + y].
 * d <= 64 uses the linear scan (the definition); larger d bisect, which is
 * identical on non-decreasing rows. */
void or_mc_chain(int nstreams, uint32_t seed0, int per, int nk, int d, int y0,
                 const float *ctpk, const int *m_k, int layout_rows, int16_t *y_sk)
{
    size_t nscen = (size_t)nstreams * per;
    size_t d2 = (size_t)d * d;
#pragma omp parallel for num_threads(nthreads()) schedule(dynamic, 16)
    for (int s = 0; s < nstreams; s++) {
        uint32_t st[625];
        or_mt19937_init(seed0 + (uint32_t)s, st);
        int ycur[64];
        for (int j0 = 0; j0 < per; j0 += 64) {
            const int n = per - j0 < 64 ? per - j0 : 64;
            for (int l = 0; l < n; l++) ycur[l] = y0;
            size_t sbase = (size_t)s * per + j0;
            for (int k = 0; k < nk; k++) {
                const float *tab = ctpk + (size_t)m_k[k] * d2;
                for (int l = 0; l < n; l++) {
                    float u = or_u01_from_bits(or_mt19937_next(st));
                    int yp = ycur[l], y;
                    if (layout_rows) {
                        const float *row = tab + (size_t)yp * d;
                        y = (d <= 64) ? or_sample_linear(row, d, u)
                                      : sample_bisect(row, 1, d, u);
                    } else {
                        y = (d <= 64) ? or_sample_colmajor(tab, d, yp, u)
                                      : sample_bisect(tab + yp, (size_t)d, d, u);
                    }
                    ycur[l] = y;
                    y_sk[(size_t)k * nscen + sbase + l] = (int16_t)y;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* payoff / pricing (A15)                                               */
/* ------------------------------------------------------------------ */

double or_ncdf(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

/* Black-Scholes call (BSPricer.cs:7-13), with an erfc-based N(x) rather than
 * the reference's Hart-style polynomial (BSPricer.cs:15-65). */
double or_bs_call(double S, double K, double r, double sigma, double T)
{
    double sq = sigma * sqrt(T);
    double d1 = (log(S / K) + (r + 0.5 * sigma * sigma) * T) / sq;
    double d2 = d1 - sq;
    return S * or_ncdf(d1) - K * exp(-r * T) * or_ncdf(d2);
}

/* ------------------------------------------------------------------ */
/* surrogate MLP (N3)                                                   */
/* ------------------------------------------------------------------ */

/* exact GELU x*Phi(x) */
double or_gelu(double x) { return 0.5 * x * erfc(-x / sqrt(2.0)); }
double or_gelu_grad(double x)
{
    return 0.5 * erfc(-x / sqrt(2.0)) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
}

/* Layers l = 0..nlayers-1: z_l = W_l a_{l-1} + b_l (a_{-1} = x), a_l =
 * gelu(z_l) except the last layer, which is linear.  W_l is
 * dims[l+1] x dims[l] row-major (W[o*in + i]); x is batch x dims[0];
 * acts receives every z_l (batch x dims[l+1], concatenated). */
void or_mlp_forward(int nlayers, const int *dims, const double *const *W,
                    const double *const *b, int batch, const double *x,
                    double *out, double *acts)
{
    int maxd = 0;
    for (int l = 0; l <= nlayers; l++) if (dims[l] > maxd) maxd = dims[l];
    double *a = (double *)malloc(sizeof(double) * maxd);
    double *z = (double *)malloc(sizeof(double) * maxd);
    for (int n = 0; n < batch; n++) {
        for (int i = 0; i < dims[0]; i++) a[i] = x[(size_t)n * dims[0] + i];
        size_t off = 0;
        for (int l = 0; l < nlayers; l++) {
            int in = dims[l], o = dims[l + 1];
            for (int j = 0; j < o; j++) {
                double s = b[l][j];
                for (int i = 0; i < in; i++) s += W[l][(size_t)j * in + i] * a[i];
                z[j] = s;
            }
            if (acts) memcpy(acts + off + (size_t)n * o, z, sizeof(double) * o);
            off += (size_t)batch * o;
            for (int j = 0; j < o; j++) a[j] = (l < nlayers - 1) ? or_gelu(z[j]) : z[j];
        }
        memcpy(out + (size_t)n * dims[nlayers], a, sizeof(double) * dims[nlayers]);
    }
    free(a);
    free(z);
}

/* Gradients of sum(out .* dout) w.r.t. W and b (summed over the batch). */
void or_mlp_backward(int nlayers, const int *dims, const double *const *W,
                     const double *const *b, int batch, const double *x,
                     const double *acts, const double *dout, double *const *dW,
                     double *const *db)
{
    (void)b;
    int maxd = 0;
    for (int l = 0; l <= nlayers; l++) if (dims[l] > maxd) maxd = dims[l];
    size_t *off = (size_t *)malloc(sizeof(size_t) * nlayers);
    size_t o = 0;
    for (int l = 0; l < nlayers; l++) { off[l] = o; o += (size_t)batch * dims[l + 1]; }
    for (int l = 0; l < nlayers; l++) {
        memset(dW[l], 0, sizeof(double) * (size_t)dims[l] * dims[l + 1]);
        memset(db[l], 0, sizeof(double) * dims[l + 1]);
    }
    double *g = (double *)malloc(sizeof(double) * maxd);
    double *gp = (double *)malloc(sizeof(double) * maxd);
    double *ain = (double *)malloc(sizeof(double) * maxd);
    for (int n = 0; n < batch; n++) {
        int L = nlayers - 1;
        for (int j = 0; j < dims[nlayers]; j++) g[j] = dout[(size_t)n * dims[nlayers] + j];
        for (int l = L; l >= 0; l--) {
            int in = dims[l], out = dims[l + 1];
            /* g = dLoss/dz_l */
            if (l < L) {
                const double *z = acts + off[l] + (size_t)n * out;
                for (int j = 0; j < out; j++) g[j] *= or_gelu_grad(z[j]);
            }
            if (l == 0) {
                for (int i = 0; i < in; i++) ain[i] = x[(size_t)n * in + i];
            } else {
                const double *zp = acts + off[l - 1] + (size_t)n * in;
                for (int i = 0; i < in; i++) ain[i] = or_gelu(zp[i]);
            }
            for (int j = 0; j < out; j++) {
                db[l][j] += g[j];
                for (int i = 0; i < in; i++) dW[l][(size_t)j * in + i] += g[j] * ain[i];
            }
            if (l > 0) {
                for (int i = 0; i < in; i++) {
                    double s = 0.0;
                    for (int j = 0; j < out; j++) s += W[l][(size_t)j * in + i] * g[j];
                    gp[i] = s;
                }
                memcpy(g, gp, sizeof(double) * in);
            }
        }
    }
    free(off);
    free(g);
    free(gp);
    free(ain);
}

/* Adam (Kingma & Ba 2015, Algorithm 1) with bias correction, step t >= 1. */
void or_adam_step(int n, double *p, const double *g, double *m, double *v,
                  double lr, double beta1, double beta2, double eps, int t)
{
    double c1 = 1.0 - pow(beta1, t), c2 = 1.0 - pow(beta2, t);
    for (int i = 0; i < n; i++) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}

/* ------------------------------------------------------------------ */
/* Stochastic-volatility lattice generator (CSVModel.host_d_gen,        */
/* CSVModel.cs:187-379, with the 4x4 stencil of S2DGrid.pregen,         */
/* S2DGrid.cs:123-245).  States y = x + nx*r; per state the x-variance  */
/* svol^2, regime drift voldrift and regime variance volvol^2 set the   */
/* four neighbour rates (the 4x4 system is block diagonal: a 2x2 in x   */
/* with zero drift, a 2x2 in r; at r = 0 / nr-1 the one-sided rate      */
/* voldrift/delta, at x = 0 / nx-1 no x moves).  A positive rate up to  */
/* regime r+1 is mirrored by jumps into regime r+1 (x1 < x0-1 with      */
/* weight exp((x1-x0)/(jm x0)), x1 > x0 with exp(-(x1-x0)/(jp x0)),     */
/* normalised to the same total).  Negative rates are clipped, the row  */
/* drift sum(L (x1 - x0)) is set to sdrift by one neighbour, the        */
/* diagonal closes the row.  Two fixes against the reference: the jump  */
/* weights use the scaled sizes jm x0, jp x0 both in the normalisation  */
/* and in the assignment (CSVModel.cs:296-297 vs 315,321; sgsvg.cu:111- */
/* 112 uses the minus size for both sides), and the drift correction    */
/* added to the down neighbour is (drift - sdrift)/(x0 - x1) > 0 (the   */
/* reference divides by x1 - x0, CSVModel.cs:356-357, sgsvg.cu:186,     */
/* which doubles the excess instead of removing it).  Entries are       */
/* summed in state order.  gen: d x d column-major, gen[y0 + d*y1].     */
/* ------------------------------------------------------------------ */
void or_sv_generator(int nx, int nr, const double *xval, const double *rval,
                     const double *sdrift, const double *svol, const double *volvol,
                     const double *voldrift, const double *jm, const double *jp,
                     double *gen)
{
    int d = nx * nr;
    memset(gen, 0, sizeof(double) * (size_t)d * d);
    for (int r0 = 0; r0 < nr; r0++) {
        for (int x0 = 0; x0 < nx; x0++) {
            int y0 = x0 + nx * r0;
            double *row = gen + y0; /* row[d*y1] */
            double X0 = xval[y0];
            /* x block: [dm dp; dm^2 dp^2] [am ap]' = [0 s2]' (interior x) */
            if (x0 > 0 && x0 < nx - 1) {
                double dm = xval[y0 - 1] - X0, dp = xval[y0 + 1] - X0;
                double s2 = svol[y0] * svol[y0];
                double det = dm * dp * dp - dp * dm * dm;
                row[(size_t)d * (y0 - 1)] = (-dp * s2) / det;
                row[(size_t)d * (y0 + 1)] = (dm * s2) / det;
            }
            /* r block */
            if (nr > 1) {
                double bm = 0.0, bp = 0.0;
                if (r0 > 0 && r0 < nr - 1) {
                    double em = rval[y0 - nx] - rval[y0], ep = rval[y0 + nx] - rval[y0];
                    double v2 = volvol[y0] * volvol[y0], mu = voldrift[y0];
                    double det = em * ep * ep - ep * em * em;
                    bm = (ep * ep * mu - ep * v2) / det;
                    bp = (-em * em * mu + em * v2) / det;
                } else if (r0 == 0) {
                    bp = voldrift[y0] / (rval[y0 + nx] - rval[y0]);
                } else {
                    bm = voldrift[y0] / (rval[y0 - nx] - rval[y0]);
                }
                if (r0 > 0) row[(size_t)d * (y0 - nx)] = bm;
                if (r0 < nr - 1) row[(size_t)d * (y0 + nx)] = bp;
                /* jumps into regime r0+1 mirroring the rate up */
                if (r0 < nr - 1 && bp > 0.0) {
                    double sm = jm[y0] * X0, sp = jp[y0] * X0, sum0 = 0.0;
                    for (int x1 = 0; x1 < x0 - 1; x1++)
                        if (sm > 0.0) sum0 += exp((xval[x1 + nx * (r0 + 1)] - X0) / sm);
                    for (int x1 = x0 + 1; x1 < nx; x1++)
                        if (sp > 0.0) sum0 += exp(-(xval[x1 + nx * (r0 + 1)] - X0) / sp);
                    if (sum0 > 0.0) {
                        double ratio = bp / sum0;
                        for (int x1 = 0; x1 < x0 - 1; x1++)
                            if (sm > 0.0) {
                                int y1 = x1 + nx * (r0 + 1);
                                row[(size_t)d * y1] = ratio * exp((xval[y1] - X0) / sm);
                            }
                        for (int x1 = x0 + 1; x1 < nx; x1++)
                            if (sp > 0.0) {
                                int y1 = x1 + nx * (r0 + 1);
                                row[(size_t)d * y1] = ratio * exp(-(xval[y1] - X0) / sp);
                            }
                    }
                }
            }
            /* clip, drift and mass in state order */
            double drift = 0.0, sum0 = 0.0;
            for (int y1 = 0; y1 < d; y1++) {
                if (y1 == y0) continue;
                double v = row[(size_t)d * y1];
                if (v < 0.0) {
                    row[(size_t)d * y1] = 0.0;
                } else {
                    drift += v * (xval[y1] - X0);
                    sum0 += v;
                }
            }
            double drift0 = sdrift[y0];
            if (drift > drift0) {
                if (x0 > 0) {
                    double ratio = (drift - drift0) / (X0 - xval[y0 - 1]);
                    row[(size_t)d * (y0 - 1)] += ratio;
                    sum0 += ratio;
                }
            } else if (x0 < nx - 1) {
                double ratio = (drift0 - drift) / (xval[y0 + 1] - X0);
                row[(size_t)d * (y0 + 1)] += ratio;
                sum0 += ratio;
            }
            row[(size_t)d * y0] = -sum0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Twisted GFSR lanes of the reference's GPU generator (smc.cu:118-133, */
/* 664-728; MersenneTwister.h:8-29): per lane a state of n words, the   */
/* recurrence x_i = x_{i+m} ^ ((x_i & umask | x_{i+1} & lmask) >> 1) ^  */
/* (odd ? matrix_a : 0) and the tempering with shifts (s0, 7, 15, s1)    */
/* and masks (mask_b, mask_c), stepped through a circular buffer whose  */
/* index restarts at 0 on every call (the reference keeps no index).   */
/* The DCMT-607 lanes use n = 19, m = 9, umask = 0xfffffffe, lmask = 1, */
/* s0 = 12, s1 = 18; with n = 624, m = 397, umask 0x80000000, s0 = 11   */
/* and the MT19937 constants the same code is MT19937 (pinned by its    */
/* known answer).  state: lane-major, st[lane + nlanes * i].            */
/* ------------------------------------------------------------------ */
void or_tgfsr_run(int nlanes, int n, int m, uint32_t umask, uint32_t lmask, int s0, int s1,
                  const uint32_t *matrix_a, const uint32_t *mask_b, const uint32_t *mask_c,
                  uint32_t *st, int count, uint32_t *out)
{
    uint32_t *s = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int l = 0; l < nlanes; l++) {
        for (int i = 0; i < n; i++) s[i] = st[l + (size_t)nlanes * i];
        int i = 0;
        for (int c = 0; c < count; c++) {
            int i1 = i + 1 >= n ? i + 1 - n : i + 1;
            int im = i + m >= n ? i + m - n : i + m;
            uint32_t y = (s[i] & umask) | (s[i1] & lmask);
            uint32_t x = s[im] ^ (y >> 1) ^ ((y & 1u) ? matrix_a[l] : 0u);
            s[i] = x;
            i = i1;
            x ^= x >> s0;
            x ^= (x << 7) & mask_b[l];
            x ^= (x << 15) & mask_c[l];
            x ^= x >> s1;
            out[l + (size_t)nlanes * c] = x;
        }
        for (int k = 0; k < n; k++) st[l + (size_t)nlanes * k] = s[k];
    }
    free(s);
}

/* lane states from seeds (smc.cu:118-133): st[l] = seed[l], then the
 * init_genrand recurrence over n words, lane-major */
void or_tgfsr_seed(int nlanes, int n, const uint32_t *seed, uint32_t *st)
{
    for (int l = 0; l < nlanes; l++) {
        st[l] = seed[l];
        for (int i = 1; i < n; i++) {
            uint32_t p = st[l + (size_t)nlanes * (i - 1)];
            st[l + (size_t)nlanes * i] = 1812433253u * (p ^ (p >> 30)) + (uint32_t)i;
        }
    }
}

/* ------------------------------------------------------------------ */
/* The reference's CPU chain sampler (F4): hash-bracket tables           */
/* (CModel Host Double.cs:403-458) and the sampler (rangen.cpp:96-153).  */
/* ctpk[m][y1 + d*y2]: running FP64 row sums divided by the row total,  */
/* capped at 1, column 0 left as the raw first entry (the reference     */
/* starts the normalisation at y2 = 1); hash[y1 + d*s + d*d*m] = lb |    */
/* ub << 16 brackets the states whose cumulative value crosses          */
/* [s/d, (s+1)/d).  A draw u = x / 2^32 takes bracket s = (int)(u d),   */
/* state 0 when u < ctpk[y][0] and lb == 0, else bisects (lb, ub] to ub. */
/* Uniforms: MT19937 streams, one per block of nscen paths (seed0 + b;  */
/* the reference's MKL MT2203 streams are not available), nscen words   */
/* per step, step-major (viRngUniformBits per step, rangen.cpp:118).    */
/* ------------------------------------------------------------------ */
void or_mc_hash_build(int d, int nm, const double *tpk, double *ctpk, int *hash)
{
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int my = 0; my < nm * d; my++) {
        int m = my / d, y1 = my % d;
        const double *t = tpk + (size_t)d * d * m;
        double *c = ctpk + (size_t)d * d * m;
        double acc = t[y1];
        c[y1] = acc;
        for (int y2 = 1; y2 < d; y2++) {
            acc += t[y1 + (size_t)d * y2];
            c[y1 + (size_t)d * y2] = acc;
        }
        for (int y2 = 1; y2 < d; y2++) {
            double v = c[y1 + (size_t)d * y2] / acc;
            c[y1 + (size_t)d * y2] = v < 1.0 ? v : 1.0;
        }
        int lb = 0, ub = 0;
        for (int s = 0; s < d; s++) {
            for (; lb < d; lb++)
                if (c[y1 + (size_t)d * lb] > (double)((float)s / (float)d)) break;
            if (lb > 0) lb -= 1;
            for (; ub < d; ub++)
                if (c[y1 + (size_t)d * ub] > (double)((float)(s + 1) / (float)d)) break;
            if (ub == d) ub -= 1;
            if (c[y1 + (size_t)d * lb] >= 1.0) ub = lb;
            hash[y1 + (size_t)d * s + (size_t)d * d * m] = lb + (ub << 16);
        }
    }
}

void or_mc_hash_run(int d, const double *ctpk, const int *hash, int nblocks, int nscen, int nk, int y0,
                    const int *m_k, uint32_t seed0, int16_t *y_sk)
{
    const int niter = (int)(log((double)d) / log(2.0)) + 1;
#pragma omp parallel for num_threads(nthreads()) schedule(dynamic, 1)
    for (int b = 0; b < nblocks; b++) {
        uint32_t st[625];
        or_mt19937_init(seed0 + (uint32_t)b, st);
        int16_t *yb = y_sk + (size_t)b * nk * nscen;
        for (int k = 0; k < nk; k++) {
            const int m = m_k[k];
            const int *hs = hash + (size_t)d * d * m;
            const double *cm = ctpk + (size_t)d * d * m;
            for (int sc = 0; sc < nscen; sc++) {
                const double u = (double)or_mt19937_next(st) / 4294967296.0;
                const int yp = k > 0 ? yb[(size_t)(k - 1) * nscen + sc] : y0;
                const int s = (int)(u * d);
                const int h = hs[yp + (size_t)d * s];
                int ub = h >> 16, lb = h - (ub << 16);
                if (lb == 0 && u < cm[yp]) ub = 0;
                if (ub - lb > 1) {
                    const double *cp = cm + yp;
                    for (int it = 0; it < niter; it++) {
                        const int mid = (ub + lb) / 2;
                        if (u < cp[(size_t)d * mid]) ub = mid;
                        else lb = mid;
                        if (ub - lb <= 1) break;
                    }
                }
                yb[(size_t)k * nscen + sc] = (int16_t)ub;
            }
        }
    }
}
