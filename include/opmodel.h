/*
 * opmodel.h -- host-side model layer over the ophip C-ABI (libopmodel.so).
 *
 * A C++ restatement of the part of OPLib's C# opmodel layer that drives the
 * pricing path (CModel / CLVModel, oplib/opmodel/models/): lattice
 * construction (SGrid.cs:492-593, S1DGrid.cs:78-147), the step planner
 * (CModel Common.cs:289-307), generator assembly and first step
 * (CLVModel.cs:86-167, CModel Device.cs:186-196), repeated squaring
 * (device_mc_sfex, CModel Device.cs:216-228), the propagation chain
 * (device_mc_smk, CModel Device.cs:269-279) and pricing.  It calls libophip
 * only through the functions declared in ophip.h, as the C# layer calls
 * opcuda through P/Invoke.
 *
 * Lattice: nx price nodes uniform on [xlo,xhi], rescaled so that the node
 * nearest to spot lands on spot (SGrid.cs:540-557), times nr regimes,
 * state y = x + nx*r (S2DGrid.cs:89).
 */
#ifndef OPMODEL_H
#define OPMODEL_H

#ifdef __cplusplus
extern "C" {
#endif

/* local-vol families */
#define OPM_VOL_REGIME 0   /* sigma(x, r) = sigma_r[r]                           */
#define OPM_VOL_C1     1   /* clip(0.25*(x/100)^-0.5, 0.05, 1)                   */
#define OPM_VOL_CEV    2   /* sigma0*(x/100)^(beta-1)*(1+skew*l+curv*l^2), l=ln(x/100) */

/* pricing methods */
#define OPM_PRICE_BACKWARD 0  /* V <- df*U*V for nsteps steps, prices V_0[y0 + nx*r] */
#define OPM_PRICE_PROPAGATE 1 /* U_k = U_{k-1}*U, prices at every k from row y0 of U_k */

/* squaring arithmetic */
#define OPM_PREC_FP32   0     /* plain FP32 on U (the reference's device method) */
#define OPM_PREC_F16    1     /* U = I + W, FP16 one product, row-sum renormalised */
#define OPM_PREC_F16X3  2     /* U = I + W, FP16 hi/lo three products, renormalised */
#define OPM_PREC_BF16X3 3     /* U = I + W, BF16 hi/lo three products, renormalised */
#define OPM_PREC_F8X3   4     /* U = I + W, FP8 e4m3 hi/lo three products with power-of-two
                                 row/column scales, renormalised (single lattices only) */

typedef struct {
    int nx, nr;
    double xlo, xhi, spot;
    double rate, q_off;
    int vol_kind;
    double sigma_r[16];
    double sigma0, beta, skew, curv;
    double DeltaT;   /* step of the squared kernel, years */
    int nsq;         /* squarings: dt0 = DeltaT * 2^-nsq */
    int price_kind;  /* OPM_PRICE_* */
    int nsteps;      /* backward steps or propagation products */
    int nstrikes;
    double strikes[64];
    int precision;   /* OPM_PREC_* */
} opm_config;

/* host lattice and parameters (FP64), for inspection and tests */
int opm_grid(const opm_config *cfg, double *xval /* nx */, int *x0);
int opm_params(const opm_config *cfg, double *drift /* nx*nr */, double *vol /* nx*nr */);
/* the largest exit rate max_y |L_yy| of the configuration's generator */
double opm_max_rate(const opm_config *cfg);
/* the reference planner (CModel Common.cs:289-307) in years:
 * dt = min(0.5/max|L_yy|, dt0_cap), niter = ceil(log2(DeltaT/dt)) */
int opm_plan_nsq(const opm_config *cfg, double dt0_cap);

/* the SV generator's per-state inputs from the model's term structure,
 * CSVModel.mkgen (CSVModel.cs:103-172): for each generator date i (t_years[i]
 * = (t_i - today) in days / 365.25) the eight curves c (tau: 8 x 4
 * row-major {A, B, c2, c3}) give xi_c = A + (B - A)(1 + c3 t/c2) exp(-t/c2)
 * (A when c2 = 0), plus rho[8 i + c] when rho is given: vol, low beta, high
 * beta, vol of vol, vol mean-reversion rate and level, jump sizes - and +.
 * Per state y = x + nx r (xval[nx] nodes, rval[nr] regime values, y0 the
 * spot state): beta = (high (sqrt x_y - sqrt x_min) + low (sqrt x_max -
 * sqrt x_y)) / (sqrt x_max - sqrt x_min), svol = rval_y vol (x_y /
 * x_y0)^beta, volvol, voldrift = rate (level - rval_y), jm, jp; outputs
 * ni x d, date-major, FP64 on the host.  The arrays feed ophip_dgen_sv.
 * Returns 0, or 1 on a bad argument or a non-finite curve value (where
 * the reference throws). */
int opm_sv_params(int nx, int nr, const double *xval, const double *rval, int y0, int ni, const double *t_years,
                  const double *tau, const double *rho, double *svol, double *volvol, double *voldrift,
                  double *jm, double *jp);

/* a solver owns device buffers for one configuration */
typedef struct opm_solver opm_solver;
opm_solver *opm_create(const opm_config *cfg);
void opm_destroy(opm_solver *s);
/* run the whole solve (assembly, first step, squarings, pricing) */
int opm_solve(opm_solver *s);
/* the same solve recorded once into a HIP graph and replayed */
int opm_capture(opm_solver *s);
int opm_launch(opm_solver *s);
/* wait and copy results: prices[nr][n_mat][nstrikes] (n_mat = 1 for
 * backward induction, nsteps for propagation) */
int opm_prices(opm_solver *s, double *prices);
/* copy U(DeltaT) (FP32, d*d column-major; I + W formed on the host in the
 * W form), states in the reference's order y = x + nx*r */
int opm_kernel(opm_solver *s, float *U);
/* the squared kernel: U, or W when *is_w = 1 (same state order) */
int opm_kernel_raw(opm_solver *s, float *K, int *is_w);
/* its device handle (valid after a solve); on the device the W form keeps
 * node-major states y = nr*x + r when nr > 1 */
unsigned opm_kernel_handle(opm_solver *s);
int opm_dims(opm_solver *s, int *d, int *n_mat);

/* ---- a group of scenarios of one lattice size, priced together ----
 * Solvers (opm_create) of one d, W-form precision (F16, F16X3, BF16X3),
 * backward pricing, squaring count and step count: their squarings run in
 * one persistent launch per round and, in F16X3, every backward step of
 * all of them in one launch.  Each solver keeps its buffers; opm_prices on
 * a member reads its prices after opm_group_solve / opm_group_launch.  A
 * solver belongs to at most one group at a time; destroy the group first. */
typedef struct opm_group opm_group;
opm_group *opm_group_create(opm_solver *const *solvers, int n);
void opm_group_destroy(opm_group *g);
int opm_group_solve(opm_group *g);
int opm_group_capture(opm_group *g);
int opm_group_launch(opm_group *g);

/* ---- calibration sweep: many one-regime lattices of different size ----
 * Every configuration shares precision, nsq, DeltaT, nsteps (maturities
 * k*DeltaT, k = 1..nsteps) and the strikes; nx, the vol parameters and
 * the rate differ.  Pricing is the forward step of the spot row,
 * p_k = p_{k-1} U(DeltaT).  In the W form all lattices square together in
 * one persistent launch per round (ophip_wsq_plan); the plain FP32 method
 * squares them one launch each.  tile: 0 = autotuned/default, 64 or 128. */
typedef struct opm_batch opm_batch;
opm_batch *opm_batch_create(const opm_config *cfgs, int n, int tile);
void opm_batch_destroy(opm_batch *b);
int opm_batch_solve(opm_batch *b);
int opm_batch_capture(opm_batch *b);
int opm_batch_launch(opm_batch *b);
/* prices[n][nsteps][nstrikes] */
int opm_batch_prices(opm_batch *b, double *prices);
/* new drift/vol/rate for every lattice, same grids, squaring counts,
 * maturities and strikes: plans and the captured graph stay valid */
int opm_batch_update(opm_batch *b, const opm_config *cfgs);

/* ---- the reference's step plan (CModel Common.cs:163-644), in days ----
 * today = day 0; step dates t_k[nk] and generator dates t_i[ni] strictly
 * increasing and > 0 (t_k[nk-1] <= t_i[ni-1]); generator i covers
 * (t_{i-1}, t_i]; dt0_days caps the first step; maxdiag[i] = max_y |L_yy| of
 * generator i (per year).  Kernels j: distinct (generator, interval length)
 * pairs of the merged dates, each I + dt_j L squared niter_j times; step
 * kernels m: the product of a step's interval kernels (chain), shared by
 * later steps of the same length inside one generator period; m_k[k] maps
 * step k to its step kernel.  Returns 0 on invalid dates. */
typedef struct opm_mcplan opm_mcplan;
opm_mcplan *opm_mcplan_create(int nk, const int *t_k, int ni, const int *t_i, int dt0_days,
                              const double *maxdiag);
void opm_mcplan_destroy(opm_mcplan *p);
/* nr = the most squarings of any kernel, ns = the longest chain */
int opm_mcplan_sizes(const opm_mcplan *p, int *nq, int *nj, int *nm, int *nr, int *ns);
/* any pointer may be 0: t_q/i_q/j_q[nq], i_j/niter_j/dt_j[nj] (dt_j in
 * years), m_k[nk], chain[nm*ns] (-1 padded), chain_len[nm] */
int opm_mcplan_arrays(const opm_mcplan *p, int *t_q, int *i_q, int *j_q, int *i_j, int *niter_j,
                      double *dt_j, int *m_k, int *chain, int *chain_len);
/* the plan on the device in FP32 through opcuda_sek / opcuda_ssqmm pointer
 * tables / opcuda_scopy1 (CModel Device.cs:186-295): gen = ni generators
 * (d*d column-major each, per year), xtpk = nj*d*d, xbuf = max(nj,nm)*d*d,
 * tpk = nm*d*d floats; leaves tpk_m = the step kernel of m (each chain
 * factor once, in date order) */
int opm_mcplan_exec(opm_mcplan *p, int d, unsigned gen, unsigned xtpk, unsigned xbuf, unsigned tpk);

#ifdef __cplusplus
}
#endif
#endif
