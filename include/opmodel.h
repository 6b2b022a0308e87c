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
#define OPM_PREC_FP32 0       /* plain FP32 (the reference's device method) */

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
/* copy U(DeltaT) (FP32, d*d column-major) */
int opm_kernel(opm_solver *s, float *U);
int opm_dims(opm_solver *s, int *d, int *n_mat);

#ifdef __cplusplus
}
#endif
#endif
