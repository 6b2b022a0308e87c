/*
 * oracle.h -- CPU restatement of OPLib's operator-method pricing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (libophip.so,
 * libopmodel.so, oplib_amd/) links, loads or calls this code.  It is used by
 * tests/ (as the checker), by __graft_entry__.smoke() (as the checker) and by
 * bench.py's cpu_baseline leg (as the timed CPU path).
 *
 * Every function names the reference file:line whose algorithm it restates
 * (paths relative to the OPLib repository root).  Matrices are column-major
 * as in the reference: element (y0,y1) of a d x d block is at y0 + d*y1 and
 * row y0 holds the rates / probabilities *from* state y0.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_xgrid_uniform     closed form (h = xval0/x0 exactly), tests/test_oracle.py
 *   or_stencil_invm      closed-form 2x2 inverse, A*invm = I, tests/test_oracle.py
 *   or_lv_generator      closed-form uniform-grid rates, zero row sums
 *   or_rs_generator      zero row sums; regime marginal = expm(Q t) (closed form)
 *   or_sv_generator      zero row sums, nonnegative rates, row drift = sdrift,
 *                        second moment = svol^2 + the drift fix's share,
 *                        jump profile and mass (tests/test_oracle.py)
 *   or_first_step*       exact formula, row sums 1
 *   or_sekd_f32, or_sglvg_f32  every FP32 step one correct rounding of its
 *                        exact value (rational arithmetic, tests/test_oracle.py)
 *   or_dgemm / squaring  numpy matmul; scipy expm fixtures (tests/golden)
 *   or_sck_f32           closed form C[.,d-1]=1, monotone rows, float64 cumsum;
 *                        a negative entry lowers the running sum (hand-worked row)
 *   or_sck_rows_f32      bit-equal to the transpose of or_sck_f32 on
 *                        nonnegative tables
 *   or_sck_rows2_f32     bit-equal to or_sck_rows_f32 where every sum is exact
 *                        (dyadic entries); within float64 cumsum rounding
 *                        otherwise; C[.,d-1] = 1, monotone rows
 *   or_sgemv_gather      numpy A @ Z[:, col0] with the per-16-group sentinel rule
 *   or_lv_params_cev     closed forms: sigma(100) = sigma0, the factor at
 *                        100 e and 100/e, the 1e-4 floor
 *   or_sample_ref_bisect equals the linear-scan inverse CDF for every d <= 1023
 *   or_mt19937_*         C++ std::mt19937 known answers (5489 -> 3499211612,
 *                        10000th output 4123659995)
 *   or_sample_linear     definition (min j: u < C[j]) checked by brute force
 *   or_bs_call           bstest's published value (S=K=100,r=.05,s=.25,T=1:
 *                        12.336) and put-call parity
 *   or_mlp_*             finite differences and the closed-form Adam step
 *   or_tgfsr_*           with MT19937's parameters the MT19937 known answers
 */
#ifndef OPLIB_ORACLE_H
#define OPLIB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lattice (A1) ---- */
void or_xgrid_uniform(int nx, double xlo, double xhi, double xval0,
                      double *xval, int *x0);
void or_stencil_invm(int nx, const double *xval, double *invm);

/* ---- generators (A2, A3) ---- */
void or_lv_generator(int d, const double *invm, const double *drift,
                     const double *vol, double *gen);
void or_rs_generator(int nx, int nr, const double *invm, const double *xval,
                     const double *sigma_r, double rate, double q_off,
                     double *gen);
void or_lv_params_c1(int nx, const double *xval, double rate, double *drift,
                     double *vol);
/* stochastic-volatility generator (CSVModel.cs:187-379), two fixes (jump
 * sizes, drift-correction sign); per-state arrays of d = nx*nr */
void or_sv_generator(int nx, int nr, const double *xval, const double *rval,
                     const double *sdrift, const double *svol, const double *volvol,
                     const double *voldrift, const double *jm, const double *jp,
                     double *gen);
void or_lv_params_cev(int nx, const double *xval, double rate, double sigma0,
                      double beta, double skew, double curv, double *drift,
                      double *vol);

/* ---- planner (A4) ---- */
int or_plan_niter(double DeltaT, double maxdiag, double dt0, double *dt_out);
double or_max_abs_diag(int d, const double *gen);

/* ---- first step (A5) ---- */
void or_first_step(int d, const double *gen, double dt, double *U);
void or_sek_f32(int d, const float *gen, float dt, float *U);
void or_sekd_f32(int d, const float *gen, float dt, const float *r, const float *lambda, float *U);
void or_sglvg_f32(int ni, int d, const float *invm, const float *pars, float *gen);

/* ---- GEMM / squaring / propagation (A6-A9) ---- */
void or_dgemm(int m, int n, int k, const double *A, int lda, const double *B,
              int ldb, double *C, int ldc);
void or_sgemm_f64acc(int m, int n, int k, float alpha, const float *A, int lda,
                     const float *B, int ldb, float beta, float *C, int ldc);
void or_square_chain(int d, int nsq, double *U, double *work);
void or_row_renorm(int d, double *U);
int or_set_threads(int nth);

/* ---- backward induction / gather GEMV (A10) ---- */
void or_sgemv_gather(int d, const float *A, float *Z, int ncol,
                     const int *col0, const int *col1, int sentinel);
void or_backward_induction(int d, const double *U, int nsteps, double df,
                           int ncol, double *V);

/* ---- cumulative tables (A11) ---- */
void or_sck_f32(int d, int nm, const float *tpk, float *ctpk);
void or_sck_rows_f32(int d, int nm, const float *tpk, float *crow);
void or_sck_rows2_f32(int d, int nm, int chunk, int plus_identity, const float *tpk, float *crow);

/* ---- MT19937 (A12, A14) ---- */
void or_mt19937_init(uint32_t seed, uint32_t *state /* 625 words */);
uint32_t or_mt19937_next(uint32_t *state);
float or_u01_from_bits(uint32_t x);
void or_mt_uniforms(uint32_t seed, int n, float *out);

/* ---- chain sampler (A13) ---- */
int or_sample_linear(const float *crow, int d, float u);
int or_sample_colmajor(const float *ctpk_m, int d, int yprev, float u);
int or_sample_ref_bisect(const float *ctpk_m, int d, int yprev, float u);
void or_mc_chain(int nstreams, uint32_t seed0, int per, int nk, int d, int y0,
                 const float *ctpk, const int *m_k, int layout_rows, int16_t *y_sk);

/* ---- payoff / pricing (A15) ---- */
double or_bs_call(double S, double K, double r, double sigma, double T);
double or_ncdf(double x);

/* ---- surrogate MLP (N3) ---- */
void or_mlp_forward(int nlayers, const int *dims, const double *const *W,
                    const double *const *b, int batch, const double *x,
                    double *out, double *acts);
void or_mlp_backward(int nlayers, const int *dims, const double *const *W,
                     const double *const *b, int batch, const double *x,
                     const double *acts, const double *dout, double *const *dW,
                     double *const *db);
void or_adam_step(int n, double *p, const double *g, double *m, double *v,
                  double lr, double beta1, double beta2, double eps, int t);
double or_gelu(double x);
double or_gelu_grad(double x);

/* ---- twisted GFSR lanes of the reference GPU generator (F4) ---- */
void or_tgfsr_run(int nlanes, int n, int m, uint32_t umask, uint32_t lmask, int s0, int s1,
                  const uint32_t *matrix_a, const uint32_t *mask_b, const uint32_t *mask_c,
                  uint32_t *st, int count, uint32_t *out);
void or_tgfsr_seed(int nlanes, int n, const uint32_t *seed, uint32_t *st);

/* ---- the reference's CPU hash-bracket chain sampler (F4) ---- */
void or_mc_hash_build(int d, int nm, const double *tpk, double *ctpk, int *hash);
void or_mc_hash_run(int d, const double *ctpk, const int *hash, int nblocks, int nscen, int nk, int y0,
                    const int *m_k, uint32_t seed0, int16_t *y_sk);

#ifdef __cplusplus
}
#endif
#endif
