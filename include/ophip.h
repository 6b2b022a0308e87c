/*
 * ophip.h -- C-ABI of libophip, the MI355X (gfx950) operator-method pricing
 * library.  Drop-in for OPLib's opcuda (the .cu files of oplib/opcuda), reached from the
 * C# opmodel layer through [DllImport]; every opcuda_* entry point keeps the
 * reference's name, argument order and meaning.  Each declaration cites the
 * reference definition it replaces.
 *
 * Device buffers are 32-bit handles: byte offsets into one per-device arena
 * (0 = null).  The reference's 32-bit "device pointers" (CArray.cs:46,
 * util.cu:209-214) and the caller's arithmetic on them (e.g.
 * CModel Device.cs:193, ptr + j*d*d*4) carry over unchanged.  Device-side
 * pointer tables and descriptors (opcuda_ssqmm, opcuda_sgemm4,
 * opcuda_sgemv4) hold handles too.  Host pointers are 64-bit.
 *
 * Status: every call sets a thread-local status read by opcuda_get_status();
 * entry points that return int return that status (0 = success).  Codes are
 * hipError_t values or OPHIP_ERR_* below.
 *
 * Streams: all work goes to one non-blocking stream per device;
 * opcuda_thread_synchronize() waits for it; copies are synchronous.
 *
 * Matrices are column-major; element (y0,y1) of a d x d block is at
 * y0 + d*y1 and row y0 holds the rates/probabilities out of state y0.
 */
#ifndef OPHIP_H
#define OPHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define OPHIP_ERR_BAD_HANDLE  0x10001u
#define OPHIP_ERR_OUT_OF_ARENA 0x10002u
#define OPHIP_ERR_BAD_ARG     0x10003u
#define OPHIP_ERR_NO_DEVICE   0x10004u

/* ---- context, device, status (util.cu) ---- */
int      opcuda_cublas_init(void);                     /* util.cu:123 (creates arena + stream) */
int      opcuda_shutdown(void);                        /* util.cu:287 */
void     opcuda_thread_synchronize(void);              /* util.cu:131 */
void     opcuda_set_device(int dev);                   /* util.cu:139 */
int      opcuda_get_device(void);                      /* util.cu:145 */
int      opcuda_device_get_count(void);                /* util.cu:156 */
int      opcuda_device_count(void);                    /* util.cu:240 */
int      opcuda_multi_processor_count(int dev);        /* util.cu:248 */
int      opcuda_max_threads_per_block(int dev);        /* util.cu:255 */
unsigned opcuda_get_status(void);                      /* util.cu:204 */

/* ---- memory (util.cu:182-236) ---- */
unsigned      opcuda_mem_alloc(unsigned sz);                          /* util.cu:209 */
void          opcuda_mem_free_device(unsigned p);                     /* util.cu:217 */
unsigned long opcuda_mem_alloc_host(unsigned sz);                     /* util.cu:223 */
void          opcuda_mem_free_host(unsigned long h);                  /* util.cu:232 */
int opcuda_memcpy_h2d(unsigned dptr, unsigned long hptr, unsigned sz);/* util.cu:182 */
int opcuda_memcpy_d2h(unsigned dptr, unsigned long hptr, unsigned sz);/* util.cu:190 */
int opcuda_memcpy_d2d(unsigned dst, unsigned src, unsigned sz);       /* util.cu:198 */

/* ---- timers (util.cu:62-113) ---- */
void   opcuda_create_timer(void);
void   opcuda_reset_timer(void);
void   opcuda_stop_timer(void);
double opcuda_get_timer_value(void);

/* ---- copies and fills ---- */
void opcuda_scopy(int n, unsigned x, int incx, unsigned y, int incy); /* util.cu:278 */
int  opcuda_scopy1(unsigned dst, unsigned src, unsigned n);           /* scopy2.cu:137 (returns status; C# declares int) */
void opcuda_ssetall(unsigned x, int n, float c, int incx);            /* sset.cu:54 */
void opcuda_ssetone(unsigned x, int n, float c, int i);               /* sset.cu:69 */

/* ---- BLAS (sgemm.cu, sgemv.cu) ---- */
void opcuda_sgemm(int m, int n, int k, float alpha, unsigned A, int lda, unsigned B, int ldb,
                  float beta, unsigned C, int ldc);                    /* sgemm.cu:41 */
void opcuda_sgemv(int m, int n, float alpha, unsigned A, int lda, unsigned x, int incx,
                  float beta, unsigned y, int incy);                   /* sgemv.cu:44 */
void opcuda_sgemvT(int m, int n, float alpha, unsigned A, int lda, unsigned x, int incx,
                   float beta, unsigned y, int incy);                  /* sgemv.cu:55 */

/* ---- batched products (ssqmm.cu, sgemm4.cu) ---- */
void opcuda_ssqmm(int d, int ni, unsigned A_i, unsigned B_i, unsigned C_i); /* ssqmm.cu:174 */
void opcuda_sgemm4(int nblocks, unsigned argptr_bid);                 /* sgemm4.cu:148; 64x16 tiles */

/* ---- gather GEMV (sgemv3.cu, sgemv4.cu) ---- */
void opcuda_sgemv3(int d, int nz, int ncol, unsigned A, unsigned Z, unsigned col0,
                   unsigned col1);                                     /* sgemv3.cu:163 */
void opcuda_sgemv4(int nblocks, unsigned argptr_bid);                 /* sgemv4.cu:174 */

/* ---- first step, tables, generator (sek.cu, sck.cu, sglvg.cu) ---- */
void opcuda_sek(int d, unsigned gen, float dt, unsigned tpk);          /* sek.cu:106 */
void opcuda_sekd(int d, unsigned gen, float dt, unsigned tpk, unsigned r);          /* sek.cu:117 */
void opcuda_sekd_lambda(int d, unsigned gen, float dt, unsigned tpk, unsigned r,
                        unsigned lambda);                               /* sek.cu:129 */
void opcuda_sck(int d, int nm, unsigned tpk, unsigned ctpk);           /* sck.cu:88 */
void opcuda_sglvg(unsigned gen, int ni, int d, unsigned invm, unsigned pars); /* sglvg.cu:104 */
/* sgsvg.cu:215 (bound by CSVModel.cs:383-384): ni FP32 SV generators from
 * the caller's 4x4 stencil inverses (16 floats per state) and pars = xval
 * (d), then per generator SDrift, SVol, VolDrift, VolVol, Jumpsz_minus,
 * Jumpsz_plus (ni*d each); evaluated in FP64 as ophip_dgen_sv (same two
 * fixes), rounded once; writes every entry of each d x d block.
 * nx >= 3, nr >= 2. */
void opcuda_sgsvg(unsigned gen, int ni, int nx, int nr, unsigned invm, unsigned pars);

/* ---- Monte Carlo (smc.cu) ---- */
int  opcuda_mc_load_mt_gpu(char *stream, long sz);                     /* smc.cu:87 (MT19937: no-op) */
int  opcuda_mc_nrng(void);                                             /* smc.cu:98: number of streams */
int  opcuda_mc_status_sz(void);                                        /* smc.cu:103: 625 words/stream */
void opcuda_mc_setseed(unsigned long host_seedptr, unsigned mtptr);    /* smc.cu:137 */
int  opcuda_mc1f(unsigned mt, int y0, unsigned y_sk, int nscen, int nk, int d, unsigned ctpk,
                 unsigned m_k, unsigned long yhost);                   /* smc.cu:610 (yhost 0: no copy) */
int  opcuda_mt_benchmark(unsigned mt, unsigned unif, int nscen);       /* smc.cu:731 */

/* ==== MI355X additions (no reference counterpart unless cited) ==== */

/* sets the calling thread's status, read by opcuda_get_status (a layer
 * above the library reports its own failures through it); returns it */
unsigned ophip_set_status(unsigned status);

/* raw device address of a handle; arena size; library stream */
unsigned long ophip_device_address(unsigned h);
unsigned long ophip_arena_bytes(void);
unsigned long ophip_stream(void);

/* FP64 regime-switching local-vol lattice on y = x + nx*r (S2DGrid.cs:89):
 * price rates from the 2x2 stencil inverse of S1DGrid.cs:78-147 applied to
 * (drift, vol^2) as CLVModel.cs:99-114, regime rate q_off at equal x,
 * diagonal = -sum.  xval: nx doubles; drift, vol: nx*nr doubles (price
 * units).  nr = 1 is the 1-D LV generator of CLVModel.host_d_gen. */
int ophip_dgen_rs(int nx, int nr, unsigned xval, unsigned drift, unsigned vol, double q_off,
                  unsigned gen);
/* the same generator fused with the first step, evaluated in FP64:
 * out_kind 0: FP32 I+dt*L, 1: FP64 I+dt*L, 2: FP32 dt*L, 3: FP64 dt*L
 * (host_d_mc_ek, CModel Host Double.cs:89-121); out_kind + 4: the same in
 * node-major state order y = nr*x + r (one band; used by the W form) */
int ophip_sek_rs(int nx, int nr, unsigned xval, unsigned drift, unsigned vol, double q_off,
                 double dt, int out_kind, unsigned out);
/* FP64 stochastic-volatility generator (CSVModel.host_d_gen, CSVModel.cs:
 * 187-379; the reference's FP32 device version opcuda_sgsvg, sgsvg.cu:215):
 * per-state handles of d = nx*nr doubles (state y = x + nx*r): grid values
 * xval, regime values rval, the row drift sdrift (r x), x vol svol (price
 * units), volvol, voldrift, relative jump sizes jm, jp.  out_kind 0 FP32
 * I+dt*L, 1 FP64 I+dt*L, 2 FP32 dt*L, 3 FP64 dt*L, 4 FP64 L.  nr >= 2.
 * The jump weights use the scaled sizes in normalisation and assignment,
 * and the down-neighbour drift correction has the sign that removes the
 * excess (both differ from the reference, which mixes the sizes and
 * doubles the excess). */
int ophip_dgen_sv(int nx, int nr, unsigned xval, unsigned rval, unsigned sdrift, unsigned svol,
                  unsigned volvol, unsigned voldrift, unsigned jm, unsigned jp, double dt, int out_kind,
                  unsigned out);
/* ophip_sek_rs for n lattices in one launch per 32 (lattice k takes the
 * k-th entry of every array, one out_kind for all); the same values */
int ophip_sek_rs_batch(int n, const int *nx, const int *nr, const unsigned *xval,
                       const unsigned *drift, const unsigned *vol, const double *q_off,
                       const double *dt, int out_kind, const unsigned *out);
/* cumulative tables stored row-contiguous: crow[m*d*d + row*d + y]; the
 * same values as opcuda_sck (sck.cu:43-84) */
int ophip_sck_rows(int d, int nm, unsigned tpk, unsigned crow);
/* the same for U = I + W given W (plus_identity = 1 adds 1 to each
 * diagonal entry before the running sum) */
int ophip_sck_rows_ex(int d, int nm, unsigned tpk, unsigned crow, int plus_identity);

/* strided batched FP32 GEMM on the matrix cores */
int ophip_sgemm_strided_batched(int m, int n, int k, float alpha, unsigned A, int lda,
                                long long strideA, unsigned B, int ldb, long long strideB,
                                float beta, unsigned C, int ldc, long long strideC, int batch);
/* Y = alpha*A*X, A m x K column-major, X (K x ncp) and Y (m x ncp)
 * row-major, ncp in {16,32,48}, m % 64 == 0: one backward-induction
 * step over many payoff columns (the role of sgemv3/sgemv4) */
int ophip_sgemm_skinny(int m, int K, int ncp, float alpha, unsigned A, int lda, unsigned X,
                       int ldx, unsigned Y, int ldy);

/* Y = alpha*(A*X + X): one backward step with U = I + W held as W */
int ophip_sgemm_skinny_ex(int m, int K, int ncp, float alpha, unsigned A, int lda, unsigned X,
                          int ldx, unsigned Y, int ldy, int plus_x);
/* C = a*A + b*B elementwise over n floats */
int ophip_saxpby(int n, float a, unsigned A, float b, unsigned B, unsigned C);

/* Repeated squaring in the U = I + W form on FP16/BF16 matrix cores with
 * FP32 accumulation and per-step row-sum renormalisation (the reference
 * squares U in FP32, ssqmm.cu:119-170, driven by CModel Device.cs:216-228).
 * mode 0: FP16, one product H*H (H = fp16(W*2^15));
 * mode 1: FP16, three products H*H + 2^-11 (H*L + L*H);
 * mode 2: BF16, three products.
 * A state (ophip_wsq_bytes(d, mode) bytes, d % 128 == 0) holds two FP32 W
 * buffers, their 16-bit copies and partial row sums.  Write W0 = dt*L into
 * buffer 0 (ophip_sek_rs out_kind 2); ophip_wsq_exec leaves W(2^nsq dt) in
 * buffer nsq & 1.  A plan batches many states of different d into the same
 * launches; executing it only launches kernels (graph-capturable).  nsqs
 * (optional) gives each lattice its own squaring count, used when exec is
 * called with nsq < 0: round r squares the lattices with nsqs[i] > r (the
 * reference's per-round active lists, CModel Common.cs:508-545). */
unsigned long ophip_wsq_bytes(int d, int mode);
unsigned      ophip_wsq_w(unsigned state, int d, int mode, int buf);
/* any array of a state: which = 0 W of buffer buf (FP32, column-major),
 * 2 and 4 the backward layout's fragment-major hi and lo parts of W (buf
 * ignored; modes 1 and 2; the squarings split W as they stage it and store
 * no copies), 5 partial row sums, 6 cell flags of buffer buf (d/64 x d/64
 * bytes; 0 = every W value of that 64 x 64 cell is +0; otherwise 96 + E with
 * the cell's largest |W| below 2^E, or 255 where no bound is known), 7
 * backward-step tile lists (d/16 + d/64 +
 * d/32 lists of d/32 + 1 ints, per 16-row tile, 64-row block and 32-row
 * block: count, then the visited k tiles, kt + 65536 where the lo part is
 * read too, + 2^(17+q) in a block of nr row tiles where its tile q visits
 * kt and + 2^(17+nr+q) where that tile reads the lo part; buf ignored)
 * (0 for which = 1 or 3 and when the mode has no such array) */
unsigned      ophip_wsq_array(unsigned state, int d, int mode, int buf, int which);
/* Zero-cell skipping: the squaring visits only the K cells whose A and B
 * cells are flagged nonzero (the skipped products are exact zeros: same
 * bits).  1 = on (default), 0 = every K-step; returns the previous value.
 * With skipping on, ophip_wsq_set_kcut(k) also leaves out the K cells whose
 * operand cells' bounds (the flags above) multiply to at most 2^k: such a
 * product moves no entry of W by more than 64 2^k (default k = -36; -192 or
 * below keeps every cell that may be nonzero, the dense loop's bits);
 * returns the previous k and applies to rounds launched or captured
 * afterwards.  No reference counterpart (the reference's ssqmm.cu:119-170
 * is dense). */
int           ophip_wsq_set_skip(int on);
int           ophip_wsq_set_kcut(int log2_bound);
unsigned long ophip_wsq_plan(int n, const unsigned *states, const int *ds, const int *nsqs,
                             int mode, int tile, int grid);
int           ophip_wsq_exec(unsigned long plan, int nsq);
void          ophip_wsq_plan_free(unsigned long plan);
/* ophip_sek_rs_batch's FP64 assembly of W0 = dt*L (out_kind 2; node_major:
 * 6) into buffer 0 of each lattice's state, fused with the first stage of
 * ophip_wsq_exec: the flush rule, the partial row sums and buffer 0's cell
 * flags.  ophip_wsq_exec_ex(plan, nsq, 1) then leaves that stage out; the
 * results are the same bits as ophip_sek_rs_batch + ophip_wsq_exec.  Cells
 * that are structurally zero and already flagged zero are not written, so
 * buffer 0's flags must describe its contents (as plan creation and every
 * exec leave them).  d = nx*nr % 64 == 0; mode as the plan's.  The
 * reference assembles and exponentiates in separate passes (sek.cu:40-103,
 * CModel Device.cs:216-228). */
int           ophip_wsq_assemble_rs(int n, const int *nx, const int *nr, const unsigned *xval,
                                    const unsigned *drift, const unsigned *vol, const double *q_off,
                                    const double *dt, int node_major, int mode,
                                    const unsigned *states);
/* ophip_wsq_exec with flags: 1 = buffer 0 was written by
 * ophip_wsq_assemble_rs (no split) */
int           ophip_wsq_exec_ex(unsigned long plan, int nsq, int flags);
/* The W form with FP8 e4m3 operands: per round, row / column power-of-two
 * scales, hi/lo e4m3 copies (fragment-major), W' = 2W + W*W on
 * v_mfma_f32_16x16x32_fp8_fp8 (three products, FP32 accumulate) and the
 * FP64 row-sum diagonal reset.  d % 64 == 0.  Write W0 = dt*L into buffer
 * 0 (ophip_sek_rs out_kind 2); W(2^nsq dt) ends in buffer nsq & 1.
 * which: 0 Ah, 1 Al, 2 Bh, 3 Bl, 4 row scales, 5 column scales. */
unsigned long ophip_w8_bytes(int d);
unsigned      ophip_w8_w(unsigned state, int d, int buf);
unsigned      ophip_w8_array(unsigned state, int d, int which);
int           ophip_w8_exec(unsigned state, int d, int nsq);

/* Backward induction in the W form on the matrix cores (modes 1, 2):
 * Y = df (X + W X).  ophip_wsq_bwd_layout splits W (buffer `buf`, the one
 * the squarings ended in) into its hi/lo parts in the state's MFMA
 * fragment-major arrays (ophip_wsq_array 2 and 4) and builds the tile
 * lists; the layouts in W's own state order and in the regime order do not
 * read the 64 x 64 cells that buffer's flags mark zero, so the flags must
 * describe W (as plan creation and every exec leave them); X/Y are d x ncp
 * row-major FP32 (ncp 16, 32 or 48) with 16-bit hi/lo copies XT/YT
 * (2 * ncp * d values, fragment-major); ophip_wsq_bwd_prep writes XT from
 * X.  A step given X = 0 takes X from XT (hi + lo 2^-G: 22 significant bits
 * in FP16, 16 in BF16); given Y = 0 it writes YT only, so a chain of steps
 * moves the FP32 values only where they are wanted (the last step).  The
 * reference's step is the FP32 SGEMV of U (sgemv.cu:44-51, driven per step
 * by CModel Device.cs). */
int ophip_wsq_bwd_layout(unsigned state, int d, int mode, int buf);
/* the same with the states reordered: fragment state i is the copies'
 * state perm[i] (perm: handle of d ints, a permutation), e.g. a lattice
 * squared in node-major order stepped in the reference's order (the step's
 * FP32 accumulation depends on the order; its algebra does not) */
int ophip_wsq_bwd_layout_perm(unsigned state, int d, int mode, int buf, unsigned perm);
/* the same for the regime permutation perm[x + nx*r] = nr*x + r (node-major
 * squaring, reference-order steps) without a permutation array, through
 * LDS tiles: nr in {1, 2, 4, 8}, nx % 8 == 0, nx*nr = d */
int ophip_wsq_bwd_layout_regime(unsigned state, int d, int mode, int buf, int nx, int nr);
/* the same for n lattices of one d, nx and nr in three launches (per 32
 * lattices): layout, flags, lists.  No reference counterpart (the
 * reference lays out nothing; it steps one model at a time). */
int ophip_wsq_bwd_layout_regime_batch(int n, const unsigned *states, int d, int mode, int buf, int nx,
                                      int nr);
int ophip_wsq_bwd_prep(int d, int ncp, int mode, unsigned X, unsigned XT);
int ophip_wsq_bwd_step(unsigned state, int d, int mode, int buf, int ncp, float df, unsigned X,
                       unsigned XT, unsigned Y, unsigned YT);
/* The backward step of nb lattices of one d in one launch (per 32): Y[i] =
 * df[i] (X[i] + W_i X[i]), W_i from states[i] laid out as above.  The
 * band-sparse step kernels read the hi part of a 16 x 32 fragment only
 * where some |W| >= 2^visit_exp and its lo part only where some |W| >=
 * 2^lo_exp (FP16) or 2^(lo_exp - 3) (BF16) (lists built by
 * ophip_wsq_bwd_layout*, with the cutoffs set then): what they leave out of
 * an entry of W is the entry where it is under 2^visit_exp, else its lo part
 * where it is under 2^lo_exp, at most max(2^(lo_exp - 12), 2^-40); a row of
 * Y moves by at most the left-out parts times |X|.  The dense kernel reads
 * every fragment.  ophip_wsq_set_bwd_cutoffs sets the two exponents (-40 <=
 * visit_exp <= lo_exp <= -1; default -24 and -12, which move the 2048-state
 * daily workload's prices by under 1e-5 relative; -40 and -28 leave out
 * nothing above 2^-40) for lists built afterwards.
 * ophip_wsq_set_bwd_kernel(k) selects the kernel: 0 the
 * dense one (every fragment), 1 automatic (default), 2 per 16-row tile, 3
 * per 64-row block, 4 per 32-row block (3 and 4: ncp 48); it returns the
 * previous setting and applies to kernels launched afterwards.  No
 * reference counterpart: the reference steps one model at a time
 * (sgemv.cu:44-51). */
int ophip_wsq_bwd_step_batch(int nb, const unsigned *states, int d, int mode, int buf, int ncp,
                             const float *df, const unsigned *X, const unsigned *XT,
                             const unsigned *Y, const unsigned *YT);
int ophip_wsq_set_bwd_kernel(int k);
int ophip_wsq_set_bwd_cutoffs(int visit_exp, int lo_exp);
int ophip_wsq_get_bwd_cutoffs(int *visit_exp, int *lo_exp);
/* the kernel (0, 2, 3, 4 as above) the current setting gives a step of nb
 * lattices of size d with ncp value columns */
int ophip_wsq_bwd_kernel_for(int nb, int d, int ncp);
/* the same plan with a tile per lattice (64 or 128 each): one persistent
 * launch per tile class and round */
unsigned long ophip_wsq_plan_tiles(int n, const unsigned *states, const int *ds, const int *nsqs,
                                   int mode, const int *tiles, int grid);
/* autotuner: per size bucket the faster of 64- and 128-tiles (each bucket
 * timed alone), then the fastest persistent grid (1, 2, 4 workgroups per
 * CU) for the combined plan; cached per (mode, d, count) and per size
 * signature.  Runs kernels on the states (contents are overwritten) and
 * synchronizes: call it before writing W0, never under capture. */
int           ophip_wsq_autotune(int n, const unsigned *states, const int *ds, const int *nsqs,
                                 int mode, int *tiles, int *grid);

/* Forward pricing of many lattices in one launch per step (calibration
 * sweep): p_k = p_{k-1} U, or p_{k-1}(I + W) when plus_identity[i]; prices
 * df_k * sum_j p_k[j] max(x_j - K_c, 0) for k = 1..nmat (maturity k*dT).
 * P[i]: (nmat+1) x d row-major, row 0 = initial distribution; out[i]:
 * nmat x nstrikes.  exec only launches kernels (graph-capturable). */
unsigned long ophip_fwd_plan(int n, const int *d, const unsigned *K, const int *plus_identity,
                             const unsigned *P, const unsigned *x, const unsigned *out,
                             const float *rate, float dT, int nmat, const float *strikes,
                             int nstrikes);
int  ophip_fwd_exec(unsigned long plan);
/* per lattice the 64 x 64 cell flags of its kernel (handles of (d/64)^2
 * bytes as ophip_wsq_array which = 6 gives; 0 = none): the forward steps
 * then skip zero cells (exact: their entries are 0).  d % 64 == 0. */
int  ophip_fwd_set_cells(unsigned long plan, const unsigned *flags);
void ophip_fwd_plan_free(unsigned long plan);
/* prices at up to 16 spot rows of one kernel (U, or W with plus_identity):
 * out[r*nstrikes + c] = df (plus_identity pay(y0_r, c) + sum_j U[y0_r][j]
 * pay(j, c)), pay(j, c) = payT[c + nstrikes*j]: the payoff GEMV of each
 * row (OPPricer.cs:143-146) in one launch */
int  ophip_spot_prices(int d, unsigned U, int nrow, const int *rows, int plus_identity,
                       unsigned payT, int nstrikes, float df, unsigned out);
/* new rates for every lattice of a plan (waits for the library stream) */
int  ophip_fwd_set_rates(unsigned long plan, const float *rate);

/* Learned surrogate (no reference counterpart): an ensemble of MLPs
 * 5 -> 256 -> 256 -> 256 -> 210, exact-erf GELU on the hidden layers, BF16
 * operands on the matrix cores with FP32 accumulation, FP32 master
 * weights and Adam moments.  Batches B % 64 == 0, B <= capacity.
 * Parameter arrays (ophip_mlp_param_count() floats per member):
 * W1[256][32] b1[256] W2[256][256] b2[256] W3[256][256] b3[256]
 * W4[256][256] b4[256] with W_l[out][in]; W1 columns 5..31 and W4/b4 rows
 * 210..255 are padding and stay 0. */
unsigned long ophip_mlp_create(int members, int batch_cap);
void ophip_mlp_free(unsigned long net);
int  ophip_mlp_param_count(void);
/* which: 0 weights, 1 gradient, 2/3 Adam moments (member's kPar floats),
 * 4/5 output mean/scale (256 floats); uploading weights resets Adam's t */
int  ophip_mlp_params(unsigned long net, int member, int which, float *host, int to_device);
/* Y[e][b][256] for every member e from normalized inputs X (B x 5) */
int  ophip_mlp_forward(unsigned long net, int B, unsigned X, unsigned Y);
/* gradients: member e trains on rows idx[e*B + b] of X (rows x 5) and T
 * (rows x 210), loss sum (y - T)^2 / (B * 210); or, with G != 0, the
 * gradient of sum(y .* G), G members x B x 256 */
int  ophip_mlp_backward(unsigned long net, int B, unsigned idx, unsigned X, unsigned T,
                        unsigned G);
int  ophip_mlp_adam(unsigned long net, float lr, float beta1, float beta2, float eps);
/* squared-error sums of the last MSE backward, one double per member */
int  ophip_mlp_loss(unsigned long net, int B, double *out);
/* surrogate evaluation of parameter draws prm (B x 5: sigma0, beta, r, a,
 * b): prices (B x 210, maturity-major) = Black-Scholes control variate at
 * the local vol of sqrt(spot K) + mean + scale * ensemble mean; spread =
 * max over outputs of the ensemble std / max(|price|, 1e-3 max|price|);
 * flags = spread > tol (draws for the operator solver) */
int  ophip_surr_eval(unsigned long net, int B, unsigned prm, const float *lo, const float *hi,
                     float spot, float dT, int nmat, const float *strikes, int nstrikes,
                     float tol, unsigned prices, unsigned spread, unsigned flags);
/* the control variate alone (B x 210) */
int  ophip_surr_base(int B, unsigned prm, float spot, float dT, int nmat, const float *strikes,
                     int nstrikes, unsigned base);

/* Monte Carlo */
int ophip_mc_set_nrng(int nstreams);
int ophip_mc1f_rows(unsigned mt, int y0, unsigned y_sk, int nscen, int nk, int d,
                    unsigned crow, unsigned m_k, unsigned long yhost);
/* guide tables over row-contiguous cumulative tables (Chen & Asau
 * indexed search, 256 buckets): guide = nm * d * 257 int16, guide[b] =
 * min{j : b/256 < C[j]}; the guided sampler returns the same states as
 * ophip_mc1f_rows, bit for bit, with fewer dependent table probes */
int ophip_mc_guide(int d, int nm, unsigned crow, unsigned guide);
/* The reference's GPU generator (smc.cu:87-156, 664-736; MersenneTwister.h):
 * 4096 twisted-GFSR lanes of period 2^607-1 (19-word states, lane-major
 * st[lane + 4096*i]) with per-lane parameters in MersenneTwister.dat's format
 * (4096 x {matrix_a, mask_b, mask_c, seed} uint32; ophip_dcmt_load returns 1
 * on a size mismatch).  ophip_dcmt_setseed: states from 4096 host seeds;
 * ophip_dcmt_uniforms: nscen (% 4096 == 0) uniforms x/2^32, lane-interleaved
 * unif[lane + 4096*c], as opcuda_mt_benchmark's stream in the reference. */
int ophip_dcmt_load(const void *table, long bytes);
int ophip_dcmt_status_sz(void);
int ophip_dcmt_setseed(const unsigned *seeds, unsigned st);
int ophip_dcmt_uniforms(unsigned st, unsigned unif, int nscen);
/* Monte-Carlo payoffs evaluated on the device instead of after a copy of
 * every batch of states to the host (smc.cu:653, CModel Device.cs:352-363;
 * VanillaPayoffEvaluator OPPricer.cs:22-53, discounting OPPricer.cs:143-146):
 * for the listed steps (each < nk; y_sk holds nk steps of nscen paths,
 * time-major int16) and strikes, out[i*nstrikes + c] = df[i] * mean over
 * paths of max(xval[y] - K_c, 0) and out[(nsteps + i)*nstrikes + c] =
 * df[i]^2 * the mean of its square (FP64, fixed summation order).  xval: d
 * FP32 state values; nsteps, nstrikes <= 64; out: 2*nsteps*nstrikes doubles */
int ophip_mc_payoff(unsigned y_sk, int nscen, int nk, int nsteps, const int *steps, const double *df,
                    unsigned xval, int d, const float *strikes, int nstrikes, unsigned out);
int ophip_mc1f_guided(unsigned mt, int y0, unsigned y_sk, int nscen, int nk, int d, unsigned crow,
                      unsigned guide, unsigned m_k);
/* ophip_mc_payoff without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops).  The partial sums live in a fixed
 * 4 MiB array of the library on each device it runs on (held until the
 * process ends): ceil(nscen / 4096) * nsteps <= 4096 (e.g. 2^20 paths at
 * 16 listed steps), larger calls return OPHIP_ERR_BAD_ARG; the library
 * must already be initialised on the device (OPHIP_ERR_NO_DEVICE
 * otherwise).  Same arguments, kernels and bits as ophip_mc_payoff.  Calls
 * on one device share the array: calls that overlap on one device
 * (issued from several host threads at once) race. */
int ophip_mc_payoff_noalloc(unsigned y_sk, int nscen, int nk, int nsteps, const int *steps,
                            const double *df, unsigned xval, int d, const float *strikes, int nstrikes,
                            unsigned out);

/* Read-bandwidth probe (a measurement aid; no reference counterpart): one
 * streaming read of `bytes` (a multiple of 16) at handle x on the library
 * stream; out: 32 KiB scratch.  Timed with ophip_event_*, a buffer that
 * fits the 256 MB Infinity Cache and was just read gives the cache's read
 * bandwidth, a larger one HBM's. */
int ophip_read_probe(unsigned x, unsigned long bytes, unsigned out);

/* graphs, events, workspace */
int           ophip_reserve_workspace(unsigned long bytes);
int           ophip_capture_begin(void);
unsigned long ophip_capture_end(void);
int           ophip_graph_launch(unsigned long graph_exec);
void          ophip_graph_destroy(unsigned long graph_exec);
int           ophip_event_record(int slot);
double        ophip_event_elapsed_ms(int a, int b);

#ifdef __cplusplus
}
#endif
#endif
