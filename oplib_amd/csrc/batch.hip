// batch.hip -- pricing kernels for many lattices in one launch (the
// calibration sweep).  No reference counterpart beyond the role they play:
// the reference prices one model at a time with GEMV (sgemv.cu:44-62) after
// its batched squarings (sgemm4.cu:84-154).
//
//   k_fwd_step   one forward step of the state distribution of every
//                lattice: p_k = p_{k-1} U (U = I + W in the W form),
//                one wavefront per (lattice, column)
//   k_price      df_k * sum_j p_k[j] * max(x_j - K, 0) for every (lattice,
//                maturity, strike), one wavefront each
#include "ophip_internal.h"

using namespace ophip;

namespace {

struct FwdScen {
    int d, plus_identity, nmat, pad;
    const float *K;   // U or W, column-major d x d
    float *P;         // (nmat + 1) x d row-major: row 0 = initial distribution
    const float *x;   // grid values (FP32), d
    float *out;       // nmat x nstrikes prices
    float rate, dT;
    const uint8_t *fl;  // optional 64 x 64 cell flags of K ([d/64][d/64], 0 = all zero)
};

__global__ __launch_bounds__(256) void k_fwd_step(const FwdScen *__restrict__ sc, int k) {
    const FwdScen s = sc[blockIdx.y];
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= s.d) return;
    const float *p = s.P + (size_t)(k - 1) * s.d;
    const float *col = s.K + (size_t)s.d * j;
    float acc = 0.0f;
    if (s.fl) {
        // only the 64-row chunks of this column whose cell may be nonzero
        // (a zero cell is exactly 0: the skipped terms add nothing); d % 64
        // == 0 and d <= 4096
        const int nc = s.d / 64;
        const uint64_t m = __ballot(lane < nc && s.fl[(size_t)lane * nc + j / 64] != 0);
        for (int c = 0; c < nc; c++)
            if ((m >> c) & 1) acc += p[64 * c + lane] * col[64 * c + lane];
    } else {
        for (int i = lane; i < s.d; i += 64) acc += p[i] * col[i];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) s.P[(size_t)k * s.d + j] = (s.plus_identity ? p[j] : 0.0f) + acc;
}

__global__ __launch_bounds__(256) void k_price(const FwdScen *__restrict__ sc,
                                               const float *__restrict__ strikes, int nstrikes) {
    const FwdScen s = sc[blockIdx.z];
    const int kk = blockIdx.y;  // maturity index 0..nmat-1
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= nstrikes || kk >= s.nmat) return;
    const float K = strikes[c];
    const float *p = s.P + (size_t)(kk + 1) * s.d;
    float acc = 0.0f;
    for (int i = lane; i < s.d; i += 64) acc += p[i] * fmaxf(s.x[i] - K, 0.0f);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) s.out[(size_t)kk * nstrikes + c] = expf(-s.rate * s.dT * (kk + 1)) * acc;
}

// prices of a few spot rows of one kernel: out[r][c] = df (plus * pay(y0_r,
// c) + sum_j U[y0_r][j] pay(j, c)) with U column-major and pay(j, c) =
// payT[c + nstrikes j].  k_spot_part: workgroup (r, s) takes the columns
// j of chunk s, one per thread, and writes the chunk's partial sums for
// every strike (its 4 waves reduced in order); k_spot_sum adds the chunks
// in order.
struct SpotRows {
    int n, pad;
    int y0[16];
};
constexpr int kSpotChunks = 16;

__global__ __launch_bounds__(256) void k_spot_part(int d, const float *__restrict__ U, SpotRows rows,
                                                   const float *__restrict__ payT, int nstrikes,
                                                   float *__restrict__ part) {
    __shared__ float red[4][64];
    const int r = blockIdx.x, sc = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int y0 = rows.y0[r];
    const int len = (d + kSpotChunks - 1) / kSpotChunks;
    const int jb = sc * len, je = min(d, jb + len);
    float acc[64];
#pragma unroll
    for (int c = 0; c < 64; c++) acc[c] = 0.0f;
    for (int j = jb + t; j < je; j += 256) {
        const float u = U[(size_t)y0 + (size_t)d * j];
        const float *p = payT + (size_t)nstrikes * j;
#pragma unroll
        for (int c = 0; c < 64; c++)
            if (c < nstrikes) acc[c] += u * p[c];
    }
#pragma unroll
    for (int c = 0; c < 64; c++) {
        if (c < nstrikes) {  // uniform
            float x = acc[c];
            for (int w = 32; w >= 1; w >>= 1) x += __shfl_xor(x, w, 64);
            if (lane == 0) red[wave][c] = x;
        }
    }
    __syncthreads();
    if (t < nstrikes)
        part[((size_t)r * kSpotChunks + sc) * 64 + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

__global__ void k_spot_sum(SpotRows rows, int plus_identity, const float *__restrict__ payT,
                           int nstrikes, float df, const float *__restrict__ part, float *__restrict__ out) {
    const int r = blockIdx.x, c = threadIdx.x;
    if (c >= nstrikes) return;
    float x = 0.0f;
    for (int sc = 0; sc < kSpotChunks; sc++) x += part[((size_t)r * kSpotChunks + sc) * 64 + c];
    if (plus_identity) x += payT[(size_t)nstrikes * rows.y0[r] + c];
    out[(size_t)r * nstrikes + c] = df * x;
}

struct FwdPlan {
    int n = 0, nmat = 0, nstrikes = 0, dmax = 0;
    void *dev = nullptr;
    FwdScen *sc = nullptr;
    float *strikes = nullptr;
    FwdScen *host = nullptr;  // host mirror of the scenario table
};

}  // namespace

// Plan forward pricing for n lattices.  Per lattice i: d[i], the handle of
// its squared kernel K[i] (U, or W when plus_identity[i]), a distribution
// buffer P[i] of (nmat+1)*d floats whose row 0 holds the initial
// distribution, its grid x[i] (d FP32 values), an output buffer out[i] of
// nmat*nstrikes floats, and its rate.  Maturities are k*dT, k = 1..nmat.
OPHIP_EXPORT unsigned long ophip_fwd_plan(int n, const int *d, const unsigned *K,
                                          const int *plus_identity, const unsigned *P,
                                          const unsigned *x, const unsigned *out,
                                          const float *rate, float dT, int nmat,
                                          const float *strikes, int nstrikes) {
    Device &D = dev();
    if (n <= 0 || nmat <= 0 || nstrikes <= 0) { set_status(kErrBadArg); return 0; }
    FwdPlan *pl = new FwdPlan();
    pl->n = n;
    pl->nmat = nmat;
    pl->nstrikes = nstrikes;
    FwdScen *h = new FwdScen[n];
    for (int i = 0; i < n; i++) {
        const size_t dd = (size_t)d[i] * d[i];
        if (d[i] <= 0 || !handle_range_ok(K[i], 4 * dd) ||
            !handle_range_ok(P[i], 4 * (size_t)(nmat + 1) * d[i]) ||
            !handle_range_ok(x[i], 4 * (size_t)d[i]) ||
            !handle_range_ok(out[i], 4 * (size_t)nmat * nstrikes)) {
            delete[] h;
            delete pl;
            set_status(kErrBadArg);
            return 0;
        }
        h[i] = FwdScen{};
        h[i].d = d[i];
        h[i].plus_identity = plus_identity[i] ? 1 : 0;
        h[i].nmat = nmat;
        h[i].K = ophip::P<const float>(K[i]);
        h[i].rate = rate[i];
        h[i].dT = dT;
        h[i].P = ophip::P<float>(P[i]);
        h[i].x = ophip::P<const float>(x[i]);
        h[i].out = ophip::P<float>(out[i]);
        pl->dmax = d[i] > pl->dmax ? d[i] : pl->dmax;
    }
    size_t bytes = sizeof(FwdScen) * n + sizeof(float) * nstrikes;
    hipError_t e = hipMalloc(&pl->dev, bytes);
    if (e == hipSuccess) {
        pl->sc = (FwdScen *)pl->dev;
        pl->strikes = (float *)(pl->sc + n);
        e = hipMemcpy(pl->sc, h, sizeof(FwdScen) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(pl->strikes, strikes, sizeof(float) * nstrikes, hipMemcpyHostToDevice);
    }
    pl->host = h;
    if (e != hipSuccess) {
        if (pl->dev) (void)hipFree(pl->dev);
        delete[] pl->host;
        delete pl;
        set_status(hip_status(e));
        return 0;
    }
    (void)D;
    set_status(kOk);
    return (unsigned long)(uintptr_t)pl;
}

// per lattice the cell flags of its kernel (handles of (d/64)^2 bytes,
// [row cell][column cell], 0 = every entry of the cell is 0; a 0 handle
// = none): the forward steps skip the zero cells.  d % 64 == 0, d <= 4096.
OPHIP_EXPORT int ophip_fwd_set_cells(unsigned long plan, const unsigned *flags) {
    FwdPlan *pl = (FwdPlan *)(uintptr_t)plan;
    if (!pl || !flags) return (int)set_status(kErrBadArg);
    for (int i = 0; i < pl->n; i++) {
        const int d = pl->host[i].d;
        if (flags[i] && (d % 64 || d > 4096 || !handle_range_ok(flags[i], (size_t)(d / 64) * (d / 64))))
            return (int)set_status(kErrBadArg);
    }
    for (int i = 0; i < pl->n; i++) pl->host[i].fl = flags[i] ? P<const uint8_t>(flags[i]) : nullptr;
    hipError_t e = hipStreamSynchronize(dev().stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pl->sc, pl->host, sizeof(FwdScen) * pl->n, hipMemcpyHostToDevice, dev().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev().stream);
    return (int)set_status(hip_status(e));
}

// nmat forward steps then the prices (kernel launches only)
OPHIP_EXPORT int ophip_fwd_exec(unsigned long plan) {
    Device &D = dev();
    FwdPlan *pl = (FwdPlan *)(uintptr_t)plan;
    if (!pl) return (int)set_status(kErrBadArg);
    for (int k = 1; k <= pl->nmat; k++)
        hipLaunchKernelGGL(k_fwd_step, dim3(cdiv(pl->dmax, 4), pl->n), dim3(256), 0, D.stream,
                           pl->sc, k);
    hipLaunchKernelGGL(k_price, dim3(cdiv(pl->nstrikes, 4), pl->nmat, pl->n), dim3(256), 0,
                       D.stream, pl->sc, pl->strikes, pl->nstrikes);
    return (int)check_launch();
}

OPHIP_EXPORT void ophip_fwd_plan_free(unsigned long plan) {
    FwdPlan *pl = (FwdPlan *)(uintptr_t)plan;
    if (!pl) return;
    (void)hipStreamSynchronize(dev().stream);
    (void)hipFree(pl->dev);
    delete[] pl->host;
    delete pl;
}

// new discount rates for every lattice of a plan (same plan, same graph):
// waits for the library stream, then rewrites the device table
OPHIP_EXPORT int ophip_fwd_set_rates(unsigned long plan, const float *rate) {
    FwdPlan *pl = (FwdPlan *)(uintptr_t)plan;
    if (!pl || !rate) return (int)set_status(kErrBadArg);
    for (int i = 0; i < pl->n; i++) pl->host[i].rate = rate[i];
    hipError_t e = hipStreamSynchronize(dev().stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pl->sc, pl->host, sizeof(FwdScen) * pl->n, hipMemcpyHostToDevice, dev().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev().stream);
    return (int)set_status(hip_status(e));
}

// prices at up to 16 spot rows of a kernel (U, or W when plus_identity):
// out[r * nstrikes + c], see k_spot_part; kernel launches only (the
// partials live in the library workspace)
OPHIP_EXPORT int ophip_spot_prices(int d, unsigned U, int nrow, const int *rows, int plus_identity,
                                   unsigned payT, int nstrikes, float df, unsigned out) {
    if (d <= 0 || nrow <= 0 || nrow > 16 || !rows || nstrikes <= 0 || nstrikes > 64)
        return (int)set_status(kErrBadArg);
    if (!handle_range_ok(U, 4 * (size_t)d * d) || !handle_range_ok(payT, 4 * (size_t)d * nstrikes) ||
        !handle_range_ok(out, 4 * (size_t)nrow * nstrikes))
        return (int)set_status(kErrBadHandle);
    SpotRows sr{};
    sr.n = nrow;
    for (int r = 0; r < nrow; r++) {
        if (rows[r] < 0 || rows[r] >= d) return (int)set_status(kErrBadArg);
        sr.y0[r] = rows[r];
    }
    float *part = (float *)workspace(sizeof(float) * 64 * kSpotChunks * nrow);
    if (!part) return (int)kErrOutOfArena;
    hipLaunchKernelGGL(k_spot_part, dim3(nrow, kSpotChunks), dim3(256), 0, dev().stream, d,
                       P<const float>(U), sr, P<const float>(payT), nstrikes, part);
    hipLaunchKernelGGL(k_spot_sum, dim3(nrow), dim3(64), 0, dev().stream, sr, plus_identity ? 1 : 0,
                       P<const float>(payT), nstrikes, df, (const float *)part, P<float>(out));
    return (int)check_launch();
}
