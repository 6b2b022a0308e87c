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
};

__global__ __launch_bounds__(256) void k_fwd_step(const FwdScen *__restrict__ sc, int k) {
    const FwdScen s = sc[blockIdx.y];
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= s.d) return;
    const float *p = s.P + (size_t)(k - 1) * s.d;
    const float *col = s.K + (size_t)s.d * j;
    float acc = 0.0f;
    for (int i = lane; i < s.d; i += 64) acc += p[i] * col[i];
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
