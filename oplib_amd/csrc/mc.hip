// mc.hip -- MT19937 streams and the Markov-chain Monte Carlo sampler.
//
// Drop-in entry points (smc.cu:87-156, 610-656, 731-736):
//   opcuda_mc_load_mt_gpu, opcuda_mc_nrng, opcuda_mc_status_sz,
//   opcuda_mc_setseed, opcuda_mc1f, opcuda_mt_benchmark
// MI355X additions:
//   ophip_mc_set_nrng   number of streams in a state buffer
//   ophip_mc1f_rows     sampler over row-contiguous cumulative tables
//
// Generator.  The reference runs 4096 per-lane DCMT generators of period
// 2^607-1 (MersenneTwister.h:8-29) whose parameter file is absent and whose
// loader is a no-op (smc.cu:87-96).  This library runs MT19937 (Matsumoto &
// Nishimura 1998), one stream per wavefront: the 624-word state lives in
// LDS and the 64 lanes twist it together in three dependency-free chunks
// (i < 227 reads only old words; 227 <= i < 454 and 454 <= i < 624 read
// words produced by the previous chunk).  State buffer: stream s occupies
// words [625*s, 625*s + 625): 624 state words and the output index.
//
// Path mapping (fixed; the CPU oracle reproduces it): with nscen paths
// over nrng streams, stream s serves per = nscen/nrng paths, scenario
// s*per + j, in groups of 64 consecutive paths (the last group holds
// per mod 64 when that is not 0).  A group's chains run step by step; at
// step k its n paths take the stream's next n outputs in path order, so a
// call consumes exactly per*nk outputs of each stream (the reference's
// lanes consume one output per path and step too, smc.cu:195-250).
// Uniform u = (x >> 8) * 2^-24 in [0,1).
// Chain: y_{-1} = y0, y_k = min{ j : u_k < C[m_k[k]][y_{k-1}, j] } clamped
// to d-1 (the inverse CDF; the reference's fixed-count bisection from
// lb = 0, smc.cu:224-238, returns the same state, 0 included).  Output is
// time-major int16 y_sk[k*nscen + scen] (smc.cu:222).
#include "ophip_internal.h"

#include <mutex>
#include <vector>

extern "C" unsigned opcuda_mem_alloc(unsigned sz);
extern "C" unsigned opcuda_get_status(void);
extern "C" int opcuda_memcpy_h2d(unsigned d, unsigned long h, unsigned sz);

using namespace ophip;

namespace {

constexpr int kMTN = 624;
constexpr int kMTW = 625;  // words per stream in the state buffer
int g_nrng = 4096;

__device__ inline uint32_t mt_f(uint32_t a, uint32_t b, uint32_t m) {
    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return m ^ (y >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
}

// dst = twist(src); 64 lanes; three chunks separated by barriers
__device__ inline void mt_twist(const uint32_t *src, uint32_t *dst, int lane) {
    for (int i = lane; i < 227; i += 64) dst[i] = mt_f(src[i], src[i + 1], src[i + 397]);
    __syncthreads();
    for (int i = 227 + lane; i < 454; i += 64) dst[i] = mt_f(src[i], src[i + 1], dst[i - 227]);
    __syncthreads();
    for (int i = 454 + lane; i < kMTN; i += 64)
        dst[i] = mt_f(src[i], (i == kMTN - 1) ? dst[0] : src[i + 1], dst[i - 227]);
    __syncthreads();
}

__device__ inline uint32_t mt_temper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// One wavefront's view of its stream: two LDS buffers and the index.
struct WaveMT {
    uint32_t *buf0, *buf1;
    int idx, cur;
    __device__ void load(const uint32_t *st, int lane) {
        for (int i = lane; i < kMTN; i += 64) buf0[i] = st[i];
        idx = (int)st[kMTN];
        cur = 0;
        __syncthreads();
    }
    __device__ void store(uint32_t *st, int lane) {
        const uint32_t *b = cur ? buf1 : buf0;
        __syncthreads();
        for (int i = lane; i < kMTN; i += 64) st[i] = b[i];
        if (lane == 0) st[kMTN] = (uint32_t)idx;
    }
    // n <= 64 consecutive outputs, lane l < n gets output idx + l (lanes
    // >= n read a value they must not use).  Whenever this call reaches
    // the end of the current state (idx + n >= 624, including landing
    // exactly on 624) the next state is twisted first, because the
    // buffers swap at the end of the call.
    __device__ uint32_t next(int lane, int n) {
        uint32_t *c = cur ? buf1 : buf0;
        uint32_t *nb = cur ? buf0 : buf1;
        if (idx + n >= kMTN) mt_twist(c, nb, lane);
        int pos = idx + lane;
        uint32_t x = (pos < kMTN) ? c[pos] : nb[pos - kMTN < kMTN ? pos - kMTN : 0];
        idx += n;
        if (idx >= kMTN) {
            idx -= kMTN;
            cur ^= 1;
        }
        return mt_temper(x);
    }
    __device__ uint32_t next64(int lane) { return next(lane, 64); }
};

__device__ inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// init_genrand for every stream: one thread per stream
__global__ void k_mt_seed(int nrng, const uint32_t *__restrict__ seeds, uint32_t *__restrict__ st) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nrng) return;
    uint32_t *p = st + (size_t)s * kMTW;
    uint32_t v = seeds[s];
    p[0] = v;
    for (uint32_t i = 1; i < kMTN; i++) {
        v = 1812433253u * (v ^ (v >> 30)) + i;
        p[i] = v;
    }
    p[kMTN] = kMTN;
}

template <bool ROWS>
__global__ __launch_bounds__(64) void k_mc1f(uint32_t *__restrict__ mt, int per, int nk, int d,
                                             int y0, const float *__restrict__ ctpk,
                                             const int *__restrict__ m_k,
                                             int16_t *__restrict__ y_sk, size_t nscen, int niter) {
    __shared__ uint32_t b0[kMTN], b1[kMTN];
    const int lane = threadIdx.x;
    const int s = blockIdx.x;
    WaveMT w{b0, b1, 0, 0};
    uint32_t *st = mt + (size_t)s * kMTW;
    w.load(st, lane);
    const size_t d2 = (size_t)d * d;
    for (int j0 = 0; j0 < per; j0 += 64) {
        const int n = min(64, per - j0);  // paths of this group
        const bool live = lane < n;
        int y = y0;
        const size_t scen = (size_t)s * per + j0 + lane;
        for (int k = 0; k < nk; k++) {
            const float u = u01(w.next(lane, n));
            const float *tab = ctpk + (size_t)m_k[k] * d2;
            const float *row = ROWS ? tab + (size_t)y * d : tab + y;
            const size_t stride = ROWS ? 1 : (size_t)d;
            int lo = -1, hi = d - 1;
            for (int it = 0; it < niter; it++) {
                if (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (u < row[stride * mid]) hi = mid;
                    else lo = mid;
                }
            }
            y = live ? hi : y0;
            if (live) y_sk[(size_t)k * nscen + scen] = (int16_t)y;
        }
    }
    w.store(st, lane);
}

// Guide tables (indexed search, Chen & Asau 1974) over row-contiguous
// cumulative tables: guide[m][row][b] = min{j : b/NB < C[j]} (clamped to
// d-1), b = 0..NB.  For a uniform u in [b/NB, (b+1)/NB) the sampled state
// min{j : u < C[j]} lies in [guide[b], guide[b+1]]: every j below guide[b]
// has C[j] <= b/NB <= u, and C[guide[b+1]] > (b+1)/NB > u.  The search on
// that range uses the same comparison, so the states are the same bits as
// the full bisection.  u = m 2^-24 and NB = 2^8, so b = floor(u NB) is exact.
constexpr int kGuideNB = 256;

__global__ void k_mc_guide(int d, size_t nrows, const float *__restrict__ crow,
                           int16_t *__restrict__ guide) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // (row, b)
    if (i >= nrows * (kGuideNB + 1)) return;
    const size_t rowg = i / (kGuideNB + 1);
    const int b = (int)(i - rowg * (kGuideNB + 1));
    const float t = (float)b * (1.0f / kGuideNB);
    const float *row = crow + rowg * d;
    int lo = -1, hi = d - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t < row[mid]) hi = mid;
        else lo = mid;
    }
    guide[i] = (int16_t)hi;
}

__global__ __launch_bounds__(64) void k_mc1f_guided(uint32_t *__restrict__ mt, int per, int nk, int d,
                                                    int y0, const float *__restrict__ crow,
                                                    const int16_t *__restrict__ guide,
                                                    const int *__restrict__ m_k,
                                                    int16_t *__restrict__ y_sk, size_t nscen) {
    __shared__ uint32_t b0[kMTN], b1[kMTN];
    const int lane = threadIdx.x;
    const int s = blockIdx.x;
    WaveMT w{b0, b1, 0, 0};
    uint32_t *st = mt + (size_t)s * kMTW;
    w.load(st, lane);
    const size_t d2 = (size_t)d * d, gstride = (size_t)d * (kGuideNB + 1);
    for (int j0 = 0; j0 < per; j0 += 64) {
        const int n = min(64, per - j0);
        const bool live = lane < n;
        int y = y0;
        const size_t scen = (size_t)s * per + j0 + lane;
        for (int k = 0; k < nk; k++) {
            const float u = u01(w.next(lane, n));
            const int m = m_k[k];
            const float *row = crow + (size_t)m * d2 + (size_t)y * d;
            const int16_t *gd = guide + (size_t)m * gstride + (size_t)y * (kGuideNB + 1);
            const int b = (int)(u * (float)kGuideNB);
            int lo = (int)gd[b] - 1, hi = gd[b + 1];
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (u < row[mid]) hi = mid;
                else lo = mid;
            }
            y = live ? hi : y0;
            if (live) y_sk[(size_t)k * nscen + scen] = (int16_t)y;
        }
    }
    w.store(st, lane);
}

// stream s writes its next per outputs to unif[s + nrng*j], j < per
// (lane-interleaved, the reference's layout, smc.cu:712), 64 at a time
__global__ __launch_bounds__(64) void k_mt_bench(uint32_t *__restrict__ mt, int per, int nrng,
                                                 float *__restrict__ unif) {
    __shared__ uint32_t b0[kMTN], b1[kMTN];
    const int lane = threadIdx.x;
    const int s = blockIdx.x;
    WaveMT w{b0, b1, 0, 0};
    uint32_t *st = mt + (size_t)s * kMTW;
    w.load(st, lane);
    for (int j0 = 0; j0 < per; j0 += 64) {
        const int n = min(64, per - j0);
        const float u = u01(w.next(lane, n));
        if (lane < n) unif[(size_t)s + (size_t)nrng * (j0 + lane)] = u;
    }
    w.store(st, lane);
}

int ceil_log2(int d) {
    int n = 0;
    while ((1 << n) < d) n++;
    return n;
}

int mc_run(bool rows, unsigned mtptr, int y0, unsigned y_skp, int nscen, int nk, int d,
           unsigned ctpk, unsigned m_kp, unsigned long yhost) {
    Device &D = dev();
    const int nrng = g_nrng;
    if (nscen <= 0 || nk <= 0 || d <= 0 || d > 32767 || y0 < 0 || y0 >= d || nscen % nrng != 0)
        return (int)set_status(kErrBadArg);
    uint32_t *mt = P<uint32_t>(mtptr);
    int16_t *ysk = P<int16_t>(y_skp);
    const float *tab = P<const float>(ctpk);
    const int *mk = P<const int>(m_kp);
    if (!mt || !ysk || !tab || !mk) return (int)set_status(kErrBadHandle);
    if (!handle_range_ok(mtptr, (size_t)nrng * kMTW * 4) ||
        !handle_range_ok(y_skp, (size_t)nscen * nk * 2) || !handle_range_ok(m_kp, (size_t)nk * 4))
        return (int)set_status(kErrBadHandle);
    if (!handle_range_ok(ctpk, (size_t)4 * d * d)) return (int)set_status(kErrBadHandle);
    const int per = nscen / nrng;
    const int niter = ceil_log2(d);
    if (rows)
        hipLaunchKernelGGL(k_mc1f<true>, dim3(nrng), dim3(64), 0, D.stream, mt, per, nk, d, y0, tab,
                           mk, ysk, (size_t)nscen, niter);
    else
        hipLaunchKernelGGL(k_mc1f<false>, dim3(nrng), dim3(64), 0, D.stream, mt, per, nk, d, y0,
                           tab, mk, ysk, (size_t)nscen, niter);
    unsigned st = check_launch();
    if (st) return (int)st;
    if (yhost) {
        // the reference copies every batch to the host (smc.cu:653)
        hipError_t e = hipMemcpyAsync((void *)(uintptr_t)yhost, ysk, (size_t)nscen * nk * 2,
                                      hipMemcpyDeviceToHost, D.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(D.stream);
        return (int)set_status(hip_status(e));
    }
    return 0;
}

}  // namespace

// MT19937 has fixed parameters: a DCMT parameter stream has nothing to
// configure.  Accepted and ignored (the reference's loader copies the
// buffer onto itself and changes nothing, smc.cu:89-94).
OPHIP_EXPORT int opcuda_mc_load_mt_gpu(char *stream, long sz) {
    (void)stream;
    (void)sz;
    return (int)set_status(kOk);
}

OPHIP_EXPORT int opcuda_mc_nrng(void) { return g_nrng; }

OPHIP_EXPORT int opcuda_mc_status_sz(void) { return g_nrng * kMTW; }

OPHIP_EXPORT int ophip_mc_set_nrng(int n) {
    if (n <= 0 || n > (1 << 20)) return (int)set_status(kErrBadArg);
    g_nrng = n;
    return (int)set_status(kOk);
}

// seeds: host array of opcuda_mc_nrng() uint32 seeds (smc.cu:137-156);
// stream s is init_genrand(seed[s]).
OPHIP_EXPORT void opcuda_mc_setseed(unsigned long host_seedptr, unsigned mtptr) {
    Device &D = dev();
    const int nrng = g_nrng;
    uint32_t *mt = P<uint32_t>(mtptr);
    if (!host_seedptr || !mt || !handle_range_ok(mtptr, (size_t)nrng * kMTW * 4)) {
        set_status(kErrBadHandle);
        return;
    }
    uint32_t *seeds = (uint32_t *)workspace((size_t)nrng * 4);
    if (!seeds) return;
    hipError_t e = hipMemcpyAsync(seeds, (const void *)(uintptr_t)host_seedptr, (size_t)nrng * 4,
                                  hipMemcpyHostToDevice, D.stream);
    if (e != hipSuccess) { set_status(hip_status(e)); return; }
    hipLaunchKernelGGL(k_mt_seed, dim3(cdiv(nrng, 64)), dim3(64), 0, D.stream, nrng, seeds, mt);
    if (check_launch()) return;
    e = hipStreamSynchronize(D.stream);
    set_status(hip_status(e));
}

// Sample nk steps of nscen chains on the reference's column-major table
// layout ctpk[m*d*d + yprev + d*y].  nscen must be a multiple of
// opcuda_mc_nrng() (4096 by default: the reference's requirement,
// CModel Device.cs:345, smc.cu:616).  d up to 32767 (int16 states; the
// reference stops at d < 1024, smc.cu:622).  yhost = 0 leaves the states
// on the device (no host round trip).
OPHIP_EXPORT int opcuda_mc1f(unsigned mtptr, int y0, unsigned y_sk, int nscen, int nk, int d,
                             unsigned ctpk, unsigned m_k, unsigned long yhost) {
    return mc_run(false, mtptr, y0, y_sk, nscen, nk, d, ctpk, m_k, yhost);
}

// Same chains on a row-contiguous table crow[m*d*d + yprev*d + y]
// (ophip_sck_rows): identical states given identical tables.
OPHIP_EXPORT int ophip_mc1f_rows(unsigned mtptr, int y0, unsigned y_sk, int nscen, int nk, int d,
                                 unsigned crow, unsigned m_k, unsigned long yhost) {
    return mc_run(true, mtptr, y0, y_sk, nscen, nk, d, crow, m_k, yhost);
}

// nscen uniforms: stream s writes its next per = nscen/nrng outputs to
// unif[s + nrng*j], j < per (lane-interleaved, as smc.cu:712 writes
// unif_s[iRng + nRng*scen]).  nscen % nrng == 0; each call consumes
// exactly per outputs of every stream.
OPHIP_EXPORT int opcuda_mt_benchmark(unsigned mtptr, unsigned unif, int nscen) {
    Device &D = dev();
    const int nrng = g_nrng;
    if (nscen <= 0 || nscen % nrng) return (int)set_status(kErrBadArg);
    uint32_t *mt = P<uint32_t>(mtptr);
    float *u = P<float>(unif);
    if (!mt || !u || !handle_range_ok(unif, (size_t)nscen * 4) ||
        !handle_range_ok(mtptr, (size_t)nrng * kMTW * 4))
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_mt_bench, dim3(nrng), dim3(64), 0, D.stream, mt, nscen / nrng, nrng, u);
    return (int)check_launch();
}

// guide tables for nm row-contiguous cumulative tables (ophip_sck_rows):
// guide = nm * d * 257 int16 (d <= 32767)
OPHIP_EXPORT int ophip_mc_guide(int d, int nm, unsigned crow, unsigned guide) {
    if (d <= 0 || d > 32767 || nm <= 0) return (int)set_status(kErrBadArg);
    const size_t nrows = (size_t)nm * d;
    if (!handle_range_ok(crow, 4 * nrows * d) || !handle_range_ok(guide, 2 * nrows * (kGuideNB + 1)))
        return (int)set_status(kErrBadHandle);
    const size_t n = nrows * (kGuideNB + 1);
    hipLaunchKernelGGL(k_mc_guide, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, dev().stream, d, nrows,
                       P<const float>(crow), P<int16_t>(guide));
    return (int)check_launch();
}

// ophip_mc1f_rows with guide tables: the same states, bit for bit, in
// fewer dependent probes
OPHIP_EXPORT int ophip_mc1f_guided(unsigned mtptr, int y0, unsigned y_skp, int nscen, int nk, int d,
                                   unsigned crow, unsigned guide, unsigned m_kp) {
    Device &D = dev();
    const int nrng = g_nrng;
    if (nscen <= 0 || nk <= 0 || d <= 0 || d > 32767 || y0 < 0 || y0 >= d || nscen % nrng != 0)
        return (int)set_status(kErrBadArg);
    uint32_t *mt = P<uint32_t>(mtptr);
    int16_t *ysk = P<int16_t>(y_skp);
    const float *tab = P<const float>(crow);
    const int16_t *gd = P<const int16_t>(guide);
    const int *mk = P<const int>(m_kp);
    if (!mt || !ysk || !tab || !gd || !mk) return (int)set_status(kErrBadHandle);
    if (!handle_range_ok(mtptr, (size_t)nrng * kMTW * 4) ||
        !handle_range_ok(y_skp, (size_t)nscen * nk * 2) || !handle_range_ok(m_kp, (size_t)nk * 4))
        return (int)set_status(kErrBadHandle);
    if (!handle_range_ok(crow, (size_t)4 * d * d) || !handle_range_ok(guide, (size_t)2 * d * (kGuideNB + 1)))
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_mc1f_guided, dim3(nrng), dim3(64), 0, D.stream, mt, nscen / nrng, nk, d, y0, tab, gd,
                       mk, ysk, (size_t)nscen);
    return (int)check_launch();
}

// ---------------------------------------------------------------------
// Monte-Carlo payoff evaluation on the device.  The reference copies every
// batch of states to the host (smc.cu:653, CModel Device.cs:352-363) and
// evaluates there: VanillaPayoffEvaluator (OPPricer.cs:22-53) takes
// max(x[y_T] - K, 0) per path at the last step, the mean over paths, and
// discounts it by df(T) (OPPricer.cs:143-146).  Here the states stay in
// HBM: for each listed step i and strike c,
//   out[i][c]         = df_i   mean_s max(x[y_sk[step_i][s]] - K_c, 0)
//   out[n + i][c]     = df_i^2 mean_s max(...)^2   (n = nsteps: the second
//                                                    moment, for the
//                                                    standard error)
// in FP64, summed in a fixed order (chunks of kPayChunk paths: per chunk
// four threads per strike over interleaved quarters, then the quarters,
// then the chunks, in index order), so results do not depend on timing.
// ---------------------------------------------------------------------
namespace {

constexpr int kPayMax = 64;      // listed steps and strikes per call
constexpr int kPayChunk = 4096;  // paths per workgroup

struct PayArgs {
    int nsteps, nstrikes;
    int steps[kPayMax];
    float strikes[kPayMax];
    double df[kPayMax];
};

__global__ __launch_bounds__(256) void k_mc_pay_part(const int16_t *__restrict__ ysk, size_t nscen,
                                                     const float *__restrict__ xval, int d, const PayArgs a,
                                                     int nchunk, double *__restrict__ part) {
    __shared__ float xs[kPayChunk];
    __shared__ double red[256][2];
    const int i = blockIdx.y, chunk = blockIdx.x, t = threadIdx.x;
    const int16_t *y = ysk + (size_t)a.steps[i] * nscen;
    const size_t s0 = (size_t)chunk * kPayChunk;
    const int n = (int)((nscen - s0) < (size_t)kPayChunk ? nscen - s0 : (size_t)kPayChunk);
    for (int p = t; p < n; p += 256) {
        int v = y[s0 + p];
        v = v < 0 ? 0 : (v >= d ? d - 1 : v);  // states outside the lattice read its edge
        xs[p] = xval[v];
    }
    __syncthreads();
    // thread t: strike c = t / 4 (when c < nstrikes), paths q, q + 4, ... with q = t % 4
    const int c = t >> 2, q = t & 3;
    double s1 = 0.0, s2 = 0.0;
    if (c < a.nstrikes) {
        const float K = a.strikes[c];
        for (int p = q; p < n; p += 4) {
            const double v = (double)fmaxf(xs[p] - K, 0.0f);
            s1 += v;
            s2 += v * v;
        }
    }
    red[t][0] = s1;
    red[t][1] = s2;
    __syncthreads();
    if (q == 0 && c < a.nstrikes) {
        const double r1 = ((red[t][0] + red[t + 1][0]) + red[t + 2][0]) + red[t + 3][0];
        const double r2 = ((red[t][1] + red[t + 1][1]) + red[t + 2][1]) + red[t + 3][1];
        double *o = part + ((size_t)i * nchunk + chunk) * 2 * kPayMax;
        o[c] = r1;
        o[kPayMax + c] = r2;
    }
}

__global__ void k_mc_pay_sum(size_t nscen, const PayArgs a, int nchunk, const double *__restrict__ part,
                             double *__restrict__ out) {
    const int i = blockIdx.x, c = threadIdx.x;
    if (c >= a.nstrikes) return;
    double r1 = 0.0, r2 = 0.0;
    for (int k = 0; k < nchunk; k++) {
        const double *o = part + ((size_t)i * nchunk + k) * 2 * kPayMax;
        r1 += o[c];
        r2 += o[kPayMax + c];
    }
    const double inv = 1.0 / (double)nscen, df = a.df[i];
    out[(size_t)i * a.nstrikes + c] = df * r1 * inv;
    out[(size_t)(a.nsteps + i) * a.nstrikes + c] = df * df * r2 * inv;
}

}  // namespace

// European call payoffs of nscen sampled paths (y_sk: time-major int16
// states, as opcuda_mc1f / ophip_mc1f_rows / ophip_mc1f_guided write them)
// at nsteps listed steps (each < the steps y_sk holds) for nstrikes
// strikes; xval: handle of d FP32 state values.  out: handle of 2 * nsteps
// * nstrikes doubles (discounted means, then discounted second moments).
// nsteps, nstrikes <= 64.  Kernel launches only (partials in the library
// workspace): graph-capturable.
OPHIP_EXPORT int ophip_mc_payoff(unsigned y_skp, int nscen, int nk, int nsteps, const int *steps,
                                 const double *df, unsigned xvalp, int d, const float *strikes, int nstrikes,
                                 unsigned outp) {
    if (nscen <= 0 || nk <= 0 || nsteps <= 0 || nsteps > kPayMax || nstrikes <= 0 || nstrikes > kPayMax ||
        d <= 0 || d > 32767 || !steps || !df || !strikes)
        return (int)set_status(kErrBadArg);
    PayArgs a{};
    a.nsteps = nsteps;
    a.nstrikes = nstrikes;
    for (int i = 0; i < nsteps; i++) {
        if (steps[i] < 0 || steps[i] >= nk) return (int)set_status(kErrBadArg);
        a.steps[i] = steps[i];
        a.df[i] = df[i];
    }
    for (int c = 0; c < nstrikes; c++) a.strikes[c] = strikes[c];
    if (!handle_range_ok(y_skp, (size_t)nscen * nk * 2) || !handle_range_ok(xvalp, (size_t)d * 4) ||
        !handle_range_ok(outp, (size_t)2 * nsteps * nstrikes * 8))
        return (int)set_status(kErrBadHandle);
    const int nchunk = (int)(((size_t)nscen + kPayChunk - 1) / kPayChunk);
    double *part = (double *)workspace(sizeof(double) * 2 * kPayMax * (size_t)nchunk * nsteps);
    if (!part) return (int)kErrOutOfArena;
    Device &D = dev();
    hipLaunchKernelGGL(k_mc_pay_part, dim3((unsigned)nchunk, (unsigned)nsteps), dim3(256), 0, D.stream,
                       P<const int16_t>(y_skp), (size_t)nscen, P<const float>(xvalp), d, a, nchunk, part);
    hipLaunchKernelGGL(k_mc_pay_sum, dim3((unsigned)nsteps), dim3(64), 0, D.stream, (size_t)nscen, a, nchunk,
                       (const double *)part, P<double>(outp));
    return (int)check_launch();
}

// ---------------------------------------------------------------------
// ophip_mc_payoff_noalloc: the same payoff evaluation for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops): the partial sums live in `scratch`, a fixed array of this
// library (kPayScratchBytes on every device the library runs on, held
// until the process ends) instead of the library workspace, which is
// grown with hipMalloc on demand.  Its device address is looked up once
// per device (hipGetSymbolAddress with that device current) and passed to
// the same kernels.  k_mc_pay_part writes every partial slot that
// k_mc_pay_sum reads, so nothing needs clearing between calls.
// ---------------------------------------------------------------------
constexpr size_t kPayScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) double scratch[kPayScratchBytes / sizeof(double)];

namespace {

double *g_scratch[16] = {};  // per device: the device address of `scratch`
std::mutex g_scratch_mu;

double *scratch_ptr(Device &D) {
    if (D.id < 0 || D.id >= 16) {
        set_status(kErrNoDevice);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[D.id]) {
        int cur = -1;
        hipError_t e0 = hipGetDevice(&cur);
        if (e0 == hipSuccess && cur != D.id) e0 = hipSetDevice(D.id);
        if (e0 != hipSuccess) {
            set_status(hip_status(e0));
            return nullptr;
        }
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != D.id) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            set_status(hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol));
            return nullptr;
        }
        g_scratch[D.id] = (double *)p;
    }
    return g_scratch[D.id];
}

}  // namespace

// ophip_mc_payoff's arguments and results (the same kernels, the same
// bits); the partial sums (2 * 64 doubles per chunk of 4096 paths and
// listed step) must fit kPayScratchBytes: ceil(nscen / 4096) * nsteps <=
// 4096, e.g. 2^20 paths at 16 steps; larger calls are refused with
// OPHIP_ERR_BAD_ARG.  The library must be initialised on the device
// (opcuda_cublas_init, as every handle the caller passes implies); the
// call allocates nothing.  Calls on one device share the array: calls that
// overlap on one device (issued from several host threads at once) race.
OPHIP_EXPORT int ophip_mc_payoff_noalloc(unsigned y_skp, int nscen, int nk, int nsteps, const int *steps,
                                         const double *df, unsigned xvalp, int d, const float *strikes,
                                         int nstrikes, unsigned outp) {
    Device *Dp = dev_ready();
    if (!Dp) return (int)set_status(kErrNoDevice);
    if (nscen <= 0 || nk <= 0 || nsteps <= 0 || nsteps > kPayMax || nstrikes <= 0 || nstrikes > kPayMax ||
        d <= 0 || d > 32767 || !steps || !df || !strikes)
        return (int)set_status(kErrBadArg);
    const size_t nchunk = ((size_t)nscen + kPayChunk - 1) / kPayChunk;
    if (sizeof(double) * 2 * kPayMax * nchunk * (size_t)nsteps > kPayScratchBytes)
        return (int)set_status(kErrBadArg);  // beyond the fixed partial-sum array
    PayArgs a{};
    a.nsteps = nsteps;
    a.nstrikes = nstrikes;
    for (int i = 0; i < nsteps; i++) {
        if (steps[i] < 0 || steps[i] >= nk) return (int)set_status(kErrBadArg);
        a.steps[i] = steps[i];
        a.df[i] = df[i];
    }
    for (int c = 0; c < nstrikes; c++) a.strikes[c] = strikes[c];
    if (!handle_range_ok(y_skp, (size_t)nscen * nk * 2) || !handle_range_ok(xvalp, (size_t)d * 4) ||
        !handle_range_ok(outp, (size_t)2 * nsteps * nstrikes * 8))
        return (int)set_status(kErrBadHandle);
    Device &D = *Dp;
    double *part = scratch_ptr(D);
    if (!part) return (int)opcuda_get_status();
    hipLaunchKernelGGL(k_mc_pay_part, dim3((unsigned)nchunk, (unsigned)nsteps), dim3(256), 0, D.stream,
                       P<const int16_t>(y_skp), (size_t)nscen, P<const float>(xvalp), d, a, (int)nchunk, part);
    hipLaunchKernelGGL(k_mc_pay_sum, dim3((unsigned)nsteps), dim3(64), 0, D.stream, (size_t)nscen, a, (int)nchunk,
                       (const double *)part, P<double>(outp));
    return (int)check_launch();
}

// ---------------------------------------------------------------------
// The reference's GPU generator: 4096 twisted-GFSR lanes of period
// 2^607 - 1 (DCMT-607: 19-word states, MersenneTwister.h:8-29) with
// per-lane parameters {matrix_a, mask_b, mask_c, seed} in the on-disk
// format of MersenneTwister.dat (16 bytes per lane, 4096 lanes).  The
// reference's loader copies its argument onto itself (smc.cu:87-96), so
// its lanes run with zero parameters; here the table is loaded.  Seeding
// (smc.cu:118-156) and the uniform stream (smc.cu:664-736: lane-
// interleaved unif[lane + 4096 c] = x / 2^32) follow the reference,
// including its circular-buffer index restarting at 0 on every call.  The
// state buffer is lane-major, st[lane + 4096 i], 19 * 4096 words.
// ---------------------------------------------------------------------
namespace {

constexpr int kDcN = 19, kDcM = 9, kDcLanes = 4096;

struct DcmtParam {
    uint32_t matrix_a, mask_b, mask_c, seed;
};
unsigned g_dcmt[16] = {0};  // per device: arena handle of the loaded table

__global__ void k_dcmt_seed(const uint32_t *__restrict__ seeds, uint32_t *__restrict__ st) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= kDcLanes) return;
    uint32_t v = seeds[l];
    st[l] = v;
    for (int i = 1; i < kDcN; i++) {
        v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
        st[l + kDcLanes * i] = v;
    }
}

// one thread per lane; the circular buffer is walked in whole turns of 19
// with compile-time indices (the state stays in registers)
__global__ __launch_bounds__(64) void k_dcmt_unif(const DcmtParam *__restrict__ prm, uint32_t *__restrict__ st,
                                                  int count, float *__restrict__ unif) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= kDcLanes) return;
    const DcmtParam p = prm[l];
    uint32_t s[kDcN];
#pragma unroll
    for (int i = 0; i < kDcN; i++) s[i] = st[l + kDcLanes * i];
    for (int c0 = 0; c0 < count; c0 += kDcN) {
#pragma unroll
        for (int i = 0; i < kDcN; i++) {
            if (c0 + i < count) {
                const uint32_t y = (s[i] & 0xfffffffeu) | (s[(i + 1) % kDcN] & 0x1u);
                uint32_t x = s[(i + kDcM) % kDcN] ^ (y >> 1) ^ ((y & 1u) ? p.matrix_a : 0u);
                s[i] = x;
                x ^= x >> 12;
                x ^= (x << 7) & p.mask_b;
                x ^= (x << 15) & p.mask_c;
                x ^= x >> 18;
                unif[(size_t)l + (size_t)kDcLanes * (c0 + i)] = (float)x / 4294967296.0f;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kDcN; i++) st[l + kDcLanes * i] = s[i];
}

}  // namespace

// the parameter table, MersenneTwister.dat's format: 4096 records of four
// uint32 {matrix_a, mask_b, mask_c, seed}; bytes must be 65536.  Returns 0,
// or 1 on a size mismatch (the reference's convention, smc.cu:87-96).
OPHIP_EXPORT int ophip_dcmt_load(const void *table, long bytes) {
    if (!table || bytes != (long)(sizeof(DcmtParam) * kDcLanes)) return 1;
    Device &D = dev();
    if (D.id < 0 || D.id >= 16) return (int)set_status(kErrNoDevice);
    unsigned &h = g_dcmt[D.id];
    if (!h) h = opcuda_mem_alloc((unsigned)bytes);
    if (!h) return (int)opcuda_get_status();
    return opcuda_memcpy_h2d(h, (unsigned long)(uintptr_t)table, (unsigned)bytes);
}

// words of a lane state buffer (19 per lane, 4096 lanes)
OPHIP_EXPORT int ophip_dcmt_status_sz(void) { return kDcN * kDcLanes; }

// lane states from 4096 host seeds (smc.cu:118-156)
OPHIP_EXPORT int ophip_dcmt_setseed(const unsigned *seeds, unsigned st) {
    if (!seeds) return (int)set_status(kErrBadArg);
    if (!handle_range_ok(st, 4 * (size_t)kDcN * kDcLanes)) return (int)set_status(kErrBadHandle);
    uint32_t *sd = (uint32_t *)workspace(4 * (size_t)kDcLanes);
    if (!sd) return (int)kErrOutOfArena;
    Device &D = dev();
    hipError_t e = hipMemcpyAsync(sd, seeds, 4 * (size_t)kDcLanes, hipMemcpyHostToDevice, D.stream);
    if (e != hipSuccess) return (int)set_status(hip_status(e));
    hipLaunchKernelGGL(k_dcmt_seed, dim3(kDcLanes / 64), dim3(64), 0, D.stream, (const uint32_t *)sd, P<uint32_t>(st));
    e = hipStreamSynchronize(D.stream);  // the seeds' staging buffer is reused
    if (e != hipSuccess) return (int)set_status(hip_status(e));
    return (int)check_launch();
}

// nscen uniforms x / 2^32 from the loaded lanes, lane-interleaved:
// unif[lane + 4096 c], c < nscen / 4096 (opcuda_mt_benchmark's stream,
// smc.cu:664-736).  nscen % 4096 == 0.
OPHIP_EXPORT int ophip_dcmt_uniforms(unsigned st, unsigned unif, int nscen) {
    Device &D = dev();
    if (nscen <= 0 || nscen % kDcLanes) return (int)set_status(kErrBadArg);
    const unsigned tab = (D.id >= 0 && D.id < 16) ? g_dcmt[D.id] : 0;
    if (!tab) return (int)set_status(kErrBadArg);  // no table loaded
    if (!handle_range_ok(st, 4 * (size_t)kDcN * kDcLanes) || !handle_range_ok(unif, 4 * (size_t)nscen))
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_dcmt_unif, dim3(kDcLanes / 64), dim3(64), 0, D.stream, P<const DcmtParam>(tab),
                       P<uint32_t>(st), nscen / kDcLanes, P<float>(unif));
    return (int)check_launch();
}
