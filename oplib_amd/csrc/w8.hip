// w8.hip -- repeated squaring in the U = I + W form with FP8 (e4m3)
// operands and power-of-two scales (no reference counterpart: the
// reference squares U in FP32, ssqmm.cu:119-170).
//
// Each round squares W' = 2W + W*W.  The product runs on
// v_mfma_f32_16x16x32_fp8_fp8 with FP32 accumulation from a hi/lo pair of
// e4m3 copies: row i of the A operand is scaled by 2^a_i and column j of
// the B operand by 2^b_j (powers of two that bring the row / column
// maximum into [64, 128), below e4m3's 448), hi = e4m3(W 2^a), lo =
// e4m3((W 2^a - hi) 2^4), and
//     (W*W)_ij ~ 2^-(a_i + b_j) (Ah Bh + 2^-4 (Ah Bl + Al Bh))_ij.
// A pair of e4m3 values carries about 8 significant bits, so a round's
// product is accurate to ~2^-8 of |W|^2: this variant is reported for its
// error, it cannot reach 1e-4 on the lattices here.  After each round the
// diagonal is reset so that every row of W sums to zero in FP64 (rows of U
// sum to 1), as in the 16-bit modes.
//
// Operands are stored fragment-major: 16 x 32 tiles, tile-row-major; lane
// l = 16 g + r of tile (it, kt) holds the 8 bytes M[16 it + r][32 kt + 8 g
// .. +7], one 8-byte MFMA operand per lane, one contiguous 512-byte load
// per wave.
#include "ophip_internal.h"

#include <stdint.h>

using namespace ophip;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline uint32_t enc8x4(float a, float b, float c, float d) {
    int v = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    v = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, v, true);
    return (uint32_t)v;
}
__device__ inline float dec8(uint32_t word, int byte) {
    switch (byte) {
    case 0: return __builtin_amdgcn_cvt_f32_fp8((int)word, 0);
    case 1: return __builtin_amdgcn_cvt_f32_fp8((int)word, 1);
    case 2: return __builtin_amdgcn_cvt_f32_fp8((int)word, 2);
    default: return __builtin_amdgcn_cvt_f32_fp8((int)word, 3);
    }
}

// power of two bringing a maximum m > 0 into [64, 128)
__device__ inline float scale_for(float m) {
    if (!(m > 0.0f)) return 1.0f;
    int e;
    (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    return ldexpf(1.0f, 7 - e);
}

// row scales: 64 rows per workgroup, 4 threads per row split the columns
__global__ __launch_bounds__(256) void k_w8_rowscale(int d, const float *__restrict__ W,
                                                     float *__restrict__ rs) {
    __shared__ float part[4][64];
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + r;
    float m = 0.0f;
    for (int k = q; k < d; k += 4) m = fmaxf(m, fabsf(W[(size_t)i + (size_t)d * k]));
    part[q][r] = m;
    __syncthreads();
    if (q == 0) rs[i] = scale_for(fmaxf(fmaxf(part[0][r], part[1][r]), fmaxf(part[2][r], part[3][r])));
}

// column scales: one wavefront per column (contiguous)
__global__ __launch_bounds__(64) void k_w8_colscale(int d, const float *__restrict__ W,
                                                    float *__restrict__ cs) {
    const int j = blockIdx.x, lane = threadIdx.x;
    float m = 0.0f;
    for (int i = lane; i < d; i += 64) m = fmaxf(m, fabsf(W[(size_t)i + (size_t)d * j]));
    for (int w = 32; w >= 1; w >>= 1) m = fmaxf(m, __shfl_xor(m, w));
    if (lane == 0) cs[j] = scale_for(m);
}

__device__ inline size_t frag8(int it, int kt, int l, int nkt) {
    return ((size_t)(it * nkt + kt) * 64 + l) * 8;
}

// 64 x 64 tile of W (column-major) through LDS -> the A fragments (rows i,
// k = columns) and the B fragments (rows = columns j, k = rows) of the
// tile, hi and lo.  256 threads, two fragments of each kind per thread.
__global__ __launch_bounds__(256) void k_w8_split(int d, const float *__restrict__ W,
                                                  const float *__restrict__ rs,
                                                  const float *__restrict__ cs,
                                                  uint8_t *__restrict__ Ah, uint8_t *__restrict__ Al,
                                                  uint8_t *__restrict__ Bh, uint8_t *__restrict__ Bl) {
    __shared__ float T[64][65];  // T[ii][jj] = W[i0 + ii][j0 + jj]
    const int t = threadIdx.x, i0 = blockIdx.x * 64, j0 = blockIdx.y * 64, nkt = d / 32;
    for (int q = 0; q < 16; q++) {
        const int jj = q * 4 + (t >> 6), ii = t & 63;
        T[ii][jj] = W[(size_t)(i0 + ii) + (size_t)d * (j0 + jj)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int f = t + 256 * q, a = f & 63, grp = f >> 6;  // a: row (A) / column (B); grp: 8 k
        // A: row i0 + a, k = j0 + 8 grp .. +7, scaled by the row's power of two
        {
            const int i = i0 + a, k = j0 + 8 * grp;
            const float s = rs[i];
            float v[8], lo[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = T[a][8 * grp + e] * s;
            const uint32_t h0 = enc8x4(v[0], v[1], v[2], v[3]), h1 = enc8x4(v[4], v[5], v[6], v[7]);
#pragma unroll
            for (int e = 0; e < 8; e++) lo[e] = (v[e] - dec8(e < 4 ? h0 : h1, e & 3)) * 16.0f;
            const size_t o = frag8(i / 16, k / 32, 16 * ((k % 32) / 8) + (i % 16), nkt);
            *(uint2 *)(Ah + o) = uint2{h0, h1};
            *(uint2 *)(Al + o) = uint2{enc8x4(lo[0], lo[1], lo[2], lo[3]), enc8x4(lo[4], lo[5], lo[6], lo[7])};
        }
        // B^T: row = column j0 + a of W, k = rows i0 + 8 grp .. +7, scaled by the column's power of two
        {
            const int j = j0 + a, k = i0 + 8 * grp;
            const float s = cs[j];
            float v[8], lo[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = T[8 * grp + e][a] * s;
            const uint32_t h0 = enc8x4(v[0], v[1], v[2], v[3]), h1 = enc8x4(v[4], v[5], v[6], v[7]);
#pragma unroll
            for (int e = 0; e < 8; e++) lo[e] = (v[e] - dec8(e < 4 ? h0 : h1, e & 3)) * 16.0f;
            const size_t o = frag8(j / 16, k / 32, 16 * ((k % 32) / 8) + (j % 16), nkt);
            *(uint2 *)(Bh + o) = uint2{h0, h1};
            *(uint2 *)(Bl + o) = uint2{enc8x4(lo[0], lo[1], lo[2], lo[3]), enc8x4(lo[4], lo[5], lo[6], lo[7])};
        }
    }
}

__device__ inline f4 mfma8(long a, long b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
}

// W' = 2 W + W*W: 64 x 64 output tile per workgroup, 2 x 2 waves of 32 x 32
// (2 x 2 MFMA tiles), fragments straight from memory with one k32 step of
// prefetch
__global__ __launch_bounds__(256) void k_w8_sq(int d, const uint8_t *__restrict__ Ah,
                                               const uint8_t *__restrict__ Al,
                                               const uint8_t *__restrict__ Bh,
                                               const uint8_t *__restrict__ Bl,
                                               const float *__restrict__ rs, const float *__restrict__ cs,
                                               const float *__restrict__ W, float *__restrict__ Wn) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int i0 = blockIdx.x * 64 + (wave & 1) * 32, j0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int nkt = d / 32;
    const long *ah[2], *al[2], *bh[2], *bl[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
        ah[a] = (const long *)(Ah + frag8(i0 / 16 + a, 0, lane, nkt));
        al[a] = (const long *)(Al + frag8(i0 / 16 + a, 0, lane, nkt));
        bh[a] = (const long *)(Bh + frag8(j0 / 16 + a, 0, lane, nkt));
        bl[a] = (const long *)(Bl + frag8(j0 / 16 + a, 0, lane, nkt));
    }
    f4 acc[2][2], cor[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            cor[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
    long xa[2][2], xl[2][2], ya[2][2], yl[2][2];  // [stage][tile]
    auto load = [&](int st, int kt) {
#pragma unroll
        for (int a = 0; a < 2; a++) {
            xa[st][a] = ah[a][(size_t)kt * 64];
            xl[st][a] = al[a][(size_t)kt * 64];
            ya[st][a] = bh[a][(size_t)kt * 64];
            yl[st][a] = bl[a][(size_t)kt * 64];
        }
    };
    auto comp = [&](int st) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = mfma8(xa[st][a], ya[st][b], acc[a][b]);
                cor[a][b] = mfma8(xa[st][a], yl[st][b], cor[a][b]);
                cor[a][b] = mfma8(xl[st][a], ya[st][b], cor[a][b]);
            }
    };
    load(0, 0);
    for (int kt = 0; kt < nkt; kt += 2) {  // nkt is even (d % 64 == 0)
        load(1, kt + 1);
        comp(0);
        load(0, min(kt + 2, nkt - 1));  // clamped: in bounds, unused at the end
        comp(1);
    }
    // lane holds rows 4g + v of column r of each 16 x 16 tile
#pragma unroll
    for (int a = 0; a < 2; a++) {
        const int row = i0 + 16 * a + 4 * g;
        f4 ra;
#pragma unroll
        for (int v = 0; v < 4; v++) ra[v] = rs[row + v];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int col = j0 + 16 * b + r;
            const float cb = cs[col];
            const size_t e = (size_t)row + (size_t)d * col;
            const f4 w = *(const f4 *)(W + e);
            f4 o;
#pragma unroll
            for (int v = 0; v < 4; v++)
                o[v] = 2.0f * w[v] + (acc[a][b][v] + cor[a][b][v] * 0.0625f) / (ra[v] * cb);
            *(f4 *)(Wn + e) = o;
        }
    }
}

// diagonal reset: W_ii = -(sum of the row's off-diagonal entries) in FP64,
// summed in column order per quarter, quarters in order.  64 rows per
// workgroup, 4 threads per row.
__global__ __launch_bounds__(256) void k_w8_fix(int d, float *__restrict__ W) {
    __shared__ double part[4][64];
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + r, len = d / 4;
    double s = 0.0;
    for (int k = q * len; k < (q + 1) * len; k++)
        if (k != i) s += (double)W[(size_t)i + (size_t)d * k];
    part[q][r] = s;
    __syncthreads();
    if (q == 0) W[(size_t)i + (size_t)d * i] = (float)(-(((part[0][r] + part[1][r]) + part[2][r]) + part[3][r]));
}

struct W8Layout {
    size_t W[2], Ah, Al, Bh, Bl, rs, cs, total;
};

W8Layout w8_layout(int d) {
    W8Layout L{};
    size_t off = 0, dd = (size_t)d * d;
    auto take = [&](size_t b) {
        size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    L.W[0] = take(4 * dd);
    L.W[1] = take(4 * dd);
    L.Ah = take(dd);
    L.Al = take(dd);
    L.Bh = take(dd);
    L.Bl = take(dd);
    L.rs = take(4 * (size_t)d);
    L.cs = take(4 * (size_t)d);
    L.total = off;
    return L;
}

}  // namespace

// bytes of an FP8 W-form state for a d x d lattice (d % 64 == 0)
OPHIP_EXPORT unsigned long ophip_w8_bytes(int d) {
    if (d <= 0 || d % 64) return 0;
    return (unsigned long)w8_layout(d).total;
}

// handle of W buffer `buf` (0/1), FP32 column-major
OPHIP_EXPORT unsigned ophip_w8_w(unsigned state, int d, int buf) {
    if (d <= 0 || d % 64 || buf < 0 || buf > 1) return 0;
    return state + (unsigned)w8_layout(d).W[buf];
}

// handle of an operand array: which = 0 Ah, 1 Al, 2 Bh, 3 Bl (fragment-
// major bytes), 4 row scales, 5 column scales
OPHIP_EXPORT unsigned ophip_w8_array(unsigned state, int d, int which) {
    if (d <= 0 || d % 64 || which < 0 || which > 5) return 0;
    const W8Layout L = w8_layout(d);
    const size_t off[6] = {L.Ah, L.Al, L.Bh, L.Bl, L.rs, L.cs};
    return state + (unsigned)off[which];
}

// diagonal reset of W0 (written by ophip_sek_rs out_kind 2 into buffer 0),
// then nsq squaring rounds; W(2^nsq dt) ends in buffer nsq & 1.  Kernel
// launches only (graph-capturable).
OPHIP_EXPORT int ophip_w8_exec(unsigned state, int d, int nsq) {
    if (d <= 0 || d % 64 || nsq < 0 || nsq > 64) return (int)set_status(kErrBadArg);
    const W8Layout L = w8_layout(d);
    if (!handle_range_ok(state, L.total)) return (int)set_status(kErrBadHandle);
    Device &D = dev();
    float *W[2] = {P<float>(state + (unsigned)L.W[0]), P<float>(state + (unsigned)L.W[1])};
    uint8_t *Ah = P<uint8_t>(state + (unsigned)L.Ah), *Al = P<uint8_t>(state + (unsigned)L.Al);
    uint8_t *Bh = P<uint8_t>(state + (unsigned)L.Bh), *Bl = P<uint8_t>(state + (unsigned)L.Bl);
    float *rs = P<float>(state + (unsigned)L.rs), *cs = P<float>(state + (unsigned)L.cs);
    hipLaunchKernelGGL(k_w8_fix, dim3(d / 64), dim3(256), 0, D.stream, d, W[0]);
    for (int r = 0; r < nsq; r++) {
        float *src = W[r & 1], *dst = W[(r + 1) & 1];
        hipLaunchKernelGGL(k_w8_rowscale, dim3(d / 64), dim3(256), 0, D.stream, d, src, rs);
        hipLaunchKernelGGL(k_w8_colscale, dim3(d), dim3(64), 0, D.stream, d, src, cs);
        hipLaunchKernelGGL(k_w8_split, dim3(d / 64, d / 64), dim3(256), 0, D.stream, d, src, rs, cs, Ah,
                           Al, Bh, Bl);
        hipLaunchKernelGGL(k_w8_sq, dim3(d / 64, d / 64), dim3(256), 0, D.stream, d, Ah, Al, Bh, Bl, rs,
                           cs, src, dst);
        hipLaunchKernelGGL(k_w8_fix, dim3(d / 64), dim3(256), 0, D.stream, d, dst);
    }
    return (int)check_launch();
}
