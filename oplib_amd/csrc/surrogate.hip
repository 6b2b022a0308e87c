// surrogate.hip -- the learned price-grid surrogate: an ensemble of E MLPs
// 5 -> 256 -> 256 -> 256 -> 210 (exact-erf GELU on the hidden layers,
// linear output), trained on prices produced by the operator solver.
//
// Arithmetic: BF16 operands on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation; FP32 master weights, biases, gradients and Adam moments.
// Every GEMM of the network (forward, backward-data, weight gradient) is
// one kernel, k_mlp_gemm, over operands stored with the reduction index
// contiguous ("k-major"): the producing epilogues write activations and
// gradients both row-major and transposed so that each consumer reads
// 16-byte fragments straight from memory.  Padded shapes: input 5 -> 32,
// output 210 -> 256; padding weights stay exactly 0 (their gradients are 0
// because the padded inputs and output gradients are 0).
//
// Determinism: no atomics.  Weight gradients are split over the batch into
// fixed chunks whose partials are added in chunk order; bias gradients are
// summed in a fixed lane/offset order.  The same inputs give the same bits.
//
// No counterpart in the reference (the learned surrogate).  The CPU
// oracle (or_mlp_forward / or_mlp_backward / or_adam_step) restates the
// same network in FP64 with W_l stored [out][in] row-major, the layout of
// the master weights here.
#include "ophip_internal.h"

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

using namespace ophip;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

constexpr int kIn = 5, kInP = 32, kHid = 256, kOut = 210, kOutP = 256;
constexpr int kInLo = kIn;  // first input column of the parameters' BF16 remainders
// padded master layout of one member (floats); weights W_l[o][i]
constexpr int oW1 = 0;
constexpr int ob1 = oW1 + kHid * kInP;
constexpr int oW2 = ob1 + kHid;
constexpr int ob2 = oW2 + kHid * kHid;
constexpr int oW3 = ob2 + kHid;
constexpr int ob3 = oW3 + kHid * kHid;
constexpr int oW4 = ob3 + kHid;
constexpr int ob4 = oW4 + kOutP * kHid;
constexpr int kPar = ob4 + kOutP;  // 205,824 (187,090 of them live)

__device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ inline f4 mfma_bf16(s8 a, s8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                   __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

enum { EPI_HID = 0, EPI_OUT = 1, EPI_BWD = 2, EPI_WGRAD = 3 };

// C[M x N] = A[M x K] * B^T with A stored [M][lda] and B stored [N][ldb]
// (both k-contiguous, BF16).  blockIdx.z = member * ks + chunk.
struct GemmOp {
    const uint16_t *A, *B;
    long long sA, sB;  // member strides (elements)
    int M, N, K, lda, ldb, kchunk, ks;
    const float *bias;  // EPI_HID / EPI_OUT: bias[n], member stride sbias
    long long sbias;
    float *Cf;  // fp32 out: Z (HID), Y (OUT), weight-gradient partials (WGRAD)
    long long sCf, sPart;
    int ldcf;
    uint16_t *Cb;  // bf16 row-major out [m][n]
    long long sCb;
    int ldcb;
    uint16_t *Ct;  // bf16 transposed out [n][m]
    long long sCt;
    int ldct;
    const float *Z;  // EPI_BWD: pre-activations [m][ldz]
    long long sZ;
    int ldz;
    int nvalid;  // OUT / WGRAD: columns >= nvalid are not real outputs
};

// 256 threads = 2 x 2 waves, workgroup tile 64 x 64, wave tile 32 x 32
// (2 x 2 MFMA tiles), one K step = 32.  Lane (r, g) holds rows r of each
// 16-row tile and k = 8g .. 8g+7 of the step; the accumulator lane holds
// C[4g + v][r].  Fragments come straight from global memory with the next
// step's loads issued before the current step's MFMAs.
template <int EPI>
__global__ __launch_bounds__(256) void k_mlp_gemm(GemmOp op) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int e = blockIdx.z / op.ks, q = blockIdx.z - e * op.ks;
    const int kb = q * op.kchunk, ke = min(op.K, kb + op.kchunk);

    const uint16_t *Ab = op.A + e * op.sA + (size_t)(m0 + wm + r) * op.lda + 8 * g;
    const uint16_t *Bb = op.B + e * op.sB + (size_t)(n0 + wn + r) * op.ldb + 8 * g;
    const size_t a16 = (size_t)16 * op.lda, b16 = (size_t)16 * op.ldb;

    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

    s8 fa0[2], fb0[2], fa1[2], fb1[2];
    auto load = [&](s8 *fa, s8 *fb, int k) {
        fa[0] = *(const s8 *)(Ab + k);
        fa[1] = *(const s8 *)(Ab + a16 + k);
        fb[0] = *(const s8 *)(Bb + k);
        fb[1] = *(const s8 *)(Bb + b16 + k);
    };
    auto mm = [&](const s8 *fa, const s8 *fb) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = mfma_bf16(fa[a], fb[b], acc[a][b]);
    };
    // K, kb and ke are multiples of 32; the last (clamped) prefetch reloads
    // a valid step and is never used
    load(fa0, fb0, kb);
    int k = kb;
    for (; k + 64 <= ke; k += 64) {
        load(fa1, fb1, k + 32);
        mm(fa0, fb0);
        load(fa0, fb0, min(k + 64, ke - 32));
        mm(fa1, fb1);
    }
    if (k < ke) mm(fa0, fb0);

    // epilogue
#pragma unroll
    for (int a = 0; a < 2; a++) {
        const int row = m0 + wm + 16 * a + 4 * g;  // rows row .. row+3
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int col = n0 + wn + 16 * b + r;
            const f4 c = acc[a][b];
            if (EPI == EPI_WGRAD) {
                if (col < op.nvalid) {
                    float *o = op.Cf + e * op.sCf + q * op.sPart;
#pragma unroll
                    for (int v = 0; v < 4; v++) o[(size_t)(row + v) * op.ldcf + col] = c[v];
                }
            } else if (EPI == EPI_OUT) {
                const float bb = op.bias[e * op.sbias + col];
                float *o = op.Cf + e * op.sCf;
#pragma unroll
                for (int v = 0; v < 4; v++)
                    o[(size_t)(row + v) * op.ldcf + col] = col < op.nvalid ? c[v] + bb : 0.0f;
            } else {
                uint16_t h[4];
                if (EPI == EPI_HID) {
                    const float bb = op.bias[e * op.sbias + col];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const float z = c[v] + bb;
                        if (op.Cf) op.Cf[e * op.sCf + (size_t)(row + v) * op.ldcf + col] = z;
                        h[v] = f2bf(gelu(z));
                    }
                } else {  // EPI_BWD: dZ = dH * gelu'(Z)
                    const float *Z = op.Z + e * op.sZ;
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        h[v] = f2bf(c[v] * gelu_grad(Z[(size_t)(row + v) * op.ldz + col]));
                }
                uint16_t *ob = op.Cb + e * op.sCb;
#pragma unroll
                for (int v = 0; v < 4; v++) ob[(size_t)(row + v) * op.ldcb + col] = h[v];
                if (op.Ct) {
                    uint2 pk;
                    pk.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
                    pk.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
                    *(uint2 *)(op.Ct + e * op.sCt + (size_t)col * op.ldct + row) = pk;
                }
            }
        }
    }
}

// normalized fp32 inputs -> bf16 X [b][32] (the rounded inputs, their
// rounded remainders in columns 5..9, zero padded) and, for training,
// X^T [64][B] (the rounded inputs; rows 5..63 zero).  idx (optional): member e's row
// b is source row idx[e*B + b]; without idx all members share row b.
__global__ void k_mlp_inputs(int B, int E, const float *__restrict__ src, const int *__restrict__ idx,
                             uint16_t *__restrict__ X, long long sX, uint16_t *__restrict__ XT,
                             long long sXT) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
    if (b >= B) return;
    const int s = idx ? idx[(size_t)e * B + b] : b;
    uint16_t v[kInP];
#pragma unroll
    for (int i = 0; i < kInP; i++) v[i] = 0;
#pragma unroll
    for (int i = 0; i < kIn; i++) {
        const float x = src[(size_t)s * kIn + i];
        v[i] = f2bf(x);
        v[kInLo + i] = f2bf(x - bf2f(v[i]));  // exact difference, rounded once
    }
    uint16_t *x = X + e * sX + (size_t)b * kInP;
#pragma unroll
    for (int i = 0; i < kInP; i += 8) {
        uint4 p;
        p.x = (uint32_t)v[i] | ((uint32_t)v[i + 1] << 16);
        p.y = (uint32_t)v[i + 2] | ((uint32_t)v[i + 3] << 16);
        p.z = (uint32_t)v[i + 4] | ((uint32_t)v[i + 5] << 16);
        p.w = (uint32_t)v[i + 6] | ((uint32_t)v[i + 7] << 16);
        *(uint4 *)(x + i) = p;
    }
    if (XT) {
        uint16_t *xt = XT + e * sXT;
        for (int i = 0; i < 64; i++) xt[(size_t)i * B + b] = i < kIn ? v[i] : 0;
    }
}

// MSE output gradient: dY = scale * (Y - T) over the 210 live columns (0 on
// the padding), or dY = given fp32 gradient G.  One thread per column,
// 64 rows per workgroup; writes dY [b][256] and dY^T [256][B] (bf16) and the
// per-(member, block, column) squared-error partial sums.
__global__ __launch_bounds__(256) void k_mlp_loss(int B, const float *__restrict__ Y, long long sY,
                                                  const float *__restrict__ T,
                                                  const int *__restrict__ idx,
                                                  const float *__restrict__ G, float scale,
                                                  uint16_t *__restrict__ dY, uint16_t *__restrict__ dYT,
                                                  long long sD, float *__restrict__ lossp) {
    const int c = threadIdx.x, e = blockIdx.y, b0 = blockIdx.x * 64;
    const float *y = Y + e * sY;
    uint16_t *dy = dY + e * sD, *dyt = dYT + e * sD;
    uint32_t pk[32];
    float se = 0.0f;
#pragma unroll
    for (int i = 0; i < 64; i++) {
        const int b = b0 + i;
        float gr = 0.0f;
        if (G) {
            if (c < kOut) gr = G[e * sY + (size_t)b * kOutP + c];
        } else if (c < kOut) {
            const int s = idx ? idx[(size_t)e * B + b] : b;
            const float diff = y[(size_t)b * kOutP + c] - T[(size_t)s * kOut + c];
            se += diff * diff;
            gr = scale * diff;
        }
        const uint16_t h = f2bf(gr);
        dy[(size_t)b * kOutP + c] = h;
        if (i & 1) pk[i >> 1] |= (uint32_t)h << 16;
        else pk[i >> 1] = h;
    }
    uint4 *o = (uint4 *)(dyt + (size_t)c * B + b0);
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = uint4{pk[4 * j], pk[4 * j + 1], pk[4 * j + 2], pk[4 * j + 3]};
    lossp[((size_t)e * gridDim.x + blockIdx.x) * kOutP + c] = se;
}

// bias gradient db[o] = sum_b dZ^T[o][b]: one wave per (o, member), lane
// sums b = lane, lane+64, ... then a fixed-order butterfly
__global__ __launch_bounds__(64) void k_mlp_bgrad(int B, const uint16_t *__restrict__ dZT, long long sD,
                                                  float *__restrict__ grad, int boff) {
    const int o = blockIdx.x, e = blockIdx.y, lane = threadIdx.x;
    const uint16_t *p = dZT + e * sD + (size_t)o * B;
    float s = 0.0f;
    for (int b = lane; b < B; b += 64) s += bf2f(p[b]);
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w);
    if (lane == 0) grad[(size_t)e * kPar + boff + o] = s;
}

// weight gradient: add the ks batch-chunk partials in chunk order
__global__ void k_mlp_wsum(int n, int ks, const float *__restrict__ part, long long sPart,
                           long long sMem, float *__restrict__ grad, int woff) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
    if (j >= n) return;
    const float *p = part + e * sMem + j;
    float s = p[0];
    for (int q = 1; q < ks; q++) s += p[q * sPart];
    grad[(size_t)e * kPar + woff + j] = s;
}

// which layer a master index belongs to: weights -> (layer, o, i) and the
// bf16 copies Wf[o][i] (forward operand) and Wb[i][o] (backward operand)
// The first layer's input carries each parameter as a BF16 pair: column
// i < 5 its rounded value, column 5 + i the rounded remainder (16
// significant bits together; one BF16 input alone quantizes the box to
// 2^-9 of its width).  Both columns take the same weight: the forward
// copy Wf mirrors W1[o][i] into column 5 + i, whose master weight stays 0
// (its gradient is 0: the transposed input keeps rows 5.. zero).

__device__ inline void bf16_copies(int j, float w, uint16_t *Wf, uint16_t *Wb) {
    int base, in, out;
    if (j < ob1) { base = oW1; in = kInP; out = kHid; }
    else if (j >= oW2 && j < ob2) { base = oW2; in = kHid; out = kHid; }
    else if (j >= oW3 && j < ob3) { base = oW3; in = kHid; out = kHid; }
    else if (j >= oW4 && j < ob4) { base = oW4; in = kHid; out = kOutP; }
    else return;  // a bias
    const int l = j - base, o = l / in, i = l - o * in;
    const uint16_t h = f2bf(w);
    if (base == oW1 && i >= kInLo && i < kInLo + kIn) return;  // written by column i - 5
    Wf[j] = h;
    if (base == oW1 && i < kIn) Wf[j + kInLo] = h;
    Wb[base + (size_t)i * out + o] = h;
}

// Adam (Kingma & Ba) with bias correction, c1 = 1/(1-beta1^t), c2 =
// 1/(1-beta2^t); refreshes the bf16 weight copies
__global__ void k_mlp_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ m, float *__restrict__ v, uint16_t *__restrict__ Wf,
                           uint16_t *__restrict__ Wb, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
    if (j >= n) return;
    const size_t k = (size_t)e * kPar + j;
    const float gg = g[k];
    const float mm = b1 * m[k] + (1.0f - b1) * gg;
    const float vv = b2 * v[k] + (1.0f - b2) * gg * gg;
    m[k] = mm;
    v[k] = vv;
    const float w = p[k] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    p[k] = w;
    bf16_copies(j, w, Wf + (size_t)e * kPar, Wb + (size_t)e * kPar);
}

__global__ void k_mlp_refresh(int n, const float *__restrict__ p, uint16_t *__restrict__ Wf,
                              uint16_t *__restrict__ Wb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
    if (j >= n) return;
    bf16_copies(j, p[(size_t)e * kPar + j], Wf + (size_t)e * kPar, Wb + (size_t)e * kPar);
}

// ---------------------------------------------------------------------
// surrogate head: parameters -> network inputs; ensemble -> prices
// ---------------------------------------------------------------------
struct SurrGrid {
    float lo[kIn], hi[kIn];  // parameter box, mapped to [-1, 1]
    float spot;
    int nmat, nstrikes;      // nmat * nstrikes == 210
    float dT;                // maturities dT, 2 dT, ..., nmat dT
    float strikes[64];
};

// network inputs: each parameter mapped onto [-1, 1] over its box, sigma0
// on a log scale (prices change fastest at low vol)
__global__ void k_surr_norm(int B, SurrGrid gr, const float *__restrict__ prm, float *__restrict__ xn) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * kIn) return;
    const int i = j % kIn;
    if (i == 0)
        xn[j] = 2.0f * logf(prm[j] / gr.lo[0]) / logf(gr.hi[0] / gr.lo[0]) - 1.0f;
    else
        xn[j] = 2.0f * (prm[j] - gr.lo[i]) / (gr.hi[i] - gr.lo[i]) - 1.0f;
}

// local vol sigma0 (x/100)^(beta-1) (1 + a l + b l^2), l = ln(x/100),
// floored at 1e-4 (the lattice's CEV family)
__device__ inline float surr_vol(const float *p, float x) {
    const float l = logf(x / 100.0f);
    const float s = p[0] * expf((p[1] - 1.0f) * l) * (1.0f + p[3] * l + p[4] * l * l);
    return fmaxf(s, 1e-4f);
}

// control variate: Black-Scholes call at the short-maturity implied vol of
// Berestycki, Busca & Florent (2002), the harmonic mean of the local vol
// along the log-price path from S to K:
//   1/sigma = int_0^1 ds / sigma_loc(S (K/S)^s),
// by 8-point Gauss-Legendre on [0, 1]
__device__ inline float surr_base(const float *p, float S, float K, float T) {
    const float gx[4] = {0.1834346424956498f, 0.5255324099163290f, 0.7966664774136267f,
                         0.9602898564975363f};
    const float gw[4] = {0.3626837833783620f, 0.3137066458778873f, 0.2223810344533745f,
                         0.1012285362903763f};
    const float lk = logf(K / S);
    float inv = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float s0 = 0.5f * (1.0f - gx[i]), s1 = 0.5f * (1.0f + gx[i]);
        inv += 0.5f * gw[i] * (1.0f / surr_vol(p, S * expf(s0 * lk)) + 1.0f / surr_vol(p, S * expf(s1 * lk)));
    }
    const float sig = 1.0f / inv;
    const float sq = sig * sqrtf(T);
    const float d1 = (logf(S / K) + (p[2] + 0.5f * sig * sig) * T) / sq, d2 = d1 - sq;
    const float n1 = 0.5f * erfcf(-d1 * 0.70710678118654752f);
    const float n2 = 0.5f * erfcf(-d2 * 0.70710678118654752f);
    return S * n1 - K * expf(-p[2] * T) * n2;
}

// base[b][k*nstrikes + c] for maturity (k+1) dT and strike c
__global__ void k_surr_base(int B, SurrGrid gr, const float *__restrict__ prm, float *__restrict__ base) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * kOut) return;
    const int b = j / kOut, o = j - b * kOut;
    const int k = o / gr.nstrikes, c = o - k * gr.nstrikes;
    float p[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) p[i] = prm[(size_t)b * kIn + i];
    base[j] = surr_base(p, gr.spot, gr.strikes[c], (k + 1) * gr.dT);
}

// one wave per draw: member outputs y_e (standardized relative residuals)
// -> price_o = base_o + w_o (mu_o + sd_o mean_e(y_e)) with the residual
// scale w_o = sqrt(base_o^2 + (1e-3 max|base|)^2) (a smooth stand-in for
// the error metric's denominator max(|p|, 1e-3 max|p|), taken on the
// control variate: within a factor sqrt(2) of it, without its kink, which
// the network cannot follow); spread = max_o w_o sd_o std_e(y_e) /
// max(|price_o|, 1e-3 max|price|); flag = spread > tol
__global__ __launch_bounds__(64) void k_surr_head(int B, int E, SurrGrid gr, const float *__restrict__ prm,
                                                  const float *__restrict__ Y, long long sY,
                                                  const float *__restrict__ mu, const float *__restrict__ sd,
                                                  float tol, float *__restrict__ prices,
                                                  float *__restrict__ spread, int *__restrict__ flag) {
    const int b = blockIdx.x, lane = threadIdx.x;
    float p[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) p[i] = prm[(size_t)b * kIn + i];
    float bs[4], bmax = 0.0f;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int o = lane + 64 * u;
        bs[u] = 0.0f;
        if (o < kOut) {
            const int k = o / gr.nstrikes, c = o - k * gr.nstrikes;
            bs[u] = surr_base(p, gr.spot, gr.strikes[c], (k + 1) * gr.dT);
            bmax = fmaxf(bmax, fabsf(bs[u]));
        }
    }
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) bmax = fmaxf(bmax, __shfl_xor(bmax, w));
    float pr[4], sp[4], pmax = 0.0f;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int o = lane + 64 * u;
        pr[u] = 0.0f;
        sp[u] = 0.0f;
        if (o < kOut) {
            float s = 0.0f, s2 = 0.0f;
            for (int e = 0; e < E; e++) {
                const float y = Y[e * sY + (size_t)b * kOutP + o];
                s += y;
                s2 += y * y;
            }
            const float mean = s / E;
            const float var = fmaxf(s2 / E - mean * mean, 0.0f);
            const float fl = 1e-3f * bmax;
            const float w = sqrtf(bs[u] * bs[u] + fl * fl);
            pr[u] = bs[u] + w * (mu[o] + sd[o] * mean);
            sp[u] = w * sd[o] * sqrtf(var);
            pmax = fmaxf(pmax, fabsf(pr[u]));
            prices[(size_t)b * kOut + o] = pr[u];
        }
    }
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) pmax = fmaxf(pmax, __shfl_xor(pmax, w));
    const float floor_ = 1e-3f * pmax;
    float rel = 0.0f;
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (lane + 64 * u < kOut) rel = fmaxf(rel, sp[u] / fmaxf(fabsf(pr[u]), floor_));
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) rel = fmaxf(rel, __shfl_xor(rel, w));
    if (lane == 0) {
        spread[b] = rel;
        flag[b] = rel > tol ? 1 : 0;
    }
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
struct Buf {
    unsigned h = 0;
    size_t bytes = 0;
};

}  // namespace

extern "C" unsigned opcuda_mem_alloc(unsigned sz);
extern "C" void opcuda_mem_free_device(unsigned p);

struct MlpNet {
    int E = 0, cap = 0, t = 0;      // members, batch capacity (multiple of 64), Adam step
    int ks = 0;                      // weight-gradient batch chunks at capacity
    std::vector<unsigned> bufs;      // every allocation, freed together
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // [E][kPar]
    uint16_t *Wf = nullptr, *Wb = nullptr;                          // [E][kPar]
    uint16_t *X = nullptr, *XT = nullptr;       // [E][cap][32], [E][64][cap]
    float *Z[3] = {};                           // [E][cap][256]
    uint16_t *H[3] = {}, *HT[3] = {};           // [E][cap][256], [E][256][cap]
    float *Y = nullptr;                         // [E][cap][256]
    uint16_t *D[2] = {}, *DT[2] = {};           // gradients w.r.t. pre-activations
    float *part = nullptr;                      // [E][ks][256*256]
    float *lossp = nullptr;                     // [E][cap/64][256]
    float *xn = nullptr;                        // [cap][5] normalized inputs
    float *mu = nullptr, *sd = nullptr;         // [256] output standardization
};

namespace {

constexpr int kChunk = 512;  // batch rows per weight-gradient partial

template <class T> bool grab(MlpNet *n, T *&out, size_t count) {
    size_t bytes = count * sizeof(T);
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes > 0xffffff00u) return false;
    unsigned h = opcuda_mem_alloc((unsigned)bytes);
    if (!h) return false;
    n->bufs.push_back(h);
    out = P<T>(h);
    return hipMemsetAsync(out, 0, bytes, dev().stream) == hipSuccess;
}

GemmOp gemm_base(int E, int M, int N, int K) {
    GemmOp op;
    memset(&op, 0, sizeof(op));
    op.M = M;
    op.N = N;
    op.K = K;
    op.kchunk = K;
    op.ks = 1;
    op.nvalid = N;
    (void)E;
    return op;
}

template <int EPI> void launch_gemm(const GemmOp &op, int E) {
    dim3 grid(op.N / 64, op.M / 64, E * op.ks);
    hipLaunchKernelGGL(k_mlp_gemm<EPI>, grid, dim3(256), 0, dev().stream, op);
}

// forward pass of every member over B rows of X (member stride sX); when
// train, keeps Z, H and H^T for the backward pass
void forward(MlpNet *n, int B, long long sX, bool train) {
    const int E = n->E;
    const long long sAct = (long long)n->cap * kHid;  // member stride of activations
    const uint16_t *in = n->X;
    long long sIn = sX;
    int K = kInP;
    int woff = oW1, boff = ob1;
    for (int l = 0; l < 3; l++) {
        GemmOp op = gemm_base(E, B, kHid, K);
        op.A = in;
        op.sA = sIn;
        op.lda = K;
        op.B = n->Wf + woff;
        op.sB = kPar;
        op.ldb = K;
        op.bias = n->p + boff;
        op.sbias = kPar;
        op.Cf = train ? n->Z[l] : nullptr;
        op.sCf = sAct;
        op.ldcf = kHid;
        // inference ping-pongs between two activation buffers
        uint16_t *Hout = train ? n->H[l] : n->H[l & 1];
        op.Cb = Hout;
        op.sCb = sAct;
        op.ldcb = kHid;
        op.Ct = train ? n->HT[l] : nullptr;
        op.sCt = sAct;
        op.ldct = B;
        launch_gemm<EPI_HID>(op, E);
        in = Hout;
        sIn = sAct;
        K = kHid;
        woff = l == 0 ? oW2 : (l == 1 ? oW3 : oW4);
        boff = l == 0 ? ob2 : (l == 1 ? ob3 : ob4);
    }
    GemmOp op = gemm_base(E, B, kOutP, kHid);
    op.A = in;
    op.sA = sIn;
    op.lda = kHid;
    op.B = n->Wf + oW4;
    op.sB = kPar;
    op.ldb = kHid;
    op.bias = n->p + ob4;
    op.sbias = kPar;
    op.Cf = n->Y;
    op.sCf = (long long)n->cap * kOutP;
    op.ldcf = kOutP;
    op.nvalid = kOut;
    launch_gemm<EPI_OUT>(op, E);
}

// weight gradient of one layer: dW[o][i] = sum_b dZ^T[o][b] * A^T[i][b]
void wgrad(MlpNet *n, int B, const uint16_t *DT, const uint16_t *AT, long long sAT, int nin,
           int nout, int woff) {
    const int E = n->E;
    const int ks = cdiv(B, kChunk);  // the last chunk may be shorter (a multiple of 64)
    GemmOp op = gemm_base(E, nout, nin < 64 ? 64 : nin, B);
    op.A = DT;
    op.sA = (long long)n->cap * kHid;
    op.lda = B;
    op.B = AT;
    op.sB = sAT;
    op.ldb = B;
    op.ks = ks;
    op.kchunk = kChunk;
    op.Cf = n->part;
    op.sCf = (long long)n->ks * kHid * kHid;
    op.sPart = (long long)nout * nin;
    op.ldcf = nin;
    op.nvalid = nin;
    launch_gemm<EPI_WGRAD>(op, E);
    const int cnt = nout * nin;
    hipLaunchKernelGGL(k_mlp_wsum, dim3(cdiv(cnt, 256), E), dim3(256), 0, dev().stream, cnt, ks,
                       n->part, op.sPart, op.sCf, n->g, woff);
}

// backward-data through layer l+1 into the GELU of layer l:
// dZ_l[b][i] = (sum_o dZ_{l+1}[b][o] W_{l+1}[o][i]) * gelu'(Z_l[b][i])
void bwd_data(MlpNet *n, int B, const uint16_t *Dn, int nout, int woff, int l, uint16_t *Dl,
              uint16_t *DTl) {
    const int E = n->E;
    const long long sAct = (long long)n->cap * kHid;
    GemmOp op = gemm_base(E, B, kHid, nout);
    op.A = Dn;
    op.sA = sAct;
    op.lda = nout;
    op.B = n->Wb + woff;  // [i][o]
    op.sB = kPar;
    op.ldb = nout;
    op.Z = n->Z[l];
    op.sZ = sAct;
    op.ldz = kHid;
    op.Cb = Dl;
    op.sCb = sAct;
    op.ldcb = kHid;
    op.Ct = DTl;
    op.sCt = sAct;
    op.ldct = B;
    launch_gemm<EPI_BWD>(op, E);
}

bool batch_ok(MlpNet *n, int B) { return n && B > 0 && B <= n->cap && (B % 64) == 0; }

}  // namespace

extern "C" {

/* ensemble of `members` networks for batches of up to batch_cap rows
 * (rounded up to a multiple of 64); weights start at zero */
OPHIP_EXPORT unsigned long ophip_mlp_create(int members, int batch_cap) {
    if (members <= 0 || members > 16 || batch_cap <= 0) {
        set_status(kErrBadArg);
        return 0;
    }
    dev();
    MlpNet *n = new MlpNet();
    n->E = members;
    n->cap = (batch_cap + 63) / 64 * 64;
    n->ks = cdiv(n->cap, kChunk);
    const size_t E = members, cap = n->cap, P = (size_t)E * kPar, act = E * cap * kHid;
    bool ok = grab(n, n->p, P) && grab(n, n->g, P) && grab(n, n->m, P) && grab(n, n->v, P) &&
              grab(n, n->Wf, P) && grab(n, n->Wb, P) && grab(n, n->X, E * cap * kInP) &&
              grab(n, n->XT, E * 64 * cap) && grab(n, n->Y, E * cap * kOutP) &&
              grab(n, n->part, E * n->ks * kHid * kHid) && grab(n, n->lossp, E * (cap / 64) * kOutP) &&
              grab(n, n->xn, cap * kIn) && grab(n, n->mu, (size_t)kOutP) && grab(n, n->sd, (size_t)kOutP);
    for (int l = 0; l < 3 && ok; l++)
        ok = grab(n, n->Z[l], act) && grab(n, n->H[l], act) && grab(n, n->HT[l], act);
    for (int l = 0; l < 2 && ok; l++) ok = grab(n, n->D[l], act) && grab(n, n->DT[l], act);
    if (!ok || hipStreamSynchronize(dev().stream) != hipSuccess) {
        for (unsigned h : n->bufs) opcuda_mem_free_device(h);
        delete n;
        set_status(kErrOutOfArena);
        return 0;
    }
    set_status(kOk);
    return (unsigned long)(uintptr_t)n;
}

OPHIP_EXPORT void ophip_mlp_free(unsigned long net) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!n) return;
    (void)hipStreamSynchronize(dev().stream);
    for (unsigned h : n->bufs) opcuda_mem_free_device(h);
    delete n;
}

OPHIP_EXPORT int ophip_mlp_param_count(void) { return kPar; }

/* copy one member's array between host and device (to_device 1: upload).
 * which: 0 master weights, 1 gradient, 2 Adam first moment, 3 second
 * moment; 4 output mean, 5 output scale (256 floats, member ignored).
 * Layout of 0..3 (kPar floats): W1[256][32] b1[256] W2[256][256] b2[256]
 * W3[256][256] b3[256] W4[256][256] b4[256], W_l[o][i]; input columns
 * 5..31 of W1 and output rows 210..255 of W4 / b4 must stay 0.  An upload
 * of master weights refreshes the bf16 copies and resets the Adam step. */
OPHIP_EXPORT int ophip_mlp_params(unsigned long net, int member, int which, float *host,
                                  int to_device) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!n || !host || which < 0 || which > 5 || ((which < 4) && (member < 0 || member >= n->E)))
        return (int)set_status(kErrBadArg);
    float *arr[6] = {n->p, n->g, n->m, n->v, n->mu, n->sd};
    const size_t cnt = which < 4 ? (size_t)kPar : (size_t)kOutP;
    float *d = arr[which] + (which < 4 ? (size_t)member * kPar : 0);
    hipError_t e;
    if (to_device) {
        e = hipMemcpyAsync(d, host, cnt * 4, hipMemcpyHostToDevice, dev().stream);
        if (e == hipSuccess && which == 0) {
            hipLaunchKernelGGL(k_mlp_refresh, dim3(cdiv(kPar, 256), n->E), dim3(256), 0, dev().stream,
                               kPar, n->p, n->Wf, n->Wb);
            n->t = 0;
        }
    } else {
        e = hipMemcpyAsync(host, d, cnt * 4, hipMemcpyDeviceToHost, dev().stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(dev().stream);
    return (int)hip_status(e);
}

/* inference: Y[e][b][0..255] (ld 256, columns 210..255 zero) for every
 * member e; X: B x 5 normalized fp32 inputs.  B % 64 == 0, B <= capacity.
 * The output handle holds members * B * 256 floats. */
OPHIP_EXPORT int ophip_mlp_forward(unsigned long net, int B, unsigned X, unsigned Yout) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!batch_ok(n, B)) return (int)set_status(kErrBadArg);
    const float *x = P<const float>(X);
    float *y = P<float>(Yout);
    if (!x || !handle_range_ok(X, (size_t)B * kIn * 4) ||
        !handle_range_ok(Yout, (size_t)n->E * B * kOutP * 4) || !y)
        return (int)set_status(kErrBadHandle);
    hipLaunchKernelGGL(k_mlp_inputs, dim3(cdiv(B, 256), 1), dim3(256), 0, dev().stream, B, 1, x,
                       (const int *)nullptr, n->X, 0LL, (uint16_t *)nullptr, 0LL);
    forward(n, B, 0, false);
    hipError_t e = hipMemcpy2DAsync(y, (size_t)B * kOutP * 4, n->Y, (size_t)n->cap * kOutP * 4,
                                    (size_t)B * kOutP * 4, n->E, hipMemcpyDeviceToDevice, dev().stream);
    if (e != hipSuccess) return (int)hip_status(e);
    return (int)check_launch();
}

/* gradients of every member into the gradient array (which = 1):
 * member e's rows are X[idx[e*B + b]] (idx 0: rows 0..B-1 for all).
 * G == 0: loss = sum over the 210 outputs of (y - T[row])^2 / (B * 210)
 * (T: rows x 210 fp32 targets).  G != 0: gradient of sum(y .* G) with G
 * given as members x B x 256 fp32 (the oracle's convention). */
OPHIP_EXPORT int ophip_mlp_backward(unsigned long net, int B, unsigned idx, unsigned X, unsigned T,
                                    unsigned G) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!batch_ok(n, B)) return (int)set_status(kErrBadArg);
    const int E = n->E;
    const int *ix = idx ? P<const int>(idx) : nullptr;
    const float *x = P<const float>(X), *tg = T ? P<const float>(T) : nullptr;
    const float *gg = G ? P<const float>(G) : nullptr;
    if ((idx && !ix) || !x || (!tg && !gg) || (G && !handle_range_ok(G, (size_t)E * B * kOutP * 4)) ||
        (idx && !handle_range_ok(idx, (size_t)E * B * 4)))
        return (int)set_status(kErrBadHandle);
    const long long sX = (long long)n->cap * kInP, sXT = (long long)64 * n->cap;
    const long long sAct = (long long)n->cap * kHid;
    hipStream_t s = dev().stream;
    hipLaunchKernelGGL(k_mlp_inputs, dim3(cdiv(B, 256), E), dim3(256), 0, s, B, E, x, ix, n->X, sX,
                       n->XT, sXT);
    // X^T rows are B apart for this batch size: the gemm reads them with ldb = B
    forward(n, B, sX, true);
    // the output gradient (into D[0] / DT[0]); G has member stride B*256
    hipLaunchKernelGGL(k_mlp_loss, dim3(B / 64, E), dim3(256), 0, s, B, n->Y,
                       gg ? (long long)B * kOutP : (long long)n->cap * kOutP, tg, ix, gg,
                       2.0f / ((float)B * kOut), n->D[0], n->DT[0], sAct, n->lossp);
    // layer 4 (output)
    wgrad(n, B, n->DT[0], n->HT[2], sAct, kHid, kOutP, oW4);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(kOutP, E), dim3(64), 0, s, B, n->DT[0], sAct, n->g, ob4);
    // through W4 into layer 3, then layers 2 and 1
    bwd_data(n, B, n->D[0], kOutP, oW4, 2, n->D[1], n->DT[1]);
    wgrad(n, B, n->DT[1], n->HT[1], sAct, kHid, kHid, oW3);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(kHid, E), dim3(64), 0, s, B, n->DT[1], sAct, n->g, ob3);
    bwd_data(n, B, n->D[1], kHid, oW3, 1, n->D[0], n->DT[0]);
    wgrad(n, B, n->DT[0], n->HT[0], sAct, kHid, kHid, oW2);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(kHid, E), dim3(64), 0, s, B, n->DT[0], sAct, n->g, ob2);
    bwd_data(n, B, n->D[0], kHid, oW2, 0, n->D[1], n->DT[1]);
    wgrad(n, B, n->DT[1], n->XT, sXT, kInP, kHid, oW1);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(kHid, E), dim3(64), 0, s, B, n->DT[1], sAct, n->g, ob1);
    return (int)check_launch();
}

/* one Adam step on every member from the gradient array */
OPHIP_EXPORT int ophip_mlp_adam(unsigned long net, float lr, float beta1, float beta2, float eps) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!n || !(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f))
        return (int)set_status(kErrBadArg);
    n->t++;
    const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, n->t)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, n->t)));
    hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(kPar, 256), n->E), dim3(256), 0, dev().stream, kPar,
                       n->p, n->g, n->m, n->v, n->Wf, n->Wb, lr, beta1, beta2, eps, c1, c2);
    return (int)check_launch();
}

/* summed squared error of the last MSE backward call, per member
 * (host array of members doubles); sums partials in a fixed order */
OPHIP_EXPORT int ophip_mlp_loss(unsigned long net, int B, double *out) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!batch_ok(n, B) || !out) return (int)set_status(kErrBadArg);
    const int nb = B / 64;
    std::vector<float> h((size_t)n->E * nb * kOutP);
    hipError_t e = hipMemcpyAsync(h.data(), n->lossp, h.size() * 4, hipMemcpyDeviceToHost, dev().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev().stream);
    if (e != hipSuccess) return (int)hip_status(e);
    for (int m = 0; m < n->E; m++) {
        double s = 0.0;
        for (size_t j = 0; j < (size_t)nb * kOutP; j++) s += h[(size_t)m * nb * kOutP + j];
        out[m] = s;
    }
    return 0;
}

/* surrogate evaluation of B parameter draws prm (B x 5 fp32: sigma0, beta,
 * r, a, b): ensemble forward on the box-normalized inputs, then per draw
 * the prices (B x 210: maturity-major, (k+1)*dT x strikes) = Black-Scholes
 * control variate + mean + scale * ensemble mean, the relative ensemble
 * spread, and flag = spread > tol (the draws to send to the solver).
 * lo/hi: the parameter box (5 each); strikes: nmat * nstrikes == 210. */
OPHIP_EXPORT int ophip_surr_eval(unsigned long net, int B, unsigned prm, const float *lo,
                                 const float *hi, float spot, float dT, int nmat,
                                 const float *strikes, int nstrikes, float tol, unsigned prices,
                                 unsigned spread, unsigned flags) {
    MlpNet *n = (MlpNet *)(uintptr_t)net;
    if (!batch_ok(n, B) || !lo || !hi || !strikes || nmat * nstrikes != kOut || nstrikes > 64)
        return (int)set_status(kErrBadArg);
    const float *pp = P<const float>(prm);
    float *pr = P<float>(prices), *sp = P<float>(spread);
    int *fl = P<int>(flags);
    if (!pp || !pr || !sp || !fl || !handle_range_ok(prm, (size_t)B * kIn * 4) ||
        !handle_range_ok(prices, (size_t)B * kOut * 4) || !handle_range_ok(spread, (size_t)B * 4) ||
        !handle_range_ok(flags, (size_t)B * 4))
        return (int)set_status(kErrBadHandle);
    SurrGrid gr;
    memset(&gr, 0, sizeof(gr));
    for (int i = 0; i < kIn; i++) {
        gr.lo[i] = lo[i];
        gr.hi[i] = hi[i];
    }
    gr.spot = spot;
    gr.dT = dT;
    gr.nmat = nmat;
    gr.nstrikes = nstrikes;
    for (int c = 0; c < nstrikes; c++) gr.strikes[c] = strikes[c];
    hipStream_t s = dev().stream;
    hipLaunchKernelGGL(k_surr_norm, dim3(cdiv(B * kIn, 256)), dim3(256), 0, s, B, gr, pp, n->xn);
    hipLaunchKernelGGL(k_mlp_inputs, dim3(cdiv(B, 256), 1), dim3(256), 0, s, B, 1, n->xn,
                       (const int *)nullptr, n->X, 0LL, (uint16_t *)nullptr, 0LL);
    forward(n, B, 0, false);
    hipLaunchKernelGGL(k_surr_head, dim3(B), dim3(64), 0, s, B, n->E, gr, pp, n->Y,
                       (long long)n->cap * kOutP, n->mu, n->sd, tol, pr, sp, fl);
    return (int)check_launch();
}

/* the control variate alone: base[b][k*nstrikes + c] for the draws prm */
OPHIP_EXPORT int ophip_surr_base(int B, unsigned prm, float spot, float dT, int nmat,
                                 const float *strikes, int nstrikes, unsigned base) {
    if (B <= 0 || !strikes || nmat * nstrikes != kOut || nstrikes > 64) return (int)set_status(kErrBadArg);
    const float *pp = P<const float>(prm);
    float *bb = P<float>(base);
    if (!pp || !bb || !handle_range_ok(prm, (size_t)B * kIn * 4) ||
        !handle_range_ok(base, (size_t)B * kOut * 4))
        return (int)set_status(kErrBadHandle);
    SurrGrid gr;
    memset(&gr, 0, sizeof(gr));
    gr.spot = spot;
    gr.dT = dT;
    gr.nmat = nmat;
    gr.nstrikes = nstrikes;
    for (int c = 0; c < nstrikes; c++) gr.strikes[c] = strikes[c];
    hipLaunchKernelGGL(k_surr_base, dim3(cdiv(B * kOut, 256)), dim3(256), 0, dev().stream, B, gr, pp, bb);
    return (int)check_launch();
}

}  // extern "C"
