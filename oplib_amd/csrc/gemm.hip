// gemm.hip -- FP32 GEMM on the CDNA4 matrix cores (v_mfma_f32_16x16x4_f32).
//
// Drop-in entry points:
//   opcuda_sgemm    C = alpha*A*B + beta*C, column-major, 'N','N'
//                   (sgemm.cu:41-44, cublasSgemm)
//   opcuda_ssqmm    batched d x d products over device pointer tables, the
//                   repeated-squaring and propagation workhorse
//                   (ssqmm.cu:119-179; driven by CModel Device.cs:216-279)
//   opcuda_sgemm4   uneven batched GEMM over per-tile descriptors
//                   {i,bx,by,m,n,k,A,B,C}, 64x16 tiles (sgemm4.cu:84-154)
// MI355X additions:
//   ophip_sgemm_strided_batched
//
// Kernel design (gfx950, wave64):
//   * workgroup tile BM x BN (128x128 or 64x64), K-step 16, 4 waves in a
//     2x2 arrangement, each wave owning (BM/2)x(BN/2) as 16x16 MFMA tiles;
//   * A staged in LDS as As[k][i] (float4 writes along i, conflict-free
//     scalar reads with a row pitch of BM+4), B staged as Bs[j][k] with a
//     pitch of 20 floats so a lane reads four consecutive k as one
//     ds_read_b128;  the K order inside a 16-step is permuted (lane group g
//     of MFMA step s takes k = 4g+s) so that those four values feed four
//     consecutive MFMAs;
//   * two LDS stages: the next tile's global loads are in flight while the
//     current one is multiplied; one barrier per K-step;
//   * workgroup ids are remapped so that consecutive tiles land on the same
//     XCD (dispatch is round-robin over the 8 XCDs, each with its own L2).
#include "ophip_internal.h"

using namespace ophip;

typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 16;
constexpr int BPITCH = BK + 4;

struct GemmArgs {
    const char *base;  // arena base for table-driven batches
    const uint32_t *tA, *tB, *tC;  // per-batch handle tables (or null)
    const float *A, *B;
    float *C;
    long long sA, sB, sC;  // per-batch element strides when no tables
    int m, n, k, lda, ldb, ldc;
    float alpha, beta;
};

__device__ inline f4 mfma4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int BM, int BN, bool FAST>
__global__ __launch_bounds__(256) void k_sgemm(GemmArgs g) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int APITCH = BM + 4;
    constexpr int A_F4 = BM * BK / 4 / 256;  // float4 loads per thread for A
    constexpr int B_F4 = BN * BK / 4 / 256;
    static_assert(A_F4 >= 1 && B_F4 >= 1, "tile too small");

    __shared__ __attribute__((aligned(16))) float As[2][BK * APITCH];
    __shared__ __attribute__((aligned(16))) float Bs[2][BN * BPITCH];

    const int z = blockIdx.z;
    const float *A;
    const float *B;
    float *C;
    if (g.tA) {
        A = (const float *)(g.base + g.tA[z]);
        B = (const float *)(g.base + g.tB[z]);
        C = (float *)(g.base + g.tC[z]);
    } else {
        A = g.A + z * g.sA;
        B = g.B + z * g.sB;
        C = g.C + z * g.sC;
    }

    const int tiles_m = (g.m + BM - 1) / BM;
    const int tiles_n = (g.n + BN - 1) / BN;
    const int ntiles = tiles_m * tiles_n;
    int bid = blockIdx.x;
    if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    const int tm_i = bid % tiles_m, tn_i = bid / tiles_m;
    const int m0 = tm_i * BM, n0 = tn_i * BN;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int r = lane & 15, grp = lane >> 4;
    const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;

    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    f4 ra[A_F4], rb[B_F4];

    auto load_tile = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A_F4; q++) {
            int idx = t + 256 * q;
            int kk = idx / (BM / 4), i4 = idx % (BM / 4);
            int gi = m0 + 4 * i4, gk = k0 + kk;
            if (FAST) {
                ra[q] = *(const f4 *)(A + gi + (size_t)g.lda * gk);
            } else {
                f4 v = {0.f, 0.f, 0.f, 0.f};
                if (gk < g.k) {
                    const float *src = A + (size_t)g.lda * gk;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (gi + e < g.m) v[e] = src[gi + e];
                }
                ra[q] = v;
            }
        }
#pragma unroll
        for (int q = 0; q < B_F4; q++) {
            int idx = t + 256 * q;
            int j = idx >> 2, kq = idx & 3;
            int gj = n0 + j, gk = k0 + 4 * kq;
            if (FAST) {
                rb[q] = *(const f4 *)(B + gk + (size_t)g.ldb * gj);
            } else {
                f4 v = {0.f, 0.f, 0.f, 0.f};
                if (gj < g.n) {
                    const float *src = B + (size_t)g.ldb * gj;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (gk + e < g.k) v[e] = src[gk + e];
                }
                rb[q] = v;
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int q = 0; q < A_F4; q++) {
            int idx = t + 256 * q;
            int kk = idx / (BM / 4), i4 = idx % (BM / 4);
            *(f4 *)&As[buf][kk * APITCH + 4 * i4] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < B_F4; q++) {
            int idx = t + 256 * q;
            int j = idx >> 2, kq = idx & 3;
            *(f4 *)&Bs[buf][j * BPITCH + 4 * kq] = rb[q];
        }
    };

    const int nk = (g.k + BK - 1) / BK;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * BK);
        f4 bq[TN];
#pragma unroll
        for (int j = 0; j < TN; j++)
            bq[j] = *(const f4 *)&Bs[buf][(wn0 + 16 * j + r) * BPITCH + 4 * grp];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            float a[TM];
#pragma unroll
            for (int i = 0; i < TM; i++) a[i] = As[buf][(4 * grp + s) * APITCH + wm0 + 16 * i + r];
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int i = 0; i < TM; i++) acc[i][j] = mfma4(a[i], bq[j][s], acc[i][j]);
        }
        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }

    // epilogue: lane holds C[4*grp + v][r] of every 16x16 tile
    const float alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < TM; i++) {
        const int row = m0 + wm0 + 16 * i + 4 * grp;
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const int col = n0 + wn0 + 16 * j + r;
            float *cp = C + row + (size_t)g.ldc * col;
            if (FAST) {
                f4 v = acc[i][j] * alpha;
                if (beta != 0.0f) v += beta * *(const f4 *)cp;
                *(f4 *)cp = v;
            } else if (col < g.n) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (row + e < g.m) {
                        float v = alpha * acc[i][j][e];
                        if (beta != 0.0f) v += beta * cp[e];
                        cp[e] = v;
                    }
                }
            }
        }
    }
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// choose the tile and the fast path, then launch `batch` products
unsigned launch_sgemm(GemmArgs g, int batch, bool tables_aligned) {
    Device &D = dev();
    if (g.m <= 0 || g.n <= 0 || batch <= 0) return set_status(kOk);
    bool big = ((long)(g.m / 128) * (g.n / 128) * batch) >= 128;
    int BM = big ? 128 : 64;
    bool fast = (g.m % BM == 0) && (g.n % BM == 0) && (g.k % BK == 0) && (g.lda % 4 == 0) &&
                (g.ldb % 4 == 0) && (g.ldc % 4 == 0) && tables_aligned &&
                (g.tA || (aligned16(g.A) && aligned16(g.B) && aligned16(g.C) && g.sA % 4 == 0 &&
                          g.sB % 4 == 0 && g.sC % 4 == 0));
    dim3 grid(cdiv(g.m, BM) * cdiv(g.n, BM), 1, batch), block(256);
    if (g.k <= 0) {
        // C = beta*C (alpha*A*B is empty): run with k = 0 through the
        // generic path, which never reads A or B
        fast = false;
    }
    if (big) {
        if (fast) hipLaunchKernelGGL((k_sgemm<128, 128, true>), grid, block, 0, D.stream, g);
        else hipLaunchKernelGGL((k_sgemm<128, 128, false>), grid, block, 0, D.stream, g);
    } else {
        if (fast) hipLaunchKernelGGL((k_sgemm<64, 64, true>), grid, block, 0, D.stream, g);
        else hipLaunchKernelGGL((k_sgemm<64, 64, false>), grid, block, 0, D.stream, g);
    }
    return check_launch();
}

// ---------------------------------------------------------------------
// sgemm4: one 64x16 output tile per descriptor, one wave per workgroup
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sgemm4(const char *__restrict__ base,
                                               const uint32_t *__restrict__ argptr_bid) {
    const uint32_t *arg = (const uint32_t *)(base + argptr_bid[blockIdx.x]);
    const int bx = (int)arg[1], by = (int)arg[2];
    const int m = (int)arg[3], n = (int)arg[4], K = (int)arg[5];
    const float *A = (const float *)(base + arg[6]);
    const float *B = (const float *)(base + arg[7]);
    float *C = (float *)(base + arg[8]);
    const int lda = m, ldb = K, ldc = m;
    const int lane = threadIdx.x, r = lane & 15, grp = lane >> 4;
    const int row0 = bx * 64, col = by * 16 + r;
    f4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 4) {
        int kk = k0 + grp;
        float b = (kk < K && col < n) ? B[kk + (size_t)ldb * col] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int row = row0 + 16 * i + r;
            float a = (kk < K && row < m) ? A[row + (size_t)lda * kk] : 0.0f;
            acc[i] = mfma4(a, b, acc[i]);
        }
    }
    if (col >= n) return;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int row = row0 + 16 * i + 4 * grp + e;
            if (row < m) C[row + (size_t)ldc * col] = acc[i][e];
        }
}

}  // namespace

OPHIP_EXPORT void opcuda_sgemm(int m, int n, int k, float alpha, unsigned A, int lda, unsigned B,
                               int ldb, float beta, unsigned C, int ldc) {
    if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) ||
        ldc < (m > 1 ? m : 1)) {
        set_status(kErrBadArg);
        return;
    }
    if (m == 0 || n == 0) { set_status(kOk); return; }
    GemmArgs g{};
    g.A = P<const float>(A);
    g.B = P<const float>(B);
    g.C = P<float>(C);
    if (!g.C || (k > 0 && (!g.A || !g.B))) { set_status(kErrBadHandle); return; }
    g.m = m; g.n = n; g.k = k; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.alpha = alpha; g.beta = beta;
    launch_sgemm(g, 1, true);
}

// C_i = A_i * B_i for i < ni; A_i/B_i/C_i are device arrays of ni handles,
// every matrix d x d column-major (ssqmm.cu:174-179).  Handles must be
// 16-byte aligned for the fast path (arena handles are 256-aligned, and
// d%4 == 0 keeps every j*d*d offset aligned).
OPHIP_EXPORT void opcuda_ssqmm(int d, int ni, unsigned A_i, unsigned B_i, unsigned C_i) {
    Device &D = dev();
    if (d <= 0 || ni <= 0) { set_status(kOk); return; }
    GemmArgs g{};
    g.base = D.base;
    g.tA = P<const uint32_t>(A_i);
    g.tB = P<const uint32_t>(B_i);
    g.tC = P<const uint32_t>(C_i);
    if (!g.tA || !g.tB || !g.tC) { set_status(kErrBadHandle); return; }
    g.m = g.n = g.k = d;
    g.lda = g.ldb = g.ldc = d;
    g.alpha = 1.0f; g.beta = 0.0f;
    // table entries are only visible on the device; the caller's
    // offsets are 4-byte granular, so alignment is guaranteed for d%4==0
    // only when every handle is 16-byte aligned -- checked on the host side
    // by the caller contract (handles from opcuda_mem_alloc plus multiples of
    // d*d*4 bytes).  d%4 != 0 takes the generic path.
    launch_sgemm(g, ni, (d % 4) == 0);
}

OPHIP_EXPORT void opcuda_sgemm4(int nblocks, unsigned argptr_bid) {
    Device &D = dev();
    if (nblocks <= 0) { set_status(kOk); return; }
    const uint32_t *tab = P<const uint32_t>(argptr_bid);
    if (!tab) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_sgemm4, dim3(nblocks), dim3(64), 0, D.stream, (const char *)D.base, tab);
    check_launch();
}

OPHIP_EXPORT int ophip_sgemm_strided_batched(int m, int n, int k, float alpha, unsigned A, int lda,
                                             long long strideA, unsigned B, int ldb,
                                             long long strideB, float beta, unsigned C, int ldc,
                                             long long strideC, int batch) {
    if (m < 0 || n < 0 || k < 0 || batch < 0) return (int)set_status(kErrBadArg);
    GemmArgs g{};
    g.A = P<const float>(A);
    g.B = P<const float>(B);
    g.C = P<float>(C);
    if (!g.C || (k > 0 && (!g.A || !g.B))) return (int)set_status(kErrBadHandle);
    g.m = m; g.n = n; g.k = k; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.sA = strideA; g.sB = strideB; g.sC = strideC;
    g.alpha = alpha; g.beta = beta;
    return (int)launch_sgemm(g, batch, true);
}
