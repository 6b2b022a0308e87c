// gemv.hip -- matrix-vector products of the pricing path.
//
// Drop-in entry points:
//   opcuda_sgemv / opcuda_sgemvT   BLAS sgemv 'N' / 'T' (sgemv.cu:44-62)
//   opcuda_sgemv3                  gather/scatter GEMV over 16-column groups,
//                                  Z[col1[c]+i] = sum_k A[i,k] Z[col0[c]+k]
//                                  (sgemv3.cu:92-176)
//   opcuda_sgemv4                  the same over per-tile descriptors
//                                  {i,bx,by,d,nz,ncol,A,Z,col0,col1}
//                                  (sgemv4.cu:90-180)
// MI355X addition:
//   ophip_sgemm_skinny   Y = alpha*A*X for n <= 64 right-hand columns,
//                        X and Y row-major (X[k*ldx + c]); the backward-
//                        induction step V_t = df*U*V_{t+1}.  One pass over
//                        A from HBM on the matrix cores, split-K partials
//                        combined deterministically.
//
// All reductions have a fixed order: results are bitwise reproducible run
// to run.
#include "ophip_internal.h"

using namespace ophip;

typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

__device__ inline f4 mfma4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------
// sgemv 'N': partial sums over K chunks, then a fixed-order combine
// ---------------------------------------------------------------------
// grid (ceil(m/256), S); 64 threads, 4 rows each
__global__ __launch_bounds__(64) void k_gemvN_part(int m, int n, const float *__restrict__ A,
                                                   int lda, const float *__restrict__ x,
                                                   int incx, int kc, float *__restrict__ part,
                                                   int vec) {
    const int row = blockIdx.x * 256 + threadIdx.x * 4;
    const int k0 = blockIdx.y * kc, k1 = min(n, k0 + kc);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    if (vec && row + 3 < m) {
        for (int k = k0; k < k1; k++) {
            float xv = x[(long)k * incx];
            f4 a = *(const f4 *)(A + row + (size_t)lda * k);
            acc += a * xv;
        }
    } else {
        for (int k = k0; k < k1; k++) {
            float xv = x[(long)k * incx];
            for (int e = 0; e < 4; e++)
                if (row + e < m) acc[e] += A[row + e + (size_t)lda * k] * xv;
        }
    }
    for (int e = 0; e < 4; e++)
        if (row + e < m) part[(size_t)blockIdx.y * m + row + e] = acc[e];
}

__global__ void k_gemv_combine(int m, int S, const float *__restrict__ part, float alpha,
                               float beta, float *__restrict__ y, int incy) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float s = 0.0f;
    for (int q = 0; q < S; q++) s += part[(size_t)q * m + i];
    float v = alpha * s;
    if (beta != 0.0f) v += beta * y[(long)i * incy];
    y[(long)i * incy] = v;
}

// sgemv 'T': one wave per output column
__global__ __launch_bounds__(256) void k_gemvT(int m, int n, const float *__restrict__ A, int lda,
                                               const float *__restrict__ x, int incx, float alpha,
                                               float beta, float *__restrict__ y, int incy) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (col >= n) return;
    const float *a = A + (size_t)lda * col;
    float s = 0.0f;
    for (int i = lane; i < m; i += 64) s += a[i] * x[(long)i * incx];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        float v = alpha * s;
        if (beta != 0.0f) v += beta * y[(long)col * incy];
        y[(long)col * incy] = v;
    }
}

// ---------------------------------------------------------------------
// sgemv3 / sgemv4: 64-row x 16-column-group tile per workgroup (one wave)
// ---------------------------------------------------------------------
__device__ void gather_tile(int d, int nz, int ncol, const float *__restrict__ A,
                            float *__restrict__ Z, const int *__restrict__ col0,
                            const int *__restrict__ col1, int bx, int by) {
    const int lane = threadIdx.x, r = lane & 15, grp = lane >> 4;
    const int sentinel = 2 * nz * d;
    // first sentinel of this group: stores stop there (sgemv3.cu:141-157)
    int c = by * 16 + r;
    bool valid = c < ncol && col1[c] < sentinel;
    unsigned long long bal = __ballot(!valid);
    int first_bad = (int)__ffsll((long long)(bal & 0xFFFFull)) - 1;  // lanes 0..15 carry the group
    if (first_bad < 0) first_bad = 16;
    const bool live = r < first_bad;
    const int c0 = live ? col0[c] : 0;
    const int row0 = bx * 64;
    f4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < d; k0 += 4) {
        int kk = k0 + grp;
        float b = (live && kk < d) ? Z[c0 + kk] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int row = row0 + 16 * i + r;
            float a = (kk < d && row < d) ? A[row + (size_t)d * kk] : 0.0f;
            acc[i] = mfma4(a, b, acc[i]);
        }
    }
    if (!live) return;
    const int c1 = col1[c];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int row = row0 + 16 * i + 4 * grp + e;
            if (row < d) Z[c1 + row] = acc[i][e];
        }
}

__global__ __launch_bounds__(64) void k_sgemv3(int d, int nz, int ncol, const float *A, float *Z,
                                               const int *col0, const int *col1) {
    gather_tile(d, nz, ncol, A, Z, col0, col1, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(64) void k_sgemv4(const char *__restrict__ base,
                                               const uint32_t *__restrict__ argptr_bid) {
    const uint32_t *arg = (const uint32_t *)(base + argptr_bid[blockIdx.x]);
    int bx = (int)arg[1], by = (int)arg[2], d = (int)arg[3], nz = (int)arg[4];
    int ncol = (int)arg[5];
    const float *A = (const float *)(base + arg[6]);
    float *Z = (float *)(base + arg[7]);
    const int *col0 = (const int *)(base + arg[8]);
    const int *col1 = (const int *)(base + arg[9]);
    // the reference leaves ncol unused in the descriptor (sgemv4.cu:100-107);
    // 0 there means "the group is full, stop only at sentinels"
    if (ncol <= 0) ncol = (by + 1) * 16;
    gather_tile(d, nz, ncol, A, Z, col0, col1, bx, by);
}

// ---------------------------------------------------------------------
// skinny GEMM: Y[m x NCP] = alpha * (A[m x K] * X[K x NCP] [+ X]), X and Y
// row-major.  Grid (m/64 row blocks, KS K-chunks of kc); workgroup = 4
// waves on the same 64 rows, splitting the chunk's k16-blocks between them
// (wave w takes blocks w, w+4, ...).  The X chunk is staged in LDS once.
// A wave covers all 64 rows at once: lane (r, g) loads the float4
// A[4r .. 4r+3][k] (256 contiguous bytes per k across the 16 lanes) and
// feeds component v to MFMA tile v, so tile v holds rows 4i + v; one X
// operand from LDS serves four MFMAs.  All A loads of a wave are issued
// before its first MFMA.  The 4 waves' partial tiles are added in wave
// order through LDS; with KS > 1 the workgroup partials go to a workspace
// and k_skinny_combine adds them in chunk order (deterministic).
// ---------------------------------------------------------------------
constexpr int kSkKB = 4;  // k16-blocks per wave per chunk (kc = 16 * 4 * kSkKB = 256)

template <int NT>
__global__ __launch_bounds__(256) void k_skinny(int m, int K, int kc, const float *__restrict__ A,
                                                int lda, const float *__restrict__ X, int ldx,
                                                float *__restrict__ out, int ldo, float alpha,
                                                int plus_x) {
    constexpr int XP = 16 * NT + 4;  // LDS row pitch (floats)
    extern __shared__ float xs[];    // kc x XP, reused for the wave reduction
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = lane & 15, grp = lane >> 4;
    const int row0 = blockIdx.x * 64;
    const int kbeg = blockIdx.y * kc, kn = min(K, kbeg + kc) - kbeg;
    // stage X[kbeg : kbeg+kn, 0 : 16NT] first (rows of 16NT floats, float4
    // loads): the barrier below then waits for these loads only, and the A
    // loads issued after them stay in flight while the first k16-blocks
    // are multiplied (in-order completion counting)
    constexpr int Q = 4 * NT;             // float4 per row
    constexpr int KC = 64 * kSkKB;        // chunk rows (kc == KC)
    constexpr int NX = KC * Q / 256;      // float4 per thread
    f4 xl[NX];
#pragma unroll
    for (int q = 0; q < NX; q++) {        // all loads first, then the LDS stores
        const int i = t + 256 * q, k = i / Q, c4 = i - k * Q;
        xl[q] = *(const f4 *)(X + (size_t)(kbeg + (k < kn ? k : kn - 1)) * ldx + 4 * c4);
    }
    // A loads: k16-block b = wave + 4u
    const float *Ar = A + row0 + 4 * r + (size_t)lda * kbeg;
    f4 av[kSkKB][4];
#pragma unroll
    for (int u = 0; u < kSkKB; u++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = 16 * (wave + 4 * u) + 4 * grp + s;
            // past the chunk: a clamped (valid, finite) column times a zero X row
            av[u][s] = *(const f4 *)(Ar + (size_t)lda * (k < kn ? k : kn - 1));
        }
    // the X chunk to LDS: waits only for the X loads, issued before A
#pragma unroll
    for (int q = 0; q < NX; q++) {
        const int i = t + 256 * q, k = i / Q, c4 = i - k * Q;
        *(f4 *)&xs[k * XP + 4 * c4] = (k < kn) ? xl[q] : f4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();
    f4 acc[4][NT];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[v][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < kSkKB; u++) {
        float xv[4][NT];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = 16 * (wave + 4 * u) + 4 * grp + s;
#pragma unroll
            for (int j = 0; j < NT; j++) xv[s][j] = xs[k * XP + 16 * j + r];
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int j = 0; j < NT; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[v][j] = mfma4(av[u][s][v], xv[s][j], acc[v][j]);
    }
    // add the 4 waves' tiles in wave order (xs is free once every wave is here)
    __syncthreads();
    f4 *red = (f4 *)xs;  // [wave][v][j][lane]
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
        for (int j = 0; j < NT; j++) red[((wave * 4 + v) * NT + j) * 64 + lane] = acc[v][j];
    __syncthreads();
    // wave w finishes tile v = w: rows 4i + w, i = 4g + e
    const int v = wave;
    float *o = out + (size_t)blockIdx.y * m * ldo;
    const bool direct = gridDim.y == 1;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        f4 s = red[((0 * 4 + v) * NT + j) * 64 + lane];
        s += red[((1 * 4 + v) * NT + j) * 64 + lane];
        s += red[((2 * 4 + v) * NT + j) * 64 + lane];
        s += red[((3 * 4 + v) * NT + j) * 64 + lane];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = row0 + 4 * (4 * grp + e) + v;
            float y = s[e];
            if (direct) {
                y *= alpha;
                if (plus_x) y += alpha * X[(size_t)row * ldx + 16 * j + r];
            }
            o[(size_t)row * ldo + 16 * j + r] = y;
        }
    }
}

__global__ void k_skinny_combine(int m, int ncp, int KS, const float *__restrict__ part,
                                 float *__restrict__ Y, int ldy, float alpha,
                                 const float *__restrict__ X, int ldx, int plus_x) {
    // four outputs per thread (ncp % 16 == 0, ldx % 4 == 0, ldy % 4 == 0);
    // partials loaded four at a time, added in chunk order
    const int idx = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (idx >= m * ncp) return;
    const int i = idx / ncp, c = idx - i * ncp;
    const size_t mn = (size_t)m * ncp;
    // every partial load is issued before the first add (KS <= 16)
    f4 p[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
        if (q < KS) p[q] = *(const f4 *)(part + q * mn + idx);
    f4 s = p[0];
#pragma unroll
    for (int q = 1; q < 16; q++)
        if (q < KS) s += p[q];
    for (int q = 16; q < KS; q++) s += *(const f4 *)(part + q * mn + idx);
    f4 y = alpha * s;
    if (plus_x) y += alpha * *(const f4 *)(X + (size_t)i * ldx + c);
    *(f4 *)(Y + (size_t)i * ldy + c) = y;
}

}  // namespace

OPHIP_EXPORT void opcuda_sgemv(int m, int n, float alpha, unsigned A, int lda, unsigned x,
                               int incx, float beta, unsigned y, int incy) {
    Device &D = dev();
    if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || incx == 0 || incy == 0) {
        set_status(kErrBadArg);
        return;
    }
    if (m == 0) { set_status(kOk); return; }
    const float *Ap = P<const float>(A), *xp = P<const float>(x);
    float *yp = P<float>(y);
    if (!yp || (n > 0 && (!Ap || !xp))) { set_status(kErrBadHandle); return; }
    if (incx < 0) xp += (long)(n - 1) * (-incx);
    if (incy < 0) yp += (long)(m - 1) * (-incy);
    int rb = cdiv(m, 256);
    int S = n > 0 ? cdiv(1024, rb) : 1;
    if (S > n && n > 0) S = n;
    int kc = n > 0 ? cdiv(n, S) : 0;
    if (n > 0) S = cdiv(n, kc);
    float *part = (float *)workspace((size_t)S * m * sizeof(float));
    if (!part) return;
    int vec = ((lda % 4) == 0 && (((uintptr_t)Ap) & 15) == 0);
    if (n > 0) {
        hipLaunchKernelGGL(k_gemvN_part, dim3(rb, S), dim3(64), 0, D.stream, m, n, Ap, lda, xp, incx,
                           kc, part, vec);
    } else {
        S = 0;
    }
    hipLaunchKernelGGL(k_gemv_combine, dim3(cdiv(m, 256)), dim3(256), 0, D.stream, m, S, part,
                       alpha, beta, yp, incy);
    check_launch();
}

OPHIP_EXPORT void opcuda_sgemvT(int m, int n, float alpha, unsigned A, int lda, unsigned x,
                                int incx, float beta, unsigned y, int incy) {
    Device &D = dev();
    if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || incx == 0 || incy == 0) {
        set_status(kErrBadArg);
        return;
    }
    if (n == 0) { set_status(kOk); return; }
    const float *Ap = P<const float>(A), *xp = P<const float>(x);
    float *yp = P<float>(y);
    if (!yp || (m > 0 && (!Ap || !xp))) { set_status(kErrBadHandle); return; }
    if (incx < 0) xp += (long)(m - 1) * (-incx);
    if (incy < 0) yp += (long)(n - 1) * (-incy);
    hipLaunchKernelGGL(k_gemvT, dim3(cdiv(n, 4)), dim3(256), 0, D.stream, m, n, Ap, lda, xp, incx,
                       alpha, beta, yp, incy);
    check_launch();
}

OPHIP_EXPORT void opcuda_sgemv3(int d, int nz, int ncol, unsigned A, unsigned Z, unsigned col0,
                                unsigned col1) {
    Device &D = dev();
    if (d <= 0 || ncol <= 0) { set_status(kOk); return; }
    if (!ptr(A) || !ptr(Z) || !ptr(col0) || !ptr(col1)) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_sgemv3, dim3(cdiv(d, 64), cdiv(ncol, 16)), dim3(64), 0, D.stream, d, nz,
                       ncol, P<const float>(A), P<float>(Z), P<const int>(col0),
                       P<const int>(col1));
    check_launch();
}

OPHIP_EXPORT void opcuda_sgemv4(int nblocks, unsigned argptr_bid) {
    Device &D = dev();
    if (nblocks <= 0) { set_status(kOk); return; }
    const uint32_t *tab = P<const uint32_t>(argptr_bid);
    if (!tab) { set_status(kErrBadHandle); return; }
    hipLaunchKernelGGL(k_sgemv4, dim3(nblocks), dim3(64), 0, D.stream, (const char *)D.base, tab);
    check_launch();
}

// Y = alpha * (A * X [+ X]).  A: m x K column-major (lda); X: K x ncp
// row-major (ldx >= ncp); Y: m x ncp row-major (ldy >= ncp).  ncp must be
// 16, 32, 48 or 64 (pad the right-hand side with zero columns); m % 16 == 0.
// plus_x = 1 adds X itself (requires m == K): one backward step with
// U = I + W stored as W.
static int skinny(int m, int K, int ncp, float alpha, unsigned A, int lda, unsigned X, int ldx,
                  unsigned Y, int ldy, int plus_x) {
    Device &D = dev();
    if (m <= 0 || K <= 0 || (m % 64) || ncp <= 0 || ncp > 64 || (ncp % 16) || ncp > 48 || ldx < ncp ||
        (ldx % 4) || (ldy % 4) || ldy < ncp || lda < m || (lda % 4) || (plus_x && m != K))
        return (int)set_status(kErrBadArg);
    const float *Ap = P<const float>(A), *Xp = P<const float>(X);
    float *Yp = P<float>(Y);
    if (!Ap || !Xp || !Yp || (((uintptr_t)Xp | (uintptr_t)Ap) & 15))
        return (int)set_status(kErrBadHandle);
    const int NT = ncp / 16;
    const int kc = 64 * kSkKB;  // 4 waves x kSkKB k16-blocks
    const int KS = cdiv(K, kc);
    float *out = Yp;
    int ldo = ldy;
    if (KS > 1) {
        out = (float *)workspace((size_t)KS * m * ncp * sizeof(float));
        if (!out) return (int)kErrOutOfArena;
        ldo = ncp;
    }
    // LDS: the X chunk, or the 4 waves' partial tiles, whichever is larger
    const size_t lds_x = sizeof(float) * (size_t)kc * (16 * NT + 4);
    const size_t lds_r = sizeof(float) * 4 * 4 * NT * 64 * 4;
    const size_t lds = lds_x > lds_r ? lds_x : lds_r;
    dim3 grid(m / 64, KS), block(256);
    switch (NT) {
    case 1: hipLaunchKernelGGL(k_skinny<1>, grid, block, lds, D.stream, m, K, kc, Ap, lda, Xp, ldx, out, ldo, alpha, plus_x); break;
    case 2: hipLaunchKernelGGL(k_skinny<2>, grid, block, lds, D.stream, m, K, kc, Ap, lda, Xp, ldx, out, ldo, alpha, plus_x); break;
    case 3: hipLaunchKernelGGL(k_skinny<3>, grid, block, lds, D.stream, m, K, kc, Ap, lda, Xp, ldx, out, ldo, alpha, plus_x); break;
    default: break;
    }
    if (KS > 1)
        hipLaunchKernelGGL(k_skinny_combine, dim3(cdiv(m * ncp / 4, 256)), dim3(256), 0,
                           D.stream, m, ncp, KS, out, Yp, ldy, alpha, Xp, ldx, plus_x);
    return (int)check_launch();
}

OPHIP_EXPORT int ophip_sgemm_skinny(int m, int K, int ncp, float alpha, unsigned A, int lda,
                                    unsigned X, int ldx, unsigned Y, int ldy) {
    return skinny(m, K, ncp, alpha, A, lda, X, ldx, Y, ldy, 0);
}

OPHIP_EXPORT int ophip_sgemm_skinny_ex(int m, int K, int ncp, float alpha, unsigned A, int lda,
                                       unsigned X, int ldx, unsigned Y, int ldy, int plus_x) {
    return skinny(m, K, ncp, alpha, A, lda, X, ldx, Y, ldy, plus_x);
}
