// lattice.hip -- generator assembly, first exponentiation step and
// cumulative transition tables.
//
// Drop-in entry points (FP32, the reference's semantics):
//   opcuda_sek / opcuda_sekd / opcuda_sekd_lambda   (sek.cu:40-138)
//   opcuda_sck                                      (sck.cu:43-93)
//   opcuda_sglvg                                    (sglvg.cu:42-125)
// MI355X additions (FP64 on the GPU where precision needs it):
//   ophip_dgen_rs     FP64 generator of a regime-switching local-vol lattice
//   ophip_sek_rs      the same generator fused with U0 = I + dt*L, evaluated
//                     in FP64 and rounded once to FP32 (or kept as FP64);
//                     the generator is never written to memory
//   ophip_sck_rows    cumulative tables stored row-contiguous for the
//                     sampler (one pass over 8-row blocks up to d = 2048)
#include "ophip_internal.h"
#include "rs_lattice.h"

using namespace ophip;

// ---------------------------------------------------------------------
// FP32 first step, drop-in (sek.cu:40-103)
// ---------------------------------------------------------------------

// mode 0: id + dt*gen; 1: (id + dt*gen)/(1 + dt*r[y1]); 2: ... /(1+dt*r*lambda)
template <int MODE>
__global__ void k_sek(int d, const float *__restrict__ gen, float dt, float *__restrict__ tpk,
                      const float *__restrict__ r, const float *__restrict__ lambda) {
    size_t n = (size_t)d * d;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        int y1 = (int)(e % d), y2 = (int)(e / d);
        float v = fmaf(dt, gen[e], (y1 == y2) ? 1.0f : 0.0f);
        if (MODE == 1) v = v / fmaf(dt, r[y1], 1.0f);
        if (MODE == 2) v = v / fmaf(dt * r[y1], lambda[0], 1.0f);
        tpk[e] = v;
    }
}

static int grid1d(size_t n, int bs) {
    size_t g = (n + bs - 1) / bs;
    if (g > 16384) g = 16384;
    return g ? (int)g : 1;
}

OPHIP_EXPORT void opcuda_sek(int d, unsigned gen, float dt, unsigned tpk) {
    Device &D = dev();
    if (d <= 0 || !ptr(gen) || !ptr(tpk)) { set_status(kErrBadArg); return; }
    hipLaunchKernelGGL(k_sek<0>, dim3(grid1d((size_t)d * d, 256)), dim3(256), 0, D.stream, d,
                       P<const float>(gen), dt, P<float>(tpk), nullptr, nullptr);
    check_launch();
}

OPHIP_EXPORT void opcuda_sekd(int d, unsigned gen, float dt, unsigned tpk, unsigned r) {
    Device &D = dev();
    if (d <= 0 || !ptr(gen) || !ptr(tpk) || !ptr(r)) { set_status(kErrBadArg); return; }
    hipLaunchKernelGGL(k_sek<1>, dim3(grid1d((size_t)d * d, 256)), dim3(256), 0, D.stream, d,
                       P<const float>(gen), dt, P<float>(tpk), P<const float>(r), nullptr);
    check_launch();
}

OPHIP_EXPORT void opcuda_sekd_lambda(int d, unsigned gen, float dt, unsigned tpk, unsigned r,
                                     unsigned lambda) {
    Device &D = dev();
    if (d <= 0 || !ptr(gen) || !ptr(tpk) || !ptr(r) || !ptr(lambda)) {
        set_status(kErrBadArg);
        return;
    }
    hipLaunchKernelGGL(k_sek<2>, dim3(grid1d((size_t)d * d, 256)), dim3(256), 0, D.stream, d,
                       P<const float>(gen), dt, P<float>(tpk), P<const float>(r),
                       P<const float>(lambda));
    check_launch();
}

// ---------------------------------------------------------------------
// FP32 LV generator, drop-in (sglvg.cu:42-125).  One thread per (i, y0);
// the row is zeroed and its two rates written.  pars = [drift(ni*d),
// vol(ni*d)], invm = 4 floats per node (column-major 2x2).
// ---------------------------------------------------------------------

__global__ void k_sglvg(float *__restrict__ gen, int ni, int d, const float *__restrict__ invm,
                        const float *__restrict__ drift, const float *__restrict__ vol) {
    size_t total = (size_t)ni * d * d;
    // zero pass, coalesced
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x)
        gen[e] = 0.0f;
}

__global__ void k_sglvg_rates(float *__restrict__ gen, int ni, int d,
                              const float *__restrict__ invm, const float *__restrict__ drift,
                              const float *__restrict__ vol) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * d) return;
    int i = t / d, y0 = t - i * d;
    if (y0 == 0 || y0 == d - 1) return;  // absorbing boundary rows stay zero
    float r0 = drift[y0 + d * i];
    float v = vol[y0 + d * i];
    float r1 = v * v;
    float *g = gen + (size_t)i * d * d;
    float lo = fmaf(invm[4 * y0 + 2], r1, invm[4 * y0 + 0] * r0);
    float hi = fmaf(invm[4 * y0 + 3], r1, invm[4 * y0 + 1] * r0);
    g[y0 + (size_t)d * (y0 - 1)] = lo;
    g[y0 + (size_t)d * (y0 + 1)] = hi;
    g[y0 + (size_t)d * y0] = -(lo + hi);
}

OPHIP_EXPORT void opcuda_sglvg(unsigned gen, int ni, int d, unsigned invm, unsigned pars) {
    Device &D = dev();
    if (ni <= 0 || d < 2 || !ptr(gen) || !ptr(invm) || !ptr(pars)) {
        set_status(kErrBadArg);
        return;
    }
    const float *drift = P<const float>(pars);
    const float *vol = drift + (size_t)ni * d;
    hipLaunchKernelGGL(k_sglvg, dim3(grid1d((size_t)ni * d * d, 256)), dim3(256), 0, D.stream,
                       P<float>(gen), ni, d, P<const float>(invm), drift, vol);
    hipLaunchKernelGGL(k_sglvg_rates, dim3(cdiv(ni * d, 256)), dim3(256), 0, D.stream,
                       P<float>(gen), ni, d, P<const float>(invm), drift, vol);
    check_launch();
}

// ---------------------------------------------------------------------
// FP64 regime-switching local-vol lattice (RSLattice, rs_elem:
// rs_lattice.h)
// ---------------------------------------------------------------------

template <int MODE, typename OUT, bool NM>
__global__ void k_rs_assemble(RSLattice L, double dt, OUT *__restrict__ out) {
    const int d = L.nx * L.nr;
    size_t n = (size_t)d * d;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % d), j = (int)(e / d);  // row i (from), column j (to)
        out[e] = (OUT)rs_elem<MODE, NM>(L, dt, i, j);
    }
}

// Many lattices in one launch (blockIdx.y = lattice): each thread writes
// four consecutive rows of one column (d % 4 == 0), 32-bit index math
// (d * d / 4 < 2^32).  The values are those of k_rs_assemble.
struct RSEntry {
    RSLattice L;
    double dt;
    void *out;
};
constexpr int kRSBatch = 32;
struct RSBatch {
    RSEntry e[kRSBatch];
};

template <int MODE, typename OUT, bool NM>
__global__ __launch_bounds__(256) void k_rs_assemble_batch(const RSBatch b) {
    const RSEntry &E = b.e[blockIdx.y];
    const unsigned d = (unsigned)(E.L.nx * E.L.nr), d4 = d / 4, n4 = d * d4;
    OUT *out = reinterpret_cast<OUT *>(E.out);
    for (unsigned q = blockIdx.x * 256u + threadIdx.x; q < n4; q += gridDim.x * 256u) {
        const unsigned j = q / d4, i = 4 * (q - j * d4);
        OUT v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = (OUT)rs_elem<MODE, NM>(E.L, E.dt, (int)i + u, (int)j);
        OUT *o = out + (size_t)j * d + i;
        if (sizeof(OUT) == 4) {
            *reinterpret_cast<float4 *>(o) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
            reinterpret_cast<double2 *>(o)[0] = make_double2(v[0], v[1]);
            reinterpret_cast<double2 *>(o)[1] = make_double2(v[2], v[3]);
        }
    }
}

// FP64 generator (d = nx*nr, d*d doubles at `gen`)
OPHIP_EXPORT int ophip_dgen_rs(int nx, int nr, unsigned xval, unsigned drift, unsigned vol,
                               double q_off, unsigned gen) {
    Device &D = dev();
    RSLattice L;
    if (!rs_args(nx, nr, xval, drift, vol, L) || !ptr(gen)) return (int)set_status(kErrBadArg);
    L.q_off = q_off;
    size_t d = (size_t)nx * nr;
    hipLaunchKernelGGL((k_rs_assemble<0, double, false>), dim3(grid1d(d * d, 256)), dim3(256), 0,
                       D.stream, L, 0.0, P<double>(gen));
    return (int)check_launch();
}

// Fused assembly + first step.  out_kind 0: FP32 I+dt*L; 1: FP64 I+dt*L;
// 2: FP32 dt*L (the U-I form); 3: FP64 dt*L; + 4: node-major states
// (y = nr*x + r instead of x + nx*r).
OPHIP_EXPORT int ophip_sek_rs(int nx, int nr, unsigned xval, unsigned drift, unsigned vol,
                              double q_off, double dt, int out_kind, unsigned out) {
    Device &D = dev();
    RSLattice L;
    if (!rs_args(nx, nr, xval, drift, vol, L) || !ptr(out)) return (int)set_status(kErrBadArg);
    L.q_off = q_off;
    size_t d = (size_t)nx * nr;
    dim3 g(grid1d(d * d, 256)), b(256);
    hipStream_t st = D.stream;
    switch (out_kind) {
    case 0: hipLaunchKernelGGL((k_rs_assemble<1, float, false>), g, b, 0, st, L, dt, P<float>(out)); break;
    case 1: hipLaunchKernelGGL((k_rs_assemble<1, double, false>), g, b, 0, st, L, dt, P<double>(out)); break;
    case 2: hipLaunchKernelGGL((k_rs_assemble<2, float, false>), g, b, 0, st, L, dt, P<float>(out)); break;
    case 3: hipLaunchKernelGGL((k_rs_assemble<2, double, false>), g, b, 0, st, L, dt, P<double>(out)); break;
    case 4: hipLaunchKernelGGL((k_rs_assemble<1, float, true>), g, b, 0, st, L, dt, P<float>(out)); break;
    case 5: hipLaunchKernelGGL((k_rs_assemble<1, double, true>), g, b, 0, st, L, dt, P<double>(out)); break;
    case 6: hipLaunchKernelGGL((k_rs_assemble<2, float, true>), g, b, 0, st, L, dt, P<float>(out)); break;
    case 7: hipLaunchKernelGGL((k_rs_assemble<2, double, true>), g, b, 0, st, L, dt, P<double>(out)); break;
    default: return (int)set_status(kErrBadArg);
    }
    return (int)check_launch();
}

// ophip_sek_rs for n lattices in one launch per 32: lattice k takes the
// k-th entry of every array; all share out_kind.  Each output holds the
// values ophip_sek_rs writes.  Every d = nx[k]*nr[k] must be a multiple of
// 4 (else each lattice is assembled by its own ophip_sek_rs launch).
OPHIP_EXPORT int ophip_sek_rs_batch(int n, const int *nx, const int *nr, const unsigned *xval,
                                    const unsigned *drift, const unsigned *vol, const double *q_off,
                                    const double *dt, int out_kind, const unsigned *out) {
    if (n <= 0 || !nx || !nr || !xval || !drift || !vol || !q_off || !dt || !out || out_kind < 0 ||
        out_kind > 7)
        return (int)set_status(kErrBadArg);
    bool vec = true;
    for (int k = 0; k < n; k++) {
        RSLattice L;
        if (!rs_args(nx[k], nr[k], xval[k], drift[k], vol[k], L) || !ptr(out[k]))
            return (int)set_status(kErrBadArg);
        vec = vec && (nx[k] * nr[k]) % 4 == 0;
    }
    if (!vec) {
        for (int k = 0; k < n; k++) {
            const int e = ophip_sek_rs(nx[k], nr[k], xval[k], drift[k], vol[k], q_off[k], dt[k], out_kind, out[k]);
            if (e) return e;
        }
        return 0;
    }
    Device &D = dev();
    for (int k0 = 0; k0 < n; k0 += kRSBatch) {
        const int m = n - k0 < kRSBatch ? n - k0 : kRSBatch;
        RSBatch b{};
        size_t most = 0;
        for (int k = 0; k < m; k++) {
            RSEntry &E = b.e[k];
            rs_args(nx[k0 + k], nr[k0 + k], xval[k0 + k], drift[k0 + k], vol[k0 + k], E.L);
            E.L.q_off = q_off[k0 + k];
            E.dt = dt[k0 + k];
            E.out = ptr(out[k0 + k]);
            const size_t d = (size_t)nx[k0 + k] * nr[k0 + k];
            most = d * d / 4 > most ? d * d / 4 : most;
        }
        const dim3 g((unsigned)grid1d(most, 256), (unsigned)m), bl(256);
        hipStream_t st = D.stream;
        switch (out_kind) {
        case 0: hipLaunchKernelGGL((k_rs_assemble_batch<1, float, false>), g, bl, 0, st, b); break;
        case 1: hipLaunchKernelGGL((k_rs_assemble_batch<1, double, false>), g, bl, 0, st, b); break;
        case 2: hipLaunchKernelGGL((k_rs_assemble_batch<2, float, false>), g, bl, 0, st, b); break;
        case 3: hipLaunchKernelGGL((k_rs_assemble_batch<2, double, false>), g, bl, 0, st, b); break;
        case 4: hipLaunchKernelGGL((k_rs_assemble_batch<1, float, true>), g, bl, 0, st, b); break;
        case 5: hipLaunchKernelGGL((k_rs_assemble_batch<1, double, true>), g, bl, 0, st, b); break;
        case 6: hipLaunchKernelGGL((k_rs_assemble_batch<2, float, true>), g, bl, 0, st, b); break;
        default: hipLaunchKernelGGL((k_rs_assemble_batch<2, double, true>), g, bl, 0, st, b); break;
        }
        const unsigned e = check_launch();
        if (e) return (int)e;
    }
    return 0;
}

// ---------------------------------------------------------------------
// cumulative tables
// ---------------------------------------------------------------------

// Drop-in layout (sck.cu:43-84): ctpk[m*d*d + row + d*y].  One thread per
// (m,row): sequential FP32 running sum of the raw entries along the row (a
// negative entry lowers the sum, as in the reference), then divide by the
// total and cap at 1.
__global__ void k_sck(int d, int nm, const float *__restrict__ tpk, float *__restrict__ ctpk) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm * d) return;
    int m = t / d, row = t - m * d;
    size_t base = (size_t)m * d * d + row;
    const float *tp = tpk + base;
    float *cp = ctpk + base;
    float s = 0.0f;
    for (int y = 0; y < d; y++) {
        s += tp[(size_t)d * y];
        cp[(size_t)d * y] = s;
    }
    for (int y = 0; y < d; y++) {
        float v = cp[(size_t)d * y] / s;
        cp[(size_t)d * y] = (v < 1.0f) ? v : 1.0f;
    }
}

OPHIP_EXPORT void opcuda_sck(int d, int nm, unsigned tpk, unsigned ctpk) {
    Device &D = dev();
    if (d <= 0 || nm <= 0 || !ptr(tpk) || !ptr(ctpk)) { set_status(kErrBadArg); return; }
    hipLaunchKernelGGL(k_sck, dim3(cdiv(nm * d, 64)), dim3(64), 0, D.stream, d, nm,
                       P<const float>(tpk), P<float>(ctpk));
    check_launch();
}

// Row-contiguous table crow[m*d*d + row*d + y] in a fixed two-level
// order: each row is cut into 64-column chunks; chunk q holds a sequential
// FP32 running sum l_q(y) of its entries (negative entries count as zero:
// a rounding-level negative would make the row non-monotone; with
// plus_identity 1 is added to the diagonal entry first), the chunk totals
// t_q = l_q(last) are chained in order (carry_0 = 0, carry_{q+1} =
// carry_q + t_q), the running sum is S(y) = carry_q + l_q(y), the row total
// T = S(d-1), and C(y) = min(1, S(y)/T): C(d-1) = 1 exactly, rows are
// non-decreasing, and the values are within a few ulps of the sequential
// order (bit for bit where every partial sum is exact).
//
// For d > 2048 (k_sck_one below takes the smaller tables in one pass):
// k_sck_part: one wavefront per 64 rows x one chunk (lane = row: the 64
// values of its chunk with 64 coalesced loads, 64 consecutive rows per
// column), writes t_q.  k_sck_fin: the same wavefronts again, each chains
// its rows' totals (the same adds in the same order in every wavefront,
// so all of them see the same T), re-forms l_q (the same adds as
// k_sck_part) and writes C through an LDS tile, so the row-contiguous
// stores are coalesced.  Grid (row blocks, chunks, tables).  Rows and
// columns past d are clamped reads that are never used.
constexpr int kSckChunk = 64;

__device__ inline void sck_chunk_load(const float *__restrict__ src, int d, int y0, float (&v)[kSckChunk]) {
#pragma unroll
    for (int c = 0; c < kSckChunk; c++) v[c] = src[(size_t)d * min(y0 + c, d - 1)];
}

__device__ inline float sck_chunk_sum(const float (&v)[kSckChunk], int ny, int row, int y0, int plus_identity,
                                      float (*run)[kSckChunk + 1], int lane) {
    float l = 0.0f;
#pragma unroll
    for (int c = 0; c < kSckChunk; c++) {
        if (c < ny) {
            float x = v[c];
            if (plus_identity && row == y0 + c) x += 1.0f;
            l += (x > 0.0f) ? x : 0.0f;
        }
        if (run) run[lane][c] = l;
    }
    return l;
}

__global__ __launch_bounds__(64) void k_sck_part(int d, const float *__restrict__ tpk, float *__restrict__ tq,
                                                  int plus_identity) {
    const int lane = threadIdx.x, nq = gridDim.y;
    const int row = blockIdx.x * 64 + lane, q = blockIdx.y, y0 = q * kSckChunk;
    const size_t mb = (size_t)blockIdx.z * d * d;
    float v[kSckChunk];
    sck_chunk_load(tpk + mb + min(row, d - 1), d, y0, v);
    const float l = sck_chunk_sum(v, min(kSckChunk, d - y0), row, y0, plus_identity, nullptr, lane);
    if (row < d) tq[((size_t)blockIdx.z * nq + q) * d + row] = l;
}

__global__ __launch_bounds__(64) void k_sck_fin(int d, const float *__restrict__ tpk,
                                                 const float *__restrict__ tq, float *__restrict__ crow,
                                                 int plus_identity) {
    __shared__ float tile[64][kSckChunk + 1];
    const int lane = threadIdx.x, nq = gridDim.y;
    const int row0 = blockIdx.x * 64, row = row0 + lane, q = blockIdx.y, y0 = q * kSckChunk;
    const size_t mb = (size_t)blockIdx.z * d * d;
    float v[kSckChunk];
    sck_chunk_load(tpk + mb + min(row, d - 1), d, y0, v);
    const float *t = tq + (size_t)blockIdx.z * nq * d + min(row, d - 1);
    float carry = 0.0f;
    for (int p = 0; p < q; p++) carry = carry + t[(size_t)p * d];
    float T = carry;
    for (int p = q; p < nq; p++) T = T + t[(size_t)p * d];
    const int ny = min(kSckChunk, d - y0);
    sck_chunk_sum(v, ny, row, y0, plus_identity, tile, lane);
#pragma unroll
    for (int c = 0; c < kSckChunk; c++) {
        const float s = (carry + tile[lane][c]) / T;
        tile[lane][c] = s < 1.0f ? s : 1.0f;
    }
    __syncthreads();
    for (int r = 0; r < 64 && row0 + r < d; r++)
        if (lane < ny) crow[mb + (size_t)(row0 + r) * d + y0 + lane] = tile[r][lane];
}

// k_sck_one: the same table in one pass for d <= 2048, the same adds in
// the same order.  A workgroup takes R rows (R = 8) and holds them whole in
// registers: lane l of wave w owns row l % R of the block and chunk
// (64 / R) w + l / R (64 values), so wave w holds 64 (64 / R) consecutive
// columns of the R rows.  Each thread forms its chunk's running sums, the
// chunk totals meet in LDS, every thread chains its row's totals in order
// (carry before its chunk, then on to T), and each wave writes its block
// through its own LDS tile with coalesced row stores.  Every entry is read
// once and written once (8 bytes per entry).  Row blocks are dealt to XCDs
// in contiguous runs (block b runs on XCD b % 8), so the blocks that share
// the 128-byte lines of 32 consecutive rows read them through one L2.
#ifndef OPHIP_SCK_R
#define OPHIP_SCK_R 8
#endif
constexpr int kSckRows = OPHIP_SCK_R, kSckCPW = 64 / kSckRows, kSckCP = kSckChunk + 1,
              kSckPitch = kSckCPW * kSckCP + 1, kSckWaves = 32 / kSckCPW;

__global__ __launch_bounds__(64 * kSckWaves) void k_sck_one(int d, const float *__restrict__ tpk,
                                                             float *__restrict__ crow, int plus_identity,
                                                             int per_xcd) {
    __shared__ float tq[kSckRows][33];
    __shared__ float st[kSckWaves][kSckRows * kSckPitch];
    const int nrb = (d + kSckRows - 1) / kSckRows;
    const int rb = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (rb >= nrb) return;  // whole workgroup, before any barrier
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nq = (d + kSckChunk - 1) / kSckChunk;  // <= 32
    const int rl = lane % kSckRows, qq = lane / kSckRows, q = kSckCPW * w + qq;
    const int row0 = rb * kSckRows, row = row0 + rl, y0 = q * kSckChunk;
    const size_t mb = (size_t)blockIdx.y * d * d;
    const bool act = q < nq;
    float v[kSckChunk];
    if (act) sck_chunk_load(tpk + mb + min(row, d - 1), d, y0, v);
    const int ny = act ? min(kSckChunk, d - y0) : 0;
    float l = 0.0f;
#pragma unroll
    for (int c = 0; c < kSckChunk; c++) {
        if (c < ny) {
            float x = v[c];
            if (plus_identity && row == y0 + c) x += 1.0f;
            l += (x > 0.0f) ? x : 0.0f;
        }
        v[c] = l;
    }
    if (act) tq[rl][q] = l;
    __syncthreads();
    float carry = 0.0f, T = 1.0f;
    if (act) {
        for (int p = 0; p < q; p++) carry = carry + tq[rl][p];
        T = carry;
        for (int p = q; p < nq; p++) T = T + tq[rl][p];
    }
    float *tile = st[w];
    {
        float *dst = tile + rl * kSckPitch + qq * kSckCP;
#pragma unroll
        for (int c = 0; c < kSckChunk; c++) {
            const float s = (carry + v[c]) / T;
            dst[c] = s < 1.0f ? s : 1.0f;
        }
    }
    __syncthreads();
    const int c0 = 64 * kSckCPW * w, ncol = min(64 * kSckCPW, d - c0);  // this wave's columns
    if (ncol > 0)
        for (int r = 0; r < kSckRows && row0 + r < d; r++)
            for (int c = lane; c < ncol; c += 64)
                crow[mb + (size_t)(row0 + r) * d + c0 + c] = tile[r * kSckPitch + (c >> 6) * kSckCP + (c & 63)];
}

static int sck_rows(int d, int nm, unsigned tpk, unsigned crow, int plus_identity) {
    Device &D = dev();
    if (d <= 0 || nm <= 0 || nm > 65535 || !ptr(tpk) || !ptr(crow)) return (int)set_status(kErrBadArg);
    if (!handle_range_ok(tpk, 4 * (size_t)nm * d * d) || !handle_range_ok(crow, 4 * (size_t)nm * d * d))
        return (int)set_status(kErrBadArg);
    const int nq = cdiv(d, kSckChunk);
    if (nq <= 32) {  // one pass: 64 / R chunks per wavefront
        const int per_xcd = cdiv(cdiv(d, kSckRows), 8);
        hipLaunchKernelGGL(k_sck_one, dim3(8 * per_xcd, nm), dim3(64 * cdiv(nq, kSckCPW)), 0, D.stream, d,
                           P<const float>(tpk), P<float>(crow), plus_identity, per_xcd);
        return (int)check_launch();
    }
    float *tq = (float *)workspace(sizeof(float) * (size_t)nm * nq * d);
    if (!tq) return (int)kErrOutOfArena;
    const dim3 g(cdiv(d, 64), nq, nm);
    hipLaunchKernelGGL(k_sck_part, g, dim3(64), 0, D.stream, d, P<const float>(tpk), tq, plus_identity);
    hipLaunchKernelGGL(k_sck_fin, g, dim3(64), 0, D.stream, d, P<const float>(tpk), (const float *)tq,
                       P<float>(crow), plus_identity);
    return (int)check_launch();
}

OPHIP_EXPORT int ophip_sck_rows(int d, int nm, unsigned tpk, unsigned crow) {
    return sck_rows(d, nm, tpk, crow, 0);
}

// the same table for U = I + W given W (the W-form kernels): 1 is added to
// each diagonal entry in FP32 before the running sum
OPHIP_EXPORT int ophip_sck_rows_ex(int d, int nm, unsigned tpk, unsigned crow, int plus_identity) {
    return sck_rows(d, nm, tpk, crow, plus_identity);
}

// ---------------------------------------------------------------------
// Stochastic-volatility lattice generator in FP64 (CSVModel.host_d_gen,
// CSVModel.cs:187-379; the reference's FP32 device version global_sgsvg,
// sgsvg.cu:40-209), with the 4x4 stencil of S2DGrid.pregen (S2DGrid.cs:
// 123-245) solved in closed form: the system is block diagonal, a 2x2 in
// x (zero drift, variance svol^2) and a 2x2 in r (drift voldrift,
// variance volvol^2; one-sided voldrift/delta at r = 0, nr-1), no x moves
// at x = 0, nx-1.  A positive rate up to regime r+1 is mirrored by jumps
// into regime r+1 (x1 < x0-1 weighted exp((x1-x0)/(jm x0)), x1 > x0
// weighted exp(-(x1-x0)/(jp x0)), normalised to that rate).  Negative
// rates are clipped; one neighbour sets the row drift sum L (x1-x0) to
// sdrift; the diagonal closes the row.  Two fixes against the reference:
// the jump weights use the scaled sizes jm x0, jp x0 in both the
// normalisation and the assignment (CSVModel.cs:296-297 vs 315, 321;
// sgsvg.cu:111-112 takes the minus size for both sides), and the drift
// correction on the down neighbour is (drift - sdrift)/(x0 - x_{-1}) > 0
// (CSVModel.cs:356-357 and sgsvg.cu:186 divide by x_{-1} - x0, which
// doubles the excess).  Sums run in state order, as the reference's.
//
// k_sv_rows: one thread per row, its few distinct values (stencil rates,
// jump normalisation, diagonal) into a row table; k_sv_fill: one thread
// per element, column-major (coalesced), the value from the row table.
// ---------------------------------------------------------------------
namespace {

struct SVLattice {
    int nx, nr;
    const double *xval, *rval, *sdrift, *svol, *volvol, *voldrift, *jm, *jp;
    // the caller's per-state 4x4 stencil inverses (S2DGrid.pregen's invm,
    // S2DGrid.cs:123-245: invm[16 y + j + 4 i] maps the moments (x drift,
    // x variance, r drift, r variance) to the rates to x-1, x+1, r-1, r+1,
    // as sgsvg.cu:72-103 applies it); nullptr: the closed form from rval
    const float *invm;
};

// row table: am, ap, bm, bp (after clipping and the drift fix), jump ratio,
// scaled jump sizes sm, sp (0: no jumps on that side), diagonal
constexpr int kSVRow = 8;

__global__ void k_sv_rows(SVLattice L, double *__restrict__ rt) {
    const int nx = L.nx, nr = L.nr, d = nx * nr;
    const int y0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (y0 >= d) return;
    const int x0 = y0 % nx, r0 = y0 / nx;
    const double X0 = L.xval[y0];
    double am = 0.0, ap = 0.0, bm = 0.0, bp = 0.0, ratio = 0.0, sm = 0.0, sp = 0.0;
    if (L.invm) {
        // rates = invm * (0, svol^2 on interior x, voldrift, volvol^2 on
        // interior r), the moments sgsvg.cu:58-68 sets, in FP64
        const bool ix = x0 > 0 && x0 < nx - 1, ir = r0 > 0 && r0 < nr - 1;
        const double rhs[4] = {0.0, ix ? L.svol[y0] * L.svol[y0] : 0.0, L.voldrift[y0],
                               ir ? L.volvol[y0] * L.volvol[y0] : 0.0};
        const float *iv = L.invm + (size_t)16 * y0;
        double rate[4];
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int i = 0; i < 4; i++) s += (double)iv[j + 4 * i] * rhs[i];
            rate[j] = s;
        }
        am = rate[0];
        ap = rate[1];
        bm = rate[2];
        bp = rate[3];
    } else if (x0 > 0 && x0 < nx - 1) {
        const double dm = L.xval[y0 - 1] - X0, dp = L.xval[y0 + 1] - X0;
        const double s2 = L.svol[y0] * L.svol[y0];
        const double det = dm * dp * dp - dp * dm * dm;
        am = (-dp * s2) / det;
        ap = (dm * s2) / det;
    }
    if (nr > 1) {
        if (L.invm) {
            // rates from the caller's stencil (above)
        } else if (r0 > 0 && r0 < nr - 1) {
            const double em = L.rval[y0 - nx] - L.rval[y0], ep = L.rval[y0 + nx] - L.rval[y0];
            const double v2 = L.volvol[y0] * L.volvol[y0], mu = L.voldrift[y0];
            const double det = em * ep * ep - ep * em * em;
            bm = (ep * ep * mu - ep * v2) / det;
            bp = (-em * em * mu + em * v2) / det;
        } else if (r0 == 0) {
            bp = L.voldrift[y0] / (L.rval[y0 + nx] - L.rval[y0]);
        } else {
            bm = L.voldrift[y0] / (L.rval[y0 - nx] - L.rval[y0]);
        }
        if (r0 < nr - 1 && bp > 0.0) {
            const double jsm = L.jm[y0] * X0, jsp = L.jp[y0] * X0;
            const double *xr = L.xval + nx * (r0 + 1);
            double s = 0.0;
            if (jsm > 0.0)
                for (int x1 = 0; x1 < x0 - 1; x1++) s += exp((xr[x1] - X0) / jsm);
            if (jsp > 0.0)
                for (int x1 = x0 + 1; x1 < nx; x1++) s += exp(-(xr[x1] - X0) / jsp);
            if (s > 0.0) {
                ratio = bp / s;
                sm = jsm > 0.0 ? jsm : 0.0;
                sp = jsp > 0.0 ? jsp : 0.0;
            }
        }
    }
    if (r0 == 0) bm = 0.0;       // no such neighbour
    if (r0 == nr - 1) bp = 0.0;
    if (x0 == 0) am = 0.0;
    if (x0 == nx - 1) ap = 0.0;
    // clip, then drift and mass over the row's entries in state order:
    // regime r0-1 (x0), regime r0 (x0-1, x0+1), regime r0+1 (jumps below
    // x0-1, x0, jumps above x0); zero entries add exact zeros
    am = am < 0.0 ? 0.0 : am;
    ap = ap < 0.0 ? 0.0 : ap;
    bm = bm < 0.0 ? 0.0 : bm;
    const double bp_kept = bp < 0.0 ? 0.0 : bp;
    double drift = 0.0, sum0 = 0.0;
    auto add = [&](double v, double x1) {
        drift += v * (x1 - X0);
        sum0 += v;
    };
    if (r0 > 0) add(bm, L.xval[y0 - nx]);
    if (x0 > 0) add(am, L.xval[y0 - 1]);
    if (x0 < nx - 1) add(ap, L.xval[y0 + 1]);
    if (r0 < nr - 1) {
        const double *xr = L.xval + nx * (r0 + 1);
        if (sm > 0.0)
            for (int x1 = 0; x1 < x0 - 1; x1++) add(ratio * exp((xr[x1] - X0) / sm), xr[x1]);
        add(bp_kept, xr[x0]);
        if (sp > 0.0)
            for (int x1 = x0 + 1; x1 < nx; x1++) add(ratio * exp(-(xr[x1] - X0) / sp), xr[x1]);
    }
    const double drift0 = L.sdrift[y0];
    if (drift > drift0) {
        if (x0 > 0) {
            const double c = (drift - drift0) / (X0 - L.xval[y0 - 1]);
            am += c;
            sum0 += c;
        }
    } else if (x0 < nx - 1) {
        const double c = (drift0 - drift) / (L.xval[y0 + 1] - X0);
        ap += c;
        sum0 += c;
    }
    double *o = rt + (size_t)kSVRow * y0;
    o[0] = am;
    o[1] = ap;
    o[2] = bm;
    o[3] = bp_kept;
    o[4] = ratio;
    o[5] = sm;
    o[6] = sp;
    o[7] = -sum0;
}

template <int MODE, typename OUT>
__global__ __launch_bounds__(256) void k_sv_fill(SVLattice L, const double *__restrict__ rt, double dt,
                                                 OUT *__restrict__ out) {
    // 32-bit index math: d <= 32767, so d * d < 2^31
    const unsigned nx = (unsigned)L.nx, d = nx * (unsigned)L.nr, n = d * d;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const unsigned y1 = e / d, y0 = e - y1 * d;
        const unsigned r0 = y0 / nx, x0 = y0 - r0 * nx, r1 = y1 / nx, x1 = y1 - r1 * nx;
        const double *o = rt + (size_t)kSVRow * y0;
        double g = 0.0;
        if (y1 == y0) g = o[7];
        else if (r1 == r0) g = x1 + 1 == x0 ? o[0] : (x1 == x0 + 1 ? o[1] : 0.0);
        else if (r1 + 1 == r0) g = x1 == x0 ? o[2] : 0.0;
        else if (r1 == r0 + 1) {
            if (x1 == x0) g = o[3];
            else if (x1 + 1 < x0 && o[5] > 0.0) g = o[4] * exp((L.xval[y1] - L.xval[y0]) / o[5]);
            else if (x1 > x0 && o[6] > 0.0) g = o[4] * exp(-(L.xval[y1] - L.xval[y0]) / o[6]);
        }
        double v;
        if (MODE == 0) v = g;
        else if (MODE == 1) v = (y0 == y1 ? 1.0 : 0.0) + dt * g;
        else v = dt * g;
        out[e] = (OUT)v;
    }
}

}  // namespace

// out_kind: 0 FP32 I + dt L, 1 FP64 I + dt L, 2 FP32 dt L, 3 FP64 dt L
// (as ophip_sek_rs), 4 the FP64 generator L.  Every array is a handle of
// d = nx*nr doubles indexed by state y = x + nx*r; nr >= 2 (the reference's
// 4x4 system is singular with one regime).
OPHIP_EXPORT int ophip_dgen_sv(int nx, int nr, unsigned xval, unsigned rval, unsigned sdrift, unsigned svol,
                               unsigned volvol, unsigned voldrift, unsigned jm, unsigned jp, double dt,
                               int out_kind, unsigned out) {
    if (nx < 3 || nr < 2 || (long)nx * nr > 32767 || out_kind < 0 || out_kind > 4)
        return (int)set_status(kErrBadArg);
    const size_t d = (size_t)nx * nr;
    const unsigned hs[8] = {xval, rval, sdrift, svol, volvol, voldrift, jm, jp};
    for (unsigned h : hs)
        if (!handle_range_ok(h, 8 * d)) return (int)set_status(kErrBadHandle);
    const size_t eb = (out_kind == 1 || out_kind >= 3) ? 8 : 4;
    if (!handle_range_ok(out, eb * d * d)) return (int)set_status(kErrBadHandle);
    SVLattice L{nx, nr, P<const double>(xval), P<const double>(rval), P<const double>(sdrift),
                P<const double>(svol), P<const double>(volvol), P<const double>(voldrift),
                P<const double>(jm), P<const double>(jp)};
    double *rt = (double *)workspace(sizeof(double) * kSVRow * d);
    if (!rt) return (int)kErrOutOfArena;
    Device &D = dev();
    hipLaunchKernelGGL(k_sv_rows, dim3(cdiv((int)d, 64)), dim3(64), 0, D.stream, L, rt);
    const dim3 g(grid1d(d * d, 256)), b(256);
    switch (out_kind) {
    case 0: hipLaunchKernelGGL((k_sv_fill<1, float>), g, b, 0, D.stream, L, (const double *)rt, dt, P<float>(out)); break;
    case 1: hipLaunchKernelGGL((k_sv_fill<1, double>), g, b, 0, D.stream, L, (const double *)rt, dt, P<double>(out)); break;
    case 2: hipLaunchKernelGGL((k_sv_fill<2, float>), g, b, 0, D.stream, L, (const double *)rt, dt, P<float>(out)); break;
    case 3: hipLaunchKernelGGL((k_sv_fill<2, double>), g, b, 0, D.stream, L, (const double *)rt, dt, P<double>(out)); break;
    default: hipLaunchKernelGGL((k_sv_fill<0, double>), g, b, 0, D.stream, L, (const double *)rt, dt, P<double>(out)); break;
    }
    return (int)check_launch();
}

namespace {

// the reference's FP32 parameter block of generator i to FP64: xval (d,
// shared), then SDrift, SVol, VolDrift, VolVol, Jumpsz_minus, Jumpsz_plus
// (ni*d each, generator i at offset i*d; sgsvg.cu:222-228); out: seven
// arrays of d doubles in the order xval, sdrift, svol, voldrift, volvol,
// jm, jp
__global__ void k_sv_pars(int d, int ni, int i, const float *__restrict__ pars, double *__restrict__ out) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= d) return;
    const size_t nd = (size_t)ni * d, gi = (size_t)i * d + y;
    out[y] = pars[y];
    out[(size_t)d + y] = pars[d + gi];                      // SDrift
    out[(size_t)2 * d + y] = pars[d + nd + gi];             // SVol
    out[(size_t)3 * d + y] = pars[d + 2 * nd + gi];         // VolDrift
    out[(size_t)4 * d + y] = pars[d + 3 * nd + gi];         // VolVol
    out[(size_t)5 * d + y] = pars[d + 4 * nd + gi];         // Jumpsz_minus
    out[(size_t)6 * d + y] = pars[d + 5 * nd + gi];         // Jumpsz_plus
}

}  // namespace

// The reference's device entry point for the stochastic-volatility
// generators (sgsvg.cu:215-236, bound by CSVModel.cs:383-384 and called
// by CSVModel.device_sgen, CSVModel.cs:417): ni FP32 generators
// gen[i*d*d + y0 + d*y1], d = nx*nr, from the caller's stencil inverses
// invm (16 floats per state) and the parameter block pars (d + 6*ni*d
// floats: xval, then per generator SDrift, SVol, VolDrift, VolVol,
// Jumpsz_minus, Jumpsz_plus).  Every entry is evaluated in FP64 by the
// kernels of ophip_dgen_sv (rates from invm, jumps into the next regime,
// clipping, drift fix, diagonal; with that generator's two fixes) and
// rounded once to FP32; the whole d x d block of each generator is
// written (the reference writes the nonzeros into a block its caller
// zeroed first, CSVModel.cs:414).  Bad sizes or handles set the status
// (the reference returns silently when d % 32 != 0, sgsvg.cu:218).
OPHIP_EXPORT void opcuda_sgsvg(unsigned gen, int ni, int nx, int nr, unsigned invm, unsigned pars) {
    const long dl = (long)nx * nr;
    if (nx < 3 || nr < 2 || ni <= 0 || dl > 32767 || (long)ni * dl * dl * 4 > 0xFFFFFFFFL) {
        set_status(kErrBadArg);
        return;
    }
    const size_t d = (size_t)dl;
    if (!handle_range_ok(gen, (size_t)4 * ni * d * d) || !handle_range_ok(invm, (size_t)64 * d) ||
        !handle_range_ok(pars, (size_t)4 * (d + 6 * (size_t)ni * d))) {
        set_status(kErrBadHandle);
        return;
    }
    double *ws = (double *)workspace(sizeof(double) * (7 + kSVRow) * d);
    if (!ws) return;
    double *rt = ws + 7 * d;
    Device &D = dev();
    SVLattice L{nx, nr, ws, nullptr, ws + d, ws + 2 * d, ws + 4 * d, ws + 3 * d, ws + 5 * d, ws + 6 * d,
                P<const float>(invm)};
    const dim3 g(grid1d(d * d, 256)), b(256);
    for (int i = 0; i < ni; i++) {
        hipLaunchKernelGGL(k_sv_pars, dim3(cdiv((int)d, 64)), dim3(64), 0, D.stream, (int)d, ni, i,
                           P<const float>(pars), ws);
        hipLaunchKernelGGL(k_sv_rows, dim3(cdiv((int)d, 64)), dim3(64), 0, D.stream, L, rt);
        hipLaunchKernelGGL((k_sv_fill<0, float>), g, b, 0, D.stream, L, (const double *)rt, 0.0,
                           P<float>(gen) + (size_t)i * d * d);
    }
    check_launch();
}
