// wsq.hip -- repeated squaring in the U = I + W form with per-step
// probability-mass renormalisation, on FP16/BF16 matrix cores with FP32
// accumulation (no reference counterpart: the reference squares U itself
// in FP32, ssqmm.cu:119-170).
//
// Why the W form.  U(dt) = I + W with |W| ~ |L| dt small.  In FP32 the
// diagonal 1 - eps of U keeps only ~2^-24/eps relative precision of the
// mass that leaves a state, and every squaring doubles the row-sum error;
// after 16 squarings and 252 backward steps the reference's FP32 method is
// off by ~10% at the 2048-state lattice.  Squaring W instead,
//     (I + W)^2 = I + (2W + W*W),
// keeps full relative precision of the small entries; the product W*W is
// the only GEMM and is rounded relative to |W|^2.  After each squaring the
// diagonal is corrected so every row of W sums to exactly the FP64 sum of
// its off-diagonal part with opposite sign (rows of U sum to 1).
//
// Operands.  W is kept in FP32 (column-major) only; a squaring splits it
// into 16-bit parts scaled by 2^S (S = 15 for FP16 so entries <= 1 stay
// normal down to 2^-24/2^15; S = 0 for BF16) as it stages its operand
// tiles into LDS: H = rnd(W 2^S) and, for the three-product variant,
// L = rnd((W 2^S - H) 2^G) (G = 11 FP16, 8 BF16).  The product is
//     W*W ~ 2^-2S (H*H + 2^-G (H*L + L*H)).
// Storing the parts would move three times the bytes of W per round (both
// layouts, both parts) for reads of the same size; splitting on the way in
// costs a few VALU operations per element instead.  A tiles are staged
// through registers as 4-row x 8-column blocks, so each row's 8
// consecutive k reach LDS as one store: both operands then feed the MFMA
// with 8 consecutive k per lane.
//
// Kernels (all tiles 128x128 or 64x64, every d a multiple of the tile):
//   k_wsplit   the first W: flush rule in place, partial row sums, flags
//   k_wsq      W' = 2W + W*W on v_mfma_f32_16x16x32_{f16,bf16}; epilogue
//              writes W' (FP32), its partial row sums and cell flags
//   k_wfix     diagonal correction from the partial row sums (fixed order)
// A scenario table lets one launch serve many lattices of different d.
#include "ophip_internal.h"
#include "rs_lattice.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

extern "C" unsigned opcuda_get_status(void);

using namespace ophip;

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));  // native vector: stays in VGPRs

namespace {

// One scenario (lattice) of a batched W-form squaring round.  Every array
// is a byte offset from the arena base, which the kernels receive as an
// argument: pointers derived from it stay in the global address space, so
// the compiler emits global (not flat) memory instructions and can overlap
// them with LDS traffic.
struct WScen {
    int d;
    int pad;
    uint64_t W, Wn;                // FP32 W (input), W' (output), column-major
    uint64_t part;                 // partial row sums of W': [d/64 column blocks][d]
    uint64_t fl, fln;              // cell flags of W, W' (below)
};

// Cell flags.  W is cut into 64 x 64 cells; flag byte [cr][cc] (d/64 x
// d/64, row-major) is 0 only if every value of cell (rows 64 cr..,
// columns 64 cc..) is +0 (so are its 16-bit parts), else 96 + E with the cell's
// largest |W| below 2^E (255: no bound known).  FP16 H and L flush entries
// |W| < 2^-51 to zero, and a banded generator keeps most cells there for
// many squarings, so the product's K loop visits only the cells k with
// A(rows, k) and B(k, columns) both flagged: the skipped MFMAs would add
// exact zeros, the result is the same bits.  Nonzero flags are only a "may
// be nonzero" (the diagonal fix flags its cells 255 unconditionally).
// Beyond the zero cells, a K cell whose two operand bounds multiply to at
// most 2^kcut (ophip_wsq_set_kcut; default 2^-36) is left out too: its
// product changes no entry of W by more than 64 2^kcut.  On the 2048-state
// daily workload that leaves out about a quarter of the K-cell products of
// the 16 squarings (half of the last rounds') and moves the prices by under
// 1e-8 relative in exact arithmetic; the error grows steeply only from
// 2^-28 (1.5e-4).
// Flush rule: an element whose 16-bit parts are both +-0 (|W| < 2^-51 for
// FP16 x3; H alone for x1) is stored as FP32 +0, so the parts the next
// squaring makes of a zero cell are zero.  A squaring item whose output cells are
// then certainly zero (no K cells, zero input cells) and whose destination
// cells already hold zeros (flags of the buffer's previous contents) writes
// nothing at all, and the diagonal fix skips the partial sums of zero cells.

// most 16-item chunks in one XCD's live list of a round
constexpr int kLiveChunks = 1024;

// work item = (scenario, tile)
struct WItem {
    int s, tm, tn, pad;
};

// a live work item of a sparse round: the item and its K cells (bit c =
// cell column c of A and cell row c of B both flagged)
struct WLive {
    int s, tm, tn, pad;
    uint64_t cells, pad2;
};

// the cell flag of a nonzero cell whose largest |W| is amax: 96 + E with
// amax < 2^E (frexp), kept in 1 .. 254
__device__ inline uint8_t cell_flag(float amax) {
    int e = 0;
    (void)frexpf(amax, &e);
    return (uint8_t)min(max(e + 96, 1), 254);
}
constexpr uint8_t kFlagAny = 255;  // may be nonzero, no bound known

template <class T> __device__ inline T *at(char *base, uint64_t off) {
    return reinterpret_cast<T *>(base + off);
}

template <bool BF> struct Num;
template <> struct Num<false> {
    static constexpr float S = 32768.0f, IS = 1.0f / 32768.0f;
    static constexpr float G = 2048.0f, IG = 1.0f / 2048.0f;
    __device__ static uint16_t enc(float x) {
        _Float16 h = (_Float16)x;
        return __builtin_bit_cast(uint16_t, h);
    }
    __device__ static float dec(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }
    __device__ static f4 mfma(s8 a, s8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                       __builtin_bit_cast(h8, b), c, 0, 0, 0);
    }
};
template <> struct Num<true> {
    static constexpr float S = 1.0f, IS = 1.0f;
    static constexpr float G = 256.0f, IG = 1.0f / 256.0f;
    __device__ static uint16_t enc(float x) {
        __bf16 h = (__bf16)x;
        return __builtin_bit_cast(uint16_t, h);
    }
    __device__ static float dec(uint16_t u) { return (float)__builtin_bit_cast(__bf16, u); }
    __device__ static f4 mfma(s8 a, s8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                        __builtin_bit_cast(b8, b), c, 0, 0, 0);
    }
};

// ---------------------------------------------------------------------
// k_wsplit: the first W of a solve, one 64x64 tile (one cell) per
// workgroup (64 threads, lane = row): the flush rule applied in place,
// the partial row sums and the cell's flag.  The squarings split W into
// its 16-bit parts as they stage their operands.
// ---------------------------------------------------------------------
template <bool BF, bool LO>
__global__ __launch_bounds__(64) void k_wsplit(char *__restrict__ base,
                                               const WScen *__restrict__ scen,
                                               const WItem *__restrict__ items) {
    using N = Num<BF>;
    const WItem it = items[blockIdx.x];
    const WScen sc = scen[it.s];
    const int d = sc.d, lane = threadIdx.x;
    const int i0 = it.tm * 64, j0 = it.tn * 64;
    const int i = i0 + lane;
    uint8_t *flag = at<uint8_t>(base, sc.fln) + (size_t)it.tm * (d / 64) + it.tn;
    double rs = 0.0;
    unsigned nz = 0;
    float amax = 0.0f;  // the lane's largest |W| after the flush
    float *Wm = at<float>(base, sc.W);  // W0 itself (flushed in place)
    // all 64 loads issued before any store (the stores may alias W)
    float wrow[64];
#pragma unroll
    for (int c = 0; c < 64; c++) wrow[c] = Wm[(size_t)i + (size_t)d * (j0 + c)];
#pragma unroll
    for (int c = 0; c < 64; c++) {
        float w = wrow[c];
        const float ws = w * N::S;
        const uint16_t h = N::enc(ws);
        const uint16_t l = LO ? N::enc((ws - N::dec(h)) * N::G) : (uint16_t)0;
        const unsigned nzv = (h | l) & 0x7fffu;
        if (!nzv) {  // flush rule: both 16-bit parts +-0 -> FP32 +0
            if (__builtin_bit_cast(uint32_t, w) != 0u) Wm[(size_t)i + (size_t)d * (j0 + c)] = 0.0f;
            w = 0.0f;
        }
        rs += (double)w;
        nz |= nzv;
        amax = fmaxf(amax, fabsf(w));
    }
    at<double>(base, sc.part)[(size_t)it.tn * d + i] = rs;  // slot = 64-column block
    const bool any = __ballot(nz != 0) != 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
    if (lane == 0) *flag = any ? cell_flag(amax) : 0;
}

// ---------------------------------------------------------------------
// k_wsq: W' = 2W + W*W.  Workgroup tile BT x BT (BT = 128: 4 waves of
// 64x64; BT = 64: 4 waves of 32x32), K-step 32, four K-steps of global
// loads in flight in named register sets (no scratch), LDS double buffer.
// Requires d % 128 == 0 (whole tiles, K-steps a multiple of 4).
// ---------------------------------------------------------------------
template <bool BF, bool LO, int BT, int NW>
__global__ __launch_bounds__(64 * NW) void k_wsq(char *__restrict__ base,
                                             const WScen *__restrict__ scen,
                                             const WItem *__restrict__ items, int nitems,
                                             const WLive *__restrict__ live,
                                             const int *__restrict__ counts, int cap) {
    // Work: the static list items[0..nitems) (every K-step of every item),
    // or, when live is given, the live items of the round (k_wlive): list x
    // = live + x cap, in chunks of 16 slots of which counts[x cap / 16 + c]
    // are used, is taken in its static order by the workgroups on XCD x
    // (b % 8 == x; the grid is a multiple of 8), with each item's K cells:
    // the workgroups of an XCD then work on neighbouring tiles of one
    // lattice at a time, whose operand panels its L2 shares.
    using N = Num<BF>;
    // K-step BK = 64 for 128-tiles (two MFMA k-steps per LDS stage, two
    // register stages in flight), 32 for 64-tiles (four stages).  LDS rows
    // hold BK 16-bit elements, unpadded; the 16-byte chunk c of row R sits at
    // c ^ swz(R) -- (R >> 1) & 3 for 4 chunks, R & 7 for 8 -- so that the
    // stores (8 lanes = 8 consecutive chunks) and the fragment reads (8
    // lanes = 8 rows, one chunk) each touch 8 distinct 16-byte bank groups
    constexpr int BK = BT == 128 ? 64 : 32, PITCH = BK, CPR = BK / 8;
    // NW waves as 2 (rows) x NW/2 (columns); wave tile WTM x WTN of TM x TN MFMA tiles
    constexpr int WGN = NW / 2, WTM = BT / 2, WTN = BT / WGN, TM = WTM / 16, TN = WTN / 16;
    constexpr int NOP = LO ? 2 : 1;
    constexpr int PF = BK == 64 ? 2 : 4;     // K-steps of loads in flight
    constexpr int SPC = 64 / BK, TPC = BT / 64;  // K-steps per cell, cells per tile side
    __shared__ __attribute__((aligned(16))) uint16_t As[2][NOP][BT * PITCH];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NOP][BT * PITCH];
    __shared__ double rsum[WGN][BT];
    __shared__ int wnz[NW];
    __shared__ float wmx[NW];  // each wave's largest |W'| (its block lies in one cell)
    __shared__ int pref[kLiveChunks + 1];  // live items before chunk c of this XCD's list

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int wm0 = (wave & 1) * WTM, wn0 = (wave >> 1) * WTN;

    const WLive *list = live ? live + (size_t)(blockIdx.x & 7) * cap : nullptr;
    int nlist = nitems, nch = 0;
    if (live) {
        // the chunk counts of this XCD chained in order (wave 0, 64 at a time)
        nch = ((nitems + 7) / 8 + 15) / 16;  // the chunks k_wlive covered
        const int *cnt = counts + (blockIdx.x & 7) * (cap / 16);
        if (wave == 0) {
            int run = 0;
            if (lane == 0) pref[0] = 0;
            for (int c0 = 0; c0 < nch; c0 += 64) {
                int v = c0 + lane < nch ? cnt[c0 + lane] : 0;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int u = __shfl_up(v, o, 64);
                    if (lane >= o) v += u;
                }
                if (c0 + lane < nch) pref[c0 + lane + 1] = run + v;
                run += __shfl(v, 63, 64);
            }
        }
        __syncthreads();
        nlist = pref[nch];
    }
    const int first = live ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const int stride = live ? (int)(gridDim.x >> 3) : (int)gridDim.x;
    auto fetch = [&](int item, WItem &it, uint64_t &cells) {
        if (live) {
            // the chunk holding list entry `item`: the last c with pref[c] <= item
            int lo = 0, hi = nch - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] <= item) lo = mid;
                else hi = mid - 1;
            }
            const WLive lv = list[16 * lo + (item - pref[lo])];
            it = WItem{lv.s, lv.tm, lv.tn, 0};
            cells = lv.cells;
        } else {
            it = items[item];
            cells = 0;
        }
    };
    // Operands come from the FP32 W (column-major) and are split into
    // their 16-bit hi/lo parts on the way to LDS, with the same operations
    // as the epilogue's flush test and the backward layout (the same bits).
    // Waves 0 .. NW/2-1 stage A, the others B, NF float4 loads per thread
    // and stage.  A: thread (rg, cg) loads a 4-row x KA-column block (rows
    // 4 rg.., k = KA cg ..; eight 128-byte column pieces per load
    // instruction) and writes each row's KA consecutive k as one LDS store.
    // B: thread (kq, jb) loads 4 consecutive k (4 kq ..) of columns jb + 16 u.
    constexpr int HALF = 32 * NW, NF = BT * BK / HALF / 4, KA = NF, CG = BK / KA, KQ = BK / 4;
    static_assert(BT / 4 * CG == HALF && HALF / KQ == 16 && 16 * NF == BT, "operand staging");
    const bool isA = wave < NW / 2;  // wave-uniform
    const int ht = t % HALF;
    const int rg = ht / CG, cg = ht % CG, kq = ht % KQ, jb = ht / KQ;
    typedef f4 Stage[NF];
    Stage s0, s1, s2, s3;
    // The load side: the item whose operand loads are being issued (an
    // item's first PF K-steps go out during the previous item's epilogue).
    // Its K-steps: every K-step (static list) or every K-step of its K
    // cells (wave-uniform: scalar registers); the cursor gives the K offset
    // of the next step to load and, past the last step, repeats the last
    // one (loads stay unconditional).
    const bool dense = live == nullptr;
    const float *ga = nullptr, *gb = nullptr;
    int ld = 0, lns = 0, lq = 0, lh = 0, klast = 0;
    uint64_t lm = 0;
    auto gload = [&](Stage &x, int k0) {
        if (isA) {
#pragma unroll
            for (int u = 0; u < NF; u++) x[u] = *(const f4 *)(ga + (size_t)ld * (k0 + u));
        } else {
#pragma unroll
            for (int u = 0; u < NF; u++) x[u] = *(const f4 *)(gb + k0 + (size_t)ld * (16 * u));
        }
    };
    auto next_k = [&]() {
        int k0;
        if (dense) {
            k0 = min(lq, lns - 1) * BK;
            lq++;
        } else if (lm) {
            k0 = 64 * __builtin_ctzll(lm) + BK * lh;
            if (SPC == 1 || lh) {
                lm &= lm - 1;
                lh = 0;
            } else {
                lh = 1;
            }
            klast = k0;
        } else {
            k0 = klast;
        }
        return k0;
    };
    // set the load side to an item and issue its first PF K-steps; returns
    // its K-step count
    auto begin = [&](const WItem &bi, uint64_t bcells, const WScen &bsc) {
        ld = bsc.d;
        const float *Wsrc = at<const float>(base, bsc.W);
        ga = Wsrc + (size_t)(bi.tm * BT + 4 * rg) + (size_t)ld * (KA * cg);
        gb = Wsrc + (size_t)(4 * kq) + (size_t)ld * (bi.tn * BT + jb);
        lm = bcells;
        lq = lh = klast = 0;
        lns = dense ? ld / BK : __popcll(bcells) * SPC;
        if (lns > 0) {
            gload(s0, next_k());
            gload(s1, next_k());
            if (PF == 4) {
                gload(s2, next_k());
                gload(s3, next_k());
            }
        }
        return lns;
    };
    WItem it{};
    WScen sc{};
    int ns = 0;
    if (first < nlist) {
        uint64_t c0 = 0;
        fetch(first, it, c0);
        sc = scen[it.s];
        ns = begin(it, c0, sc);
    }
    for (int item = first; item < nlist; item += stride) {
        const int d = sc.d;
        const int i0 = it.tm * BT, j0 = it.tn * BT;

        // the epilogue's W values, loaded before the K loop's first loads:
        // their latency hides behind the loop, and the epilogue issues
        // stores only (a load there would wait for every earlier store)
        f4 wv[TM][TN];
        {
            const float *W = at<const float>(base, sc.W);
#pragma unroll
            for (int a = 0; a < TM; a++)
#pragma unroll
                for (int b = 0; b < TN; b++)
                    wv[a][b] = *(const f4 *)(W + (size_t)(i0 + wm0 + 16 * a + 4 * g) +
                                             (size_t)d * (j0 + wn0 + 16 * b + r));
        }

        f4 acc[TM][TN], cor[TM][TN];
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
            for (int b = 0; b < TN; b++) {
                acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
                cor[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            }

        auto swz = [](int row) { return CPR == 8 ? (row & 7) : ((row >> 1) & 3); };
        auto split2 = [](float a, float b, uint32_t &hi, uint32_t &lo) {
            const float as = a * N::S, bs = b * N::S;
            const uint16_t ha = N::enc(as), hb = N::enc(bs);
            hi = (uint32_t)ha | ((uint32_t)hb << 16);
            lo = 0;
            if (LO)
                lo = (uint32_t)N::enc((as - N::dec(ha)) * N::G) |
                     ((uint32_t)N::enc((bs - N::dec(hb)) * N::G) << 16);
        };
        auto sstore = [&](const Stage &x, int buf) {
            if (isA) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = 4 * rg + v, kk = KA * cg;
                    const int sw = row * PITCH + 8 * ((kk >> 3) ^ swz(row)) + (kk & 7);
                    uint32_t h[KA / 2], l[KA / 2];
#pragma unroll
                    for (int u = 0; u < KA / 2; u++) split2(x[2 * u][v], x[2 * u + 1][v], h[u], l[u]);
                    if constexpr (KA == 8) {
                        *(u4v *)&As[buf][0][sw] = u4v{h[0], h[1], h[2], h[3]};
                        if (LO) *(u4v *)&As[buf][NOP - 1][sw] = u4v{l[0], l[1], l[2], l[3]};
                    } else {
                        *(uint2 *)&As[buf][0][sw] = uint2{h[0], h[1]};
                        if (LO) *(uint2 *)&As[buf][NOP - 1][sw] = uint2{l[0], l[1]};
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < NF; u++) {
                    const int col = jb + 16 * u, kk = 4 * kq;
                    const int sw = col * PITCH + 8 * ((kk >> 3) ^ swz(col)) + (kk & 7);
                    uint32_t h0, h1, l0, l1;
                    split2(x[u][0], x[u][1], h0, l0);
                    split2(x[u][2], x[u][3], h1, l1);
                    *(uint2 *)&Bs[buf][0][sw] = uint2{h0, h1};
                    if (LO) *(uint2 *)&Bs[buf][NOP - 1][sw] = uint2{l0, l1};
                }
            }
        };
        auto compute = [&](int buf) {
#pragma unroll
          for (int ks = 0; ks < BK / 32; ks++) {
            s8 fa[NOP][TM], fb[NOP][TN];
#pragma unroll
            for (int o = 0; o < NOP; o++)
#pragma unroll
                for (int a = 0; a < (TM > TN ? TM : TN); a++) {
                    // wm0, wn0 and 16a are multiples of 16: the row's swizzle is swz(r)
                    const int cg = 8 * ((4 * ks + g) ^ swz(r));
                    if (a < TM) fa[o][a] = *(const s8 *)&As[buf][o][(wm0 + 16 * a + r) * PITCH + cg];
                    if (a < TN) fb[o][a] = *(const s8 *)&Bs[buf][o][(wn0 + 16 * a + r) * PITCH + cg];
                }
#pragma unroll
            for (int b = 0; b < TN; b++)
#pragma unroll
                for (int a = 0; a < TM; a++) acc[a][b] = N::mfma(fa[0][a], fb[0][b], acc[a][b]);
            if (LO) {
#pragma unroll
                for (int b = 0; b < TN; b++)
#pragma unroll
                    for (int a = 0; a < TM; a++) cor[a][b] = N::mfma(fa[0][a], fb[1][b], cor[a][b]);
#pragma unroll
                for (int b = 0; b < TN; b++)
#pragma unroll
                    for (int a = 0; a < TM; a++) cor[a][b] = N::mfma(fa[1][a], fb[0][b], cor[a][b]);
            }
          }
        };

        const int nc = d / 64;
        if (ns > 0) {  // the first PF K-steps' loads are in flight (begin)
            // step kt: its registers go to LDS buffer kt&1 (last read by
            // compute(kt-2), finished by every thread before the barrier of
            // step kt-1), the registers are refilled with step kt+PF, then
            // compute.  ns is workgroup-uniform, so are the tail conditions.
#define OPHIP_WSQ_STAGE(S, KT)                                 \
    if ((KT) < ns) {                                           \
        sstore(S, (KT) & 1);                                   \
        __syncthreads();                                       \
        gload(S, next_k());                                    \
        compute((KT) & 1);                                     \
    }
            for (int kt = 0; kt < ns; kt += PF) {
                OPHIP_WSQ_STAGE(s0, kt)
                OPHIP_WSQ_STAGE(s1, kt + 1)
                if (PF == 4) {
                    OPHIP_WSQ_STAGE(s2, kt + 2)
                    OPHIP_WSQ_STAGE(s3, kt + 3)
                }
            }
#undef OPHIP_WSQ_STAGE
        }
        __syncthreads();  // all LDS reads done before the next item's stores

        // the next item's first K-steps go out now, behind this epilogue
        const int nxt = item + stride;
        WItem itn = it;
        WScen scn = sc;
        int nsn = 0;
        if (nxt < nlist) {  // workgroup-uniform
            uint64_t cn = 0;
            fetch(nxt, itn, cn);
            scn = scen[itn.s];
            nsn = begin(itn, cn, scn);
        }

        // epilogue: lane holds rows 4g+v of column r of every 16x16 tile
        float *Wn = at<float>(base, sc.Wn);
        const float inv = N::IS * N::IS;
        double rows[TM][4];
        unsigned nzl = 0;  // any nonzero 16-bit part among this lane's elements
        float amax = 0.0f;  // the lane's largest |W'| after the flush
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
            for (int v = 0; v < 4; v++) rows[a][v] = 0.0;
#pragma unroll
        for (int a = 0; a < TM; a++) {
            const int row = i0 + wm0 + 16 * a + 4 * g;
#pragma unroll
            for (int b = 0; b < TN; b++) {
                const int col = j0 + wn0 + 16 * b + r;
                const size_t e = (size_t)row + (size_t)d * col;
                const f4 w = wv[a][b];
                f4 p = acc[a][b];
                if (LO) p += cor[a][b] * N::IG;
                f4 o = 2.0f * w + p * inv;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    // the split of the stored FP32 value decides the flush:
                    // no fused multiply-add-convert, whose single rounding
                    // differs at FP16 ties
                    float ov = o[v];
                    asm volatile("" : "+v"(ov));
                    const float os = ov * N::S;
                    const uint16_t hc = N::enc(os);
                    const uint16_t lc = LO ? N::enc((os - N::dec(hc)) * N::G) : (uint16_t)0;
                    const unsigned nzv = (hc | lc) & 0x7fffu;
                    if (!nzv) o[v] = 0.0f;  // flush rule: both 16-bit parts +-0 -> FP32 +0
                    nzl |= nzv;
                    amax = fmaxf(amax, fabsf(o[v]));
                    rows[a][v] += (double)o[v];
                }
                *(f4 *)(Wn + e) = o;
            }
        }
        // row sums over this wave's WTN columns: reduce across the 16 lanes r
#pragma unroll
        for (int a = 0; a < TM; a++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double x = rows[a][v];
                x += __shfl_xor(x, 1, 64);
                x += __shfl_xor(x, 2, 64);
                x += __shfl_xor(x, 4, 64);
                x += __shfl_xor(x, 8, 64);
                rows[a][v] = x;
            }
        if (r == 0) {
#pragma unroll
            for (int a = 0; a < TM; a++)
#pragma unroll
                for (int v = 0; v < 4; v++) rsum[wave >> 1][wm0 + 16 * a + 4 * g + v] = rows[a][v];
        }
        {
            const bool wz = __ballot(nzl != 0) != 0;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
            if (lane == 0) {
                wnz[wave] = wz ? 1 : 0;
                wmx[wave] = amax;
            }
        }
        __syncthreads();
        // flags of the tile's cells: a wave's WTM x WTN block lies in one cell
        if (t < TPC * TPC) {
            const int ca = t / TPC, cb = t % TPC;
            int f = 0;
            float m = 0.0f;
#pragma unroll
            for (int w = 0; w < NW; w++)
                if ((w & 1) * WTM / 64 == ca && (w >> 1) * WTN / 64 == cb) {
                    f |= wnz[w];
                    m = fmaxf(m, wmx[w]);
                }
            at<uint8_t>(base, sc.fln)[(size_t)(it.tm * TPC + ca) * nc + it.tn * TPC + cb] =
                f ? cell_flag(m) : (uint8_t)0;
        }
        // one partial slot per 64-column block, the sum of its CPS wave
        // columns in order (CPS = 64 / WTN)
        double *part = at<double>(base, sc.part);
        if (t < BT) {
            constexpr int CPS = 64 / WTN;
            const int row = i0 + t;
#pragma unroll
            for (int h = 0; h < BT / 64; h++) {
                double x = rsum[h * CPS][t];
#pragma unroll
                for (int c = 1; c < CPS; c++) x += rsum[h * CPS + c][t];
                part[(size_t)((BT / 64) * it.tn + h) * d + row] = x;
            }
        }
        it = itn;
        sc = scn;
        ns = nsn;
    }
}

// Live items of a sparse squaring round, one wavefront per work item of
// the round's static list (lane c: cell column c of A / cell row c of B).
// An item is live if it has K cells, if its own input cells are nonzero
// (W' = 2W + ...), when wskip is 0, or if its destination cells are
// flagged (they must be overwritten with zeros).  Both buffers' flags start
// as all ones when a plan is made (k_fill8) and the split and every round
// keep them a superset of their buffers' nonzero cells, so a lattice's
// first round (and split) of a later solve writes only where the previous
// solve left values.
// Live items are appended to the list of their XCD (static position j % 8:
// the xcd_order placement); the list order within an XCD does not change
// any result.  Requires d <= 4096 (64 cells).
template <int TPC>
__global__ __launch_bounds__(1024) void k_wlive(char *__restrict__ base, const WScen *__restrict__ scen,
                                                const WItem *__restrict__ items, int nitems,
                                                WLive *__restrict__ live, int cap, int *__restrict__ counts,
                                                int wskip, int kthr) {
    // block (x, chunk): the 16 items j = x + 8 (16 chunk + w) of XCD x, one
    // per wave w; the live ones go, in w order, to slots 16 chunk .. of
    // list x, and their number to counts[x cap / 16 + chunk], so the list
    // keeps the static order (the squaring kernel chains the counts)
    __shared__ int slive[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = blockIdx.x & 7, chunk = (int)(blockIdx.x >> 3);
    const int j = x + 8 * (16 * chunk + w);
    bool ok = false;
    uint64_t cells = 0;
    WItem it{};
    if (j < nitems) {  // wave-uniform
        it = items[j];
        const WScen sc = scen[it.s];
        const int nc = sc.d / 64;
        const uint8_t *fl = at<const uint8_t>(base, sc.fl);
        int fa = 0, fb = 0;  // the largest flags of the A cells (tile rows, k) and B cells (k, tile columns)
        if (lane < nc) {
#pragma unroll
            for (int q = 0; q < TPC; q++) {
                fa = max(fa, (int)fl[(size_t)(it.tm * TPC + q) * nc + lane]);
                fb = max(fb, (int)fl[(size_t)lane * nc + it.tn * TPC + q]);
            }
        }
        // K cell `lane`: both operands may be nonzero and their bounds'
        // product 2^(fa + fb - 192) exceeds 2^kcut (kthr = kcut + 193, or 0)
        cells = __ballot(fa != 0 && fb != 0 && fa + fb >= kthr);
        bool need = (lane >= it.tn * TPC && lane < it.tn * TPC + TPC && fa != 0) || !wskip;
        if (lane < TPC * TPC)
            need = need || at<const uint8_t>(base, sc.fln)[(size_t)(it.tm * TPC + lane / TPC) * nc +
                                                          it.tn * TPC + lane % TPC] != 0;
        ok = cells != 0 || __ballot(need) != 0;
    }
    if (lane == 0) slive[w] = ok ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0;
        for (int q = 0; q < 16; q++) n += slive[q];
        counts[x * (cap / 16) + chunk] = n;
    }
    if (ok && lane == 0) {
        int rank = 0;
        for (int q = 0; q < w; q++) rank += slive[q];
        live[(size_t)x * cap + 16 * chunk + rank] = WLive{it.s, it.tm, it.tn, 0, cells, 0};
    }
}

// n bytes at p set to v (cell flags)
__global__ __launch_bounds__(256) void k_fill8(uint8_t *__restrict__ p, int n, uint8_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// diagonal correction: W'_ii -= sum_j W'_ij (fixed summation order).
// One wavefront per 64 rows; the lanes of a row group read the d/64
// partial slots of their rows in a fixed order.
__global__ void k_wfix(char *__restrict__ base, const WScen *__restrict__ scen, int nscen,
                       const int *__restrict__ nsq, int round) {
    const int s = blockIdx.y;
    if (s >= nscen) return;
    if (nsq && round >= nsq[s]) return;  // this lattice has finished its squarings
    const WScen sc = scen[s];
    const int d = sc.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    const double *part = at<const double>(base, sc.part);
    const int nslots = d / 64;  // a multiple of 2 (d % 128 == 0)
    // slots of zero cells (flag 0) hold no partial sum of this round (the
    // squaring may have skipped them: stale values, selected away) and add 0
    const uint8_t *fl = at<const uint8_t>(base, sc.fln) + (size_t)(i >> 6) * nslots;
    const size_t e = (size_t)i + (size_t)d * i;
    float *Wn = at<float>(base, sc.Wn);
    const float w0 = Wn[e];  // issued with the partial sums' loads
    double rs = 0.0;
    // loads issued sixteen at a time (every slot of d <= 1024 in one
    // round trip), added in slot order; nslots is even
    for (int q = 0; q < nslots; q += 16) {
        double v[16];
        uint8_t f[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int k = q + u < nslots ? q + u : nslots - 1;  // clamped: never added
            v[u] = part[(size_t)k * d + i];
            f[u] = fl[k];
        }
#pragma unroll
        for (int u = 0; u < 16; u++)
            if (q + u < nslots) rs += f[u] ? v[u] : 0.0;
    }
    Wn[e] = (float)((double)w0 - rs);
    // the diagonal cells may have become nonzero
    if ((i & 63) == 0) at<uint8_t>(base, sc.fln)[(size_t)(i >> 6) * (d / 64) + (i >> 6)] = kFlagAny;
}

// ---------------------------------------------------------------------
// Backward induction in the W form on the matrix cores:
//     Y = df (X + W X),  X, Y: d x ncp row-major FP32.
// W enters through its 16-bit hi/lo copies (H = rnd(W 2^S), L = rnd((W 2^S
// - H) 2^G)), X through its own (Xh = rnd(X), Xl = rnd((X - Xh) 2^G)):
//     W X ~ 2^-S (H Xh + 2^-G (H Xl + L Xh)).
// Operands are stored fragment-major: a matrix M (rows i, reduction index
// k) is cut into 16 x 32 tiles (it, kt), tile-row-major, and within a tile
// lane l = 16 g + r holds M[16 it + r][32 kt + 8 g .. +7], so one wave's
// MFMA operand is one contiguous kilobyte.  W's copies are put in that
// layout once per solve (k_wfrag); each step writes the next step's X
// copies in it.
// ---------------------------------------------------------------------
__device__ inline size_t frag_off(int it, int kt, int l, int nkt) {
    return ((size_t)(it * nkt + kt) * 64 + l) * 8;
}

// value X[row][col] from its fragment-major hi/lo copies (column tile col /
// 16, k tile row / 32, lane 16 (row % 32 / 8) + col % 16, element row % 8;
// lo copies n elements on): hi + lo 2^-G in one FP32 rounding, the split's
// own reconstruction (FP16: 22 significant bits, BF16: 16)
template <class N>
__device__ inline float x_from_copies(const uint16_t *xt, size_t n, int row, int col, int nkt) {
    const int kk = row & 31;
    const size_t o = frag_off(col >> 4, row >> 5, 16 * (kk >> 3) + (col & 15), nkt) + (kk & 7);
    return N::dec(xt[o]) + N::dec(xt[o + n]) * N::IG;
}

// the 16-bit hi/lo parts of eight FP32 values, as the squarings split W
template <bool BF>
__device__ inline void split8(const float (&w)[8], uint4 &hi, uint4 &lo) {
    using N = Num<BF>;
    uint16_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const float ws = w[e] * N::S;
        h[e] = N::enc(ws);
        l[e] = N::enc((ws - N::dec(h[e])) * N::G);
    }
    hi = *(const uint4 *)h;
    lo = *(const uint4 *)l;
}

// FP32 W (d x d column-major) -> its hi/lo parts, fragment-major; one
// thread per 16-byte fragment (8 consecutive columns of one row)
template <bool BF>
__global__ __launch_bounds__(256) void k_wfrag(int d, const float *__restrict__ W, uint16_t *__restrict__ Hf,
                                               uint16_t *__restrict__ Lf) {
    const int nkt = d / 32;
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // fragment index
    if (f >= (size_t)d * d / 8) return;
    const int l = (int)(f & 63);
    const size_t tile = f >> 6;
    const int it = (int)(tile / nkt), kt = (int)(tile % nkt);
    const int r = l & 15, g = l >> 4;
    const float *src = W + (size_t)(16 * it + r) + (size_t)d * (32 * kt + 8 * g);
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; e++) w[e] = src[(size_t)d * e];
    uint4 hi, lo;
    split8<BF>(w, hi, lo);
    *(uint4 *)(Hf + f * 8) = hi;
    *(uint4 *)(Lf + f * 8) = lo;
}

// the same in another state order: fragment rows/columns are states i of
// the output order, read from W's state perm[i] (a permutation)
template <bool BF>
__global__ __launch_bounds__(256) void k_wfragp(int d, const float *__restrict__ W,
                                                const int *__restrict__ perm,
                                                uint16_t *__restrict__ Hf, uint16_t *__restrict__ Lf) {
    const int nkt = d / 32;
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)d * d / 8) return;
    const int l = (int)(f & 63);
    const size_t tile = f >> 6;
    const int it = (int)(tile / nkt), kt = (int)(tile % nkt);
    const int r = l & 15, g = l >> 4;
    const float *row = W + perm[16 * it + r];
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; e++) w[e] = row[(size_t)d * perm[32 * kt + 8 * g + e]];
    uint4 hi, lo;
    split8<BF>(w, hi, lo);
    *(uint4 *)(Hf + f * 8) = hi;
    *(uint4 *)(Lf + f * 8) = lo;
}

// one lattice of a batched fragment layout and its tile lists (arena
// offsets): FP32 W in, fragment copies, flags and lists out
struct LayScen {
    uint64_t W, Hf, Lf, flags, tl;
    uint64_t cf;  // W's cell flags (0 = every value of a 64 x 64 cell is +0)
};
constexpr int kLayBatch = 32;
struct LayBatch {
    LayScen s[kLayBatch];
};

// The same for the regime permutation of a lattice squared node-major
// (W's state nr x + r) stepped in the reference order (fragment state
// x + nx r), through LDS: a workgroup reads one 64 x 64 block of W
// (64/nr nodes x nr regimes per side; 16-byte loads down its columns),
// keeps its hi/lo parts column by column in LDS and writes every 8-column
// run of its fragment rows as one 16-byte store.  nr in {1, 2, 4, 8},
// nx % 8 == 0, d = nx nr, d % 64 == 0.  Lattice blockIdx.y of the batch.
template <bool BF>
__global__ __launch_bounds__(256) void k_wfragk(int d, int nx, int nr, char *__restrict__ base,
                                                const LayBatch bt) {
    using N = Num<BF>;
    const LayScen &S = bt.s[blockIdx.y];
    const float *__restrict__ W = at<const float>(base, S.W);
    uint16_t *__restrict__ Hf = at<uint16_t>(base, S.Hf);
    uint16_t *__restrict__ Lf = at<uint16_t>(base, S.Lf);
    __shared__ __attribute__((aligned(16))) uint16_t sh[2][64][68];  // [hi/lo][column][row]
    const int nb = d / 64, bi = blockIdx.x / nb, bj = blockIdx.x % nb, t = threadIdx.x;
    const int nkt = d / 32, NB = 64 / nr;
    // a cell flagged zero holds +0 only: its values are not read
    const bool zero = at<const uint8_t>(base, S.cf)[(size_t)bi * nb + bj] == 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int q = t + 256 * u, col = q >> 4, rq = q & 15;  // 4 rows of one column per thread
        const f4 w = zero ? f4{0.f, 0.f, 0.f, 0.f}
                          : *(const f4 *)(W + (size_t)(64 * bi + 4 * rq) + (size_t)d * (64 * bj + col));
        uint16_t h[4], l[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float ws = w[e] * N::S;
            h[e] = N::enc(ws);
            l[e] = N::enc((ws - N::dec(h[e])) * N::G);
        }
        *(uint2 *)&sh[0][col][4 * rq] = *(const uint2 *)h;
        *(uint2 *)&sh[1][col][4 * rq] = *(const uint2 *)l;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int q = t + 256 * u;
        const int xl = q % NB, r = (q / NB) % nr, rest = q / (NB * nr);  // rest < 8
        const int rp = rest / (NB / 8), c = rest % (NB / 8);
        const int rho = nr * xl + r;
        const int orow = NB * bi + xl + nx * r, ocol = NB * bj + 8 * c + nx * rp;
        const size_t o = frag_off(orow / 16, ocol / 32, 16 * ((ocol % 32) / 8) + orow % 16, nkt);
        uint16_t h[8], l[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            h[b] = sh[0][nr * (8 * c + b) + rp][rho];
            l[b] = sh[1][nr * (8 * c + b) + rp][rho];
        }
        *(uint4 *)(Hf + o) = *(const uint4 *)h;
        *(uint4 *)(Lf + o) = *(const uint4 *)l;
    }
}

// X (d x 16NT row-major FP32) -> its hi/lo copies, fragment-major with the
// columns as rows: thread per (column, 8 consecutive k)
template <bool BF, int NT>
__global__ __launch_bounds__(256) void k_wbwd_prep(int d, const float *__restrict__ X,
                                                   uint16_t *__restrict__ XhT,
                                                   uint16_t *__restrict__ XlT) {
    using N = Num<BF>;
    constexpr int NC = 16 * NT;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (k8 group, column)
    if (i >= d / 8 * NC) return;
    const int c = i % NC, k8 = i / NC;
    const int k0 = 8 * k8, kt = k0 / 32, g = (k0 % 32) / 8;
    uint16_t h[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const float x = X[(size_t)(k0 + e) * NC + c];
        h[e] = N::enc(x);
        lo[e] = N::enc((x - N::dec(h[e])) * N::G);
    }
    const size_t o = frag_off(c / 16, kt, 16 * g + (c % 16), d / 32);
    *(uint4 *)(XhT + o) = *(const uint4 *)h;
    *(uint4 *)(XlT + o) = *(const uint4 *)lo;
}

// One kernel per step, no split-K: workgroup = 16 rows x all of K, NW
// waves each taking K/NW (groups of two k32 blocks, two groups of loads in
// flight), the waves' tiles added in wave order through LDS, then Y =
// df (X + W X) and Y's hi/lo copies for the next step.  d / 16 workgroups.
template <bool BF, int NT, int NW>
__global__ __launch_bounds__(64 * NW) void k_wbwd1(int d, float df, const uint16_t *__restrict__ Hf,
                                                   const uint16_t *__restrict__ Lf,
                                                   const uint16_t *__restrict__ XhT,
                                                   const uint16_t *__restrict__ XlT,
                                                   const float *__restrict__ X, float *__restrict__ Y,
                                                   uint16_t *__restrict__ YhT,
                                                   uint16_t *__restrict__ YlT) {
    using N = Num<BF>;
    constexpr int NC = 16 * NT, GB = 2, NOP = 2 + 2 * NT;
    __shared__ f4 red[NW][NT][64];
    __shared__ uint16_t th[NC][16], tl[NC][16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int it = blockIdx.x, row0 = 16 * it;
    const int nkt = d / 32, nblk = nkt / NW, ngrp = (nblk + GB - 1) / GB;
    const int kt0 = wave * nblk;
    // fragment of tile (it, kt) for this lane: base + kt * 512
    const uint16_t *ha = Hf + frag_off(it, 0, lane, nkt);
    const uint16_t *la = Lf + frag_off(it, 0, lane, nkt);
    const uint16_t *xh = XhT + frag_off(0, 0, lane, nkt);
    const uint16_t *xl = XlT + frag_off(0, 0, lane, nkt);
    const size_t ctile = (size_t)nkt * 512;  // elements per 16-column strip
    typedef s8 Stage[GB][NOP];
    auto load = [&](Stage &S, int grp) {
#pragma unroll
        for (int b = 0; b < GB; b++) {
            const size_t k = (size_t)512 * (kt0 + min(grp * GB + b, nblk - 1));  // clamped: unused
            S[b][0] = *(const s8 *)(ha + k);
            S[b][1] = *(const s8 *)(la + k);
#pragma unroll
            for (int j = 0; j < NT; j++) {
                S[b][2 + j] = *(const s8 *)(xh + j * ctile + k);
                S[b][2 + NT + j] = *(const s8 *)(xl + j * ctile + k);
            }
        }
    };
    f4 acc[NT], cor[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) {
        acc[j] = f4{0.f, 0.f, 0.f, 0.f};
        cor[j] = f4{0.f, 0.f, 0.f, 0.f};
    }
    auto comp = [&](const Stage &S, int grp) {
#pragma unroll
        for (int b = 0; b < GB; b++) {
            if (grp * GB + b >= nblk) continue;  // uniform
#pragma unroll
            for (int j = 0; j < NT; j++) {
                acc[j] = N::mfma(S[b][0], S[b][2 + j], acc[j]);
                cor[j] = N::mfma(S[b][0], S[b][2 + NT + j], cor[j]);
                cor[j] = N::mfma(S[b][1], S[b][2 + j], cor[j]);
            }
        }
    };
    Stage s0, s1;
    load(s0, 0);
    for (int grp = 0; grp < ngrp; grp += 2) {
        load(s1, min(grp + 1, ngrp - 1));
        comp(s0, grp);
        load(s0, min(grp + 2, ngrp - 1));
        if (grp + 1 < ngrp) comp(s1, grp + 1);
    }
#pragma unroll
    for (int j = 0; j < NT; j++) red[wave][j][lane] = (acc[j] + cor[j] * N::IG) * N::IS;
    __syncthreads();
    // wave j < NT finishes column tile j: rows row0 + 4g + e, column 16j + r
    if (wave < NT) {
        const int j = wave, col = 16 * j + r;
        f4 sum = red[0][j][lane];
#pragma unroll
        for (int w = 1; w < NW; w++) sum += red[w][j][lane];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const size_t o = (size_t)(row0 + 4 * g + e) * NC + col;
            const float xv = X ? X[o] : x_from_copies<N>(XhT, (size_t)(XlT - XhT), row0 + 4 * g + e, col, nkt);
            float y = df * (xv + sum[e]);
            // the copies split the rounded FP32 y (no fused multiply-convert,
            // whose single rounding differs from y's at FP16 ties)
            asm volatile("" : "+v"(y));
            if (Y) Y[o] = y;
            const uint16_t h = N::enc(y);
            th[col][4 * g + e] = h;
            tl[col][4 * g + e] = N::enc((y - N::dec(h)) * N::G);
        }
    }
    __syncthreads();
    // Y's copies: rows row0 .. row0+15 are k = 32 kt + 8 g .. of tile kt =
    // row0 / 32, groups g0 and g0 + 1 (g0 = (row0 % 32) / 8): two 16-byte
    // fragments per column and array
    const int kt = row0 / 32, g0 = (row0 % 32) / 8;
    for (int i = t; i < NC * 4; i += 64 * NW) {
        const int arr = i / (NC * 2), c = (i % (NC * 2)) >> 1, half = i & 1;
        const size_t o = frag_off(c / 16, kt, 16 * (g0 + half) + (c % 16), nkt);
        *(uint4 *)((arr ? YlT : YhT) + o) =
            arr ? *(const uint4 *)&tl[c][8 * half] : *(const uint4 *)&th[c][8 * half];
    }
}

// ---------------------------------------------------------------------
// Band-sparse backward step.  Outside a band around the diagonal of each
// regime block the squared W is tiny: at the 2048-state lattice (d = 2048,
// U(1/252)) 57% of the 16 x 32 fragments of its FP16 hi/lo copies are zero
// and most of the rest hold values far below FP32 resolution.
// k_wbwd_flags marks, per fragment, whether a step visits it (some |W| >=
// 2^ev: its hi part is read) and whether it also reads the lo part (some
// |W| >= 2^el in FP16, 2^(el-3) in BF16, whose hi parts carry three fewer
// bits); k_wbwd_lists lists the visited k tiles, each entry kt + 65536 lo.
// What a step leaves out of an entry of W is the entry itself where it is
// under 2^ev, or its lo part where it is under 2^el: at most half an ulp of
// the hi part, 2^(el-12) in both encodings (and at most 2^-40, FP16's
// subnormal spacing 2^-24 of W 2^15, where el < -28).  A row of the step
// then differs from the full three-product sum by at most the left-out
// parts times |X|.  The defaults ev = -24, el = -12 move the prices of the
// 2048-state daily workload by under 1e-5 relative in exact arithmetic
// (16 scenarios); ev = -40, el = -28 leave out nothing above 2^-40.
// ---------------------------------------------------------------------
// Lists: per 16-row tile (d/16 lists), then per 64-row block (d/64: the
// union of its four tiles' lists) and per 32-row block (d/32: two tiles),
// each nkt + 1 ints: the count, then the k tiles in increasing order; a
// block's entry carries bit 17 + q where its row tile q visits the k tile
// (the others skip it) and bit 17 + nr + q where that tile reads the lo
// part (bit 16: some tile of the block does).
__device__ inline size_t tl_list4(int d) { return (size_t)(d / 16) * (d / 32 + 1); }
__device__ inline size_t tl_list2(int d) { return tl_list4(d) + (size_t)(d / 64) * (d / 32 + 1); }

// the magnitude bits (sign dropped) of a hi copy at which |W| reaches 2^e:
// FP16 holds W 2^15 (exponent field f + 15 for 2^f, f >= -14; subnormal
// 2^f is the mantissa 2^(f + 24); below 2^-24 every nonzero value counts),
// BF16 holds W (exponent field e + 127)
namespace {
inline uint32_t hi_bits(int mode, int e) {
    if (mode == 2) return (uint32_t)(e + 127) << 7;
    const int f = e + 15;
    if (f >= -14) return (uint32_t)(f + 15) << 10;
    if (f >= -24) return 1u << (f + 24);
    return 1u;
}

// the cutoffs ev, el (exponents) of the flags built from now on
std::atomic<int> g_cut_visit{-24}, g_cut_lo{-12};
}  // namespace

// The 64-state cells [c0, c1] that hold the W states of fragment states
// a .. a + n - 1: pmode 1, the same states; pmode 2, the regime order
// (fragment state x + nx r is W's state nr x + r, increasing in x within
// a regime).  A range over more than two regimes gets every cell.
__device__ inline void perm_cells(int pmode, int d, int nx, int nr, int a, int n, int &c0, int &c1) {
    const int b = a + n - 1;
    if (pmode == 1) {
        c0 = a >> 6;
        c1 = b >> 6;
        return;
    }
    const int ra = a / nx, rb = b / nx;
    const int pa = nr * (a - ra * nx) + ra, pb = nr * (b - rb * nx) + rb;
    if (ra == rb) {
        c0 = pa >> 6;
        c1 = pb >> 6;
    } else if (rb == ra + 1) {  // the end of regime ra, the start of rb
        c0 = min(pa, rb) >> 6;
        c1 = max(nr * (nx - 1) + ra, pb) >> 6;
    } else {
        c0 = 0;
        c1 = (d - 1) >> 6;
    }
}

// k_wbwd_flags: byte flags[it][kt] (d/16 x d/32): bit 0 = the fragment is
// visited, bit 1 = its lo part is read; a wave per (it, four k tiles), the
// four 16-byte loads in flight; lattice blockIdx.y of the batch.  pmode 1
// (fragments in W's state order) or 2 (regime order, nx, nr): a fragment
// whose W cells are all flagged zero holds zeros and is not read; pmode 0
// (any other order): every fragment is read.
__global__ __launch_bounds__(256) void k_wbwd_flags(int d, char *__restrict__ base, const LayBatch bt,
                                                    uint32_t tvis, uint32_t tlo, int pmode, int nx, int nr) {
    const uint16_t *__restrict__ Hf = at<const uint16_t>(base, bt.s[blockIdx.y].Hf);
    uint8_t *__restrict__ flags = at<uint8_t>(base, bt.s[blockIdx.y].flags);
    const int nkt = d / 32, lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);  // (it, k0 / 4)
    if (w >= d / 16 * (nkt / 4)) return;  // wave-uniform
    const int it = w / (nkt / 4), k0 = 4 * (w % (nkt / 4));
    bool read[4] = {true, true, true, true};
    if (pmode) {  // uniform
        // lanes 16 u .. 16 u + 15 read the flags of fragment u's W cells
        // (a box of at most 16 cells, else the fragment is read) in one load
        const uint8_t *cf = at<const uint8_t>(base, bt.s[blockIdx.y].cf);
        const int nb = d / 64, u = lane >> 4, c = lane & 15;
        int r0, r1, c0, c1;
        perm_cells(pmode, d, nx, nr, 16 * it, 16, r0, r1);
        perm_cells(pmode, d, nx, nr, 32 * (k0 + u), 32, c0, c1);
        const int nrow = r1 - r0 + 1, ncol = c1 - c0 + 1;
        bool nz = nrow * ncol > 16 && c == 0;
        if (nrow * ncol <= 16 && c < nrow * ncol) nz = cf[(size_t)(r0 + c / ncol) * nb + c0 + c % ncol] != 0;
        const uint64_t m = __ballot(nz);
#pragma unroll
        for (int q = 0; q < 4; q++) read[q] = ((m >> (16 * q)) & 0xffffu) != 0;
    }
    uint4 h[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
        h[u] = read[u] ? *(const uint4 *)(Hf + frag_off(it, k0 + u, lane, nkt)) : uint4{0u, 0u, 0u, 0u};
    unsigned word = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const uint32_t q[4] = {h[u].x, h[u].y, h[u].z, h[u].w};
        uint32_t m = 0;  // the largest magnitude bits of the lane's eight values
#pragma unroll
        for (int c = 0; c < 4; c++) m = max(m, max(q[c] & 0x7fffu, (q[c] >> 16) & 0x7fffu));
        const unsigned f = (__any(m >= tvis) ? 1u : 0u) | (__any(m >= tlo) ? 2u : 0u);
        word |= f << (8 * u);
    }
    if (lane == 0) *(unsigned *)(flags + (size_t)it * nkt + k0) = word;
}

// k_wbwd_lists: one wave per list (d/16 row tiles, then d/64 64-row
// blocks, then d/32 32-row blocks), lanes over k tiles: the union of the
// list's row tiles' flags, compacted in k order with a ballot prefix;
// lattice blockIdx.y of the batch
__global__ __launch_bounds__(256) void k_wbwd_lists(int d, char *__restrict__ base, const LayBatch bt) {
    const uint8_t *__restrict__ flags = at<const uint8_t>(base, bt.s[blockIdx.y].flags);
    int *__restrict__ tl = at<int>(base, bt.s[blockIdx.y].tl);
    const int nkt = d / 32, n16 = d / 16, n64 = d / 64, n32 = d / 32;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= n16 + n64 + n32) return;  // wave-uniform
    int r0, nr;
    if (t < n16) { r0 = t; nr = 1; }
    else if (t < n16 + n64) { r0 = 4 * (t - n16); nr = 4; }
    else { r0 = 2 * (t - n16 - n64); nr = 2; }
    int *out = tl + (size_t)t * (nkt + 1);
    int n = 0;
    for (int k0 = 0; k0 < nkt; k0 += 64) {  // nkt <= 128
        const int kt = k0 + lane;
        int f = 0, vm = 0, lm = 0;  // union of the flags; which row tiles visit kt, read its lo parts
        if (kt < nkt)
            for (int q = 0; q < nr; q++) {
                const int fq = flags[(size_t)(r0 + q) * nkt + kt];
                f |= fq;
                vm |= (fq & 1) << q;
                lm |= ((fq >> 1) & 1) << q;
            }
        const uint64_t m = __ballot((f & 1) != 0);
        if (f & 1)
            out[1 + n + __popcll(m & ((1ull << lane) - 1))] =
                kt + ((f & 2) ? 65536 : 0) + (nr > 1 ? (vm << 17) + (lm << (17 + nr)) : 0);
        n += __popcll(m);
    }
    if (lane == 0) out[0] = n;
}

// one lattice of a batched backward step (arena offsets)
struct BwdScen {
    uint64_t Hf, Lf, tl;    // W's fragment-major hi/lo copies, their tile lists
    uint64_t X, XT, Y, YT;  // values and copies in, values and copies out (X 0: from XT; Y 0: not written)
    float df;
    int pad;
};
constexpr int kBwdBatch = 32;
struct BwdBatch {
    BwdScen s[kBwdBatch];
};

// Workgroup = (lattice blockIdx.y, 16-row tile it, 16-column tile j).  The
// block index b maps to (it, j) so that the column tiles of one row tile
// run on one XCD (block b goes to XCD b % 8) and its L2 serves their
// common W fragments.  NW waves take the listed k tiles round robin, E
// tiles' loads in flight per wave; the waves' sums are added in wave
// order, then Y = df (X + W X) and Y's hi/lo copies as in k_wbwd1.
template <bool BF, int NW, int E>
__global__ __launch_bounds__(64 * NW) void k_wbwd_sp(char *__restrict__ base, const BwdBatch bt, int d,
                                                     int nct) {
    using N = Num<BF>;
    __shared__ f4 red[NW][64];
    __shared__ uint16_t th[16][16], tlo[16][16];
    __shared__ __attribute__((aligned(16))) uint16_t xs[2][32][8];  // own rows' value copies (X = 0)
    const BwdScen &S = bt.s[blockIdx.y];
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int nkt = d / 32, ncp = 16 * nct;
    const int b = blockIdx.x, q = b >> 3;
    const int j = q % nct, it = (b & 7) + 8 * (q / nct);
    const int row0 = 16 * it;
    const size_t n = (size_t)d * ncp;
    const int *lst = at<const int>(base, S.tl) + (size_t)it * (nkt + 1);
    const int cnt = __builtin_amdgcn_readfirstlane(lst[0]);
    const uint16_t *ha = at<const uint16_t>(base, S.Hf) + frag_off(it, 0, lane, nkt);
    const uint16_t *la = at<const uint16_t>(base, S.Lf) + frag_off(it, 0, lane, nkt);
    const uint16_t *xh = at<const uint16_t>(base, S.XT) + frag_off(j, 0, lane, nkt);
    const uint16_t *xl = xh + n;
    // without FP32 values: this block's rows of column tile j are groups
    // g0, g0 + 1 of k tile row0 / 32, 32 fragment lanes per copy; one
    // 16-byte load per thread of the first 64, stored after the K loop
    const int g0 = (row0 % 32) / 8;
    uint4 xstage = {0u, 0u, 0u, 0u};
    if (!S.X && t < 64) {
        const int arr = t >> 5, l = 16 * g0 + (t & 31);
        xstage = *(const uint4 *)(at<const uint16_t>(base, S.XT) + (arr ? n : 0) +
                                  frag_off(j, row0 / 32, l, nkt));
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f}, cor = {0.f, 0.f, 0.f, 0.f};
    for (int e0 = wave; e0 < cnt; e0 += NW * E) {
        s8 a[E][4];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int ent = __builtin_amdgcn_readfirstlane(lst[1 + min(e0 + e * NW, cnt - 1)]);
            const size_t k = (size_t)512 * (ent & 0xffff);
            a[e][0] = *(const s8 *)(ha + k);
            a[e][1] = (ent >> 16) ? *(const s8 *)(la + k) : s8{};  // uniform
            a[e][2] = *(const s8 *)(xh + k);
            a[e][3] = *(const s8 *)(xl + k);
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (e0 + e * NW < cnt) {  // uniform (scalar operands)
                acc = N::mfma(a[e][0], a[e][2], acc);
                cor = N::mfma(a[e][0], a[e][3], cor);
                cor = N::mfma(a[e][1], a[e][2], cor);
            }
        }
    }
    red[wave][lane] = (acc + cor * N::IG) * N::IS;
    if (!S.X && t < 64) *(uint4 *)&xs[t >> 5][t & 31][0] = xstage;
    __syncthreads();
    if (wave == 0) {
        f4 sum = red[0][lane];
#pragma unroll
        for (int w = 1; w < NW; w++) sum += red[w][lane];
        const float *X = S.X ? at<const float>(base, S.X) : nullptr;
        float *Y = S.Y ? at<float>(base, S.Y) : nullptr;
        const int col = 16 * j + r;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const size_t o = (size_t)(row0 + 4 * g + e) * ncp + col;
            // row 4g + e of the block: group g / 2 past g0, element 4 (g & 1) + e
            const int fl = 16 * (g >> 1) + r, fe = 4 * (g & 1) + e;
            const float xv = X ? X[o] : N::dec(xs[0][fl][fe]) + N::dec(xs[1][fl][fe]) * N::IG;
            float y = S.df * (xv + sum[e]);
            asm volatile("" : "+v"(y));  // the copies split the rounded y
            if (Y) Y[o] = y;
            const uint16_t h = N::enc(y);
            th[r][4 * g + e] = h;
            tlo[r][4 * g + e] = N::enc((y - N::dec(h)) * N::G);
        }
    }
    __syncthreads();
    if (wave == 0) {
        // rows row0 .. row0+15 are k = 32 kt + 8 g .. of k tile kt = row0 /
        // 32, groups g0 and g0 + 1: two 16-byte fragments per column and
        // array, one per lane
        const int kt = row0 / 32, g0 = (row0 % 32) / 8;
        const int arr = lane >> 5, c = (lane >> 1) & 15, half = lane & 1;
        uint16_t *YT = at<uint16_t>(base, S.YT) + (arr ? n : 0);
        const size_t o = frag_off(j, kt, 16 * (g0 + half) + c, nkt);
        *(uint4 *)(YT + o) = arr ? *(const uint4 *)&tlo[c][8 * half] : *(const uint4 *)&th[c][8 * half];
    }
}

// Row-block variant for many lattices per launch (48 value columns): a
// workgroup takes RB row tiles (16 RB rows) and all three column tiles of
// one lattice, so every W fragment and every value fragment it loads feeds
// RB x 3 (three-product) MFMA groups instead of one.  Its k tiles are the
// union of the RB row tiles' lists (a tile that is zero for one of the row
// tiles adds exact zeros there).  Workgroups are spread so that each XCD
// (block b runs on XCD b % 8) takes a contiguous run of (lattice, row
// block) pairs: a lattice's value copies are fetched into few L2s.  NW
// waves take the listed k tiles round robin, E tiles' loads in flight; the
// sums are added in wave order.  per_xcd = ceil(nb * nrb / 8), grid 8 per_xcd.
template <bool BF, int RB, int NW, int E>
__global__ __launch_bounds__(64 * NW) void k_wbwd_blk(char *__restrict__ base, const BwdBatch bt, int nb,
                                                      int d, int per_xcd) {
    using N = Num<BF>;
    constexpr int NT = 3, NC = 48, NTILE = RB * NT, ROWS = 16 * RB;
    static_assert(RB == 2 || RB == 4, "row blocks of 32 or 64 rows");
    __shared__ f4 red[NW][NTILE][64];
    // Y's copies column by column; 8 elements of padding per column (an
    // 80-byte pitch for 32 rows) spread the columns of a wave's 2-byte
    // writes and 16-byte reads over the banks
    __shared__ __attribute__((aligned(16))) uint16_t th[NC][ROWS + 8], tlo[NC][ROWS + 8];
    constexpr int NKB = ROWS / 32;  // k tiles of this block's rows
    __shared__ __attribute__((aligned(16))) uint16_t xs[2][NKB][NT][64][8];  // own rows' value copies (X = 0)
    const int nrb = d / ROWS;
    const int v = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (v >= nb * nrb) return;  // whole workgroup, before any barrier
    const int s = v / nrb, rb = v % nrb;
    const BwdScen &S = bt.s[s];
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int nkt = d / 32, row0 = ROWS * rb;
    const size_t n = (size_t)d * NC, rstride = (size_t)nkt * 512;  // elements per 16-row strip
    const int *lst = at<const int>(base, S.tl) + (RB == 4 ? tl_list4(d) : tl_list2(d)) +
                     (size_t)rb * (nkt + 1);
    // the list's head in one load: lane l holds lst[l] (the count, then
    // the first 63 entries), read back with v_readlane; later entries
    // (d > 2016) are loaded when needed
    const int head = lane <= nkt ? lst[lane] : 0;
    const int cnt = __builtin_amdgcn_readfirstlane(head);
    auto entry = [&](int idx) {  // idx wave-uniform, < cnt
        return idx < 63 ? __builtin_amdgcn_readlane(head, 1 + idx) : __builtin_amdgcn_readfirstlane(lst[1 + idx]);
    };
    const uint16_t *ha = at<const uint16_t>(base, S.Hf) + frag_off(RB * rb, 0, lane, nkt);
    const uint16_t *la = at<const uint16_t>(base, S.Lf) + frag_off(RB * rb, 0, lane, nkt);
    const uint16_t *xh = at<const uint16_t>(base, S.XT) + frag_off(0, 0, lane, nkt);
    const uint16_t *xl = xh + n;
    // the epilogue's X values (tiles wave, wave + NW, ..) loaded up front
    constexpr int TPW = (NTILE + NW - 1) / NW;
    const float *X = S.X ? at<const float>(base, S.X) : nullptr;
    float xv[TPW][4];
    // without FP32 values: the block's rows are k tiles row0 / 32 ..; their
    // fragments (2 copies x NKB x NT x 64 lanes of 16 bytes) are loaded now,
    // one per thread, and go to LDS after the K loop
    constexpr int NXF = 2 * NKB * NT * 64, XPT = (NXF + 64 * NW - 1) / (64 * NW);
    uint4 xstage[XPT];
    if (X) {
#pragma unroll
        for (int q = 0; q < TPW; q++) {
            const int tile = wave + q * NW;
            if (tile < NTILE) {  // uniform
                const int i = tile / NT, j = tile % NT;
#pragma unroll
                for (int e = 0; e < 4; e++) xv[q][e] = X[(size_t)(row0 + 16 * i + 4 * g + e) * NC + 16 * j + r];
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < XPT; u++) {
            const int f = t + 64 * NW * u;
            xstage[u] = uint4{0u, 0u, 0u, 0u};
            if (f < NXF) {
                const int l = f & 63, j = (f >> 6) % NT, kb = (f >> 6) / NT % NKB, arr = (f >> 6) / (NT * NKB);
                xstage[u] = *(const uint4 *)(at<const uint16_t>(base, S.XT) + (arr ? n : 0) +
                                             frag_off(j, row0 / 32 + kb, l, nkt));
            }
        }
    }
    f4 acc[RB][NT], cor[RB][NT];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) {
            acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
            cor[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    for (int e0 = wave; e0 < cnt; e0 += NW * E) {
        s8 w[E][2 * RB], x[E][2 * NT];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int ent = entry(min(e0 + e * NW, cnt - 1));
            const size_t k = (size_t)512 * (ent & 0xffff);
#pragma unroll
            for (int i = 0; i < RB; i++) {  // uniform: does row tile i visit this k tile, read its lo part
                const bool vis = ((ent >> (17 + i)) & 1) != 0, lo = ((ent >> (17 + RB + i)) & 1) != 0;
                w[e][i] = vis ? *(const s8 *)(ha + i * rstride + k) : s8{};
                w[e][RB + i] = lo ? *(const s8 *)(la + i * rstride + k) : s8{};
            }
#pragma unroll
            for (int j = 0; j < NT; j++) {
                x[e][j] = *(const s8 *)(xh + j * rstride + k);
                x[e][NT + j] = *(const s8 *)(xl + j * rstride + k);
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (e0 + e * NW < cnt) {  // uniform (scalar operands)
#pragma unroll
                for (int i = 0; i < RB; i++)
#pragma unroll
                    for (int j = 0; j < NT; j++) {
                        acc[i][j] = N::mfma(w[e][i], x[e][j], acc[i][j]);
                        cor[i][j] = N::mfma(w[e][i], x[e][NT + j], cor[i][j]);
                        cor[i][j] = N::mfma(w[e][RB + i], x[e][j], cor[i][j]);
                    }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) red[wave][i * NT + j][lane] = (acc[i][j] + cor[i][j] * N::IG) * N::IS;
    if (!X) {
#pragma unroll
        for (int u = 0; u < XPT; u++) {
            const int f = t + 64 * NW * u;
            if (f < NXF) {
                const int l = f & 63, j = (f >> 6) % NT, kb = (f >> 6) / NT % NKB, arr = (f >> 6) / (NT * NKB);
                *(uint4 *)&xs[arr][kb][j][l][0] = xstage[u];
            }
        }
    }
    __syncthreads();
    if (!X)
#pragma unroll
        for (int q = 0; q < TPW; q++) {
            const int tile = wave + q * NW;
            if (tile < NTILE) {  // uniform
                const int i = tile / NT, j = tile % NT;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    // local row 16 i + 4 g + e: k tile kb, group kg, element ke
                    const int kk = 16 * i + 4 * g + e, kb = kk >> 5, fl = 16 * ((kk & 31) >> 3) + r, ke = kk & 7;
                    xv[q][e] = N::dec(xs[0][kb][j][fl][ke]) + N::dec(xs[1][kb][j][fl][ke]) * N::IG;
                }
            }
        }
    float *Y = S.Y ? at<float>(base, S.Y) : nullptr;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
        const int tile = wave + q * NW;
        if (tile >= NTILE) break;  // uniform
        const int i = tile / NT, j = tile % NT;
        f4 sum = red[0][tile][lane];
#pragma unroll
        for (int w = 1; w < NW; w++) sum += red[w][tile][lane];
        const int col = 16 * j + r;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = 16 * i + 4 * g + e;
            const size_t o = (size_t)(row0 + row) * NC + col;
            float y = S.df * (xv[q][e] + sum[e]);
            asm volatile("" : "+v"(y));  // the copies split the rounded y
            if (Y) Y[o] = y;
            const uint16_t h = N::enc(y);
            th[col][row] = h;
            tlo[col][row] = N::enc((y - N::dec(h)) * N::G);
        }
    }
    __syncthreads();
    // Y's copies: local rows 8 q .. 8 q + 7 of column c are group q % 4 of k
    // tile (row0 + 8 q) / 32; one 16-byte fragment row per thread and unit
    // (array, column tile, 8-row group, column in the tile), consecutive
    // threads on consecutive fragment lanes
    constexpr int NQ = ROWS / 8, NU = 2 * NT * NQ * 16;
    for (int f = t; f < NU; f += 64 * NW) {
        const int c16 = f & 15, q = (f >> 4) % NQ, j = ((f >> 4) / NQ) % NT, arr = (f >> 4) / (NQ * NT);
        const int c = 16 * j + c16;
        const int kt = (row0 + 8 * q) / 32;
        uint16_t *YT = at<uint16_t>(base, S.YT) + (arr ? n : 0);
        const size_t o = frag_off(j, kt, 16 * (q & 3) + c16, nkt);
        *(uint4 *)(YT + o) = arr ? *(const uint4 *)&tlo[c][8 * q] : *(const uint4 *)&th[c][8 * q];
    }
}

}  // namespace

// ---------------------------------------------------------------------
// C-ABI.  A W-form squaring state for one lattice lives in one arena
// allocation laid out by ophip_wsq_layout; ophip_wsq_* drive it.
// ---------------------------------------------------------------------

// mode: 0 = FP16 one product, 1 = FP16 three products (hi/lo),
//       2 = BF16 three products
struct WsqLayout {
    size_t W0, W1, Fh, Fl, part, fl[2], tl, tf, total;  // byte offsets; Fh/Fl: -1 in mode 0
};

static WsqLayout wsq_layout(int d, int mode) {
    WsqLayout L{};
    size_t dd = (size_t)d * d, off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    L.W0 = take(4 * dd);
    L.W1 = take(4 * dd);
    // the backward layout's fragment-major hi/lo parts of the final W (the
    // three-product modes only: the steps read both)
    L.Fh = mode ? take(2 * dd) : (size_t)-1;
    L.Fl = mode ? take(2 * dd) : (size_t)-1;
    L.part = take(8 * (size_t)(d / 64) * d);
    for (int b = 0; b < 2; b++) L.fl[b] = take((size_t)(d / 64) * (d / 64));  // cell flags
    // backward-step tile lists: per 16-row tile, 64-row block, 32-row block
    L.tl = take(4 * (size_t)(d / 16 + d / 64 + d / 32) * (d / 32 + 1));
    L.tf = take((size_t)(d / 16) * (d / 32));  // their fragment flags
    L.total = off;
    return L;
}

// bytes of device state for one lattice of size d (d % 128 == 0)
OPHIP_EXPORT unsigned long ophip_wsq_bytes(int d, int mode) {
    if (d <= 0 || d % 64 || mode < 0 || mode > 2) return 0;
    return (unsigned long)wsq_layout(d, mode).total;
}

namespace {

WScen make_scen(unsigned state, int d, int mode, int src) {
    WsqLayout L = wsq_layout(d, mode);
    const uint64_t p = state;
    WScen s{};
    s.d = d;
    const int dst = src ^ 1;
    s.W = p + (src ? L.W1 : L.W0);
    s.Wn = p + (dst ? L.W1 : L.W0);
    s.part = p + L.part;
    s.fl = p + L.fl[src];
    s.fln = p + L.fl[dst];
    return s;
}

constexpr int kMaxScen = 4096;

// zero-cell skipping in k_wsq (on by default; off gives the dense K loop,
// the same bits -- for tests and timing)
int g_wsq_skip = 1;

// K-cell cutoff of the sparse rounds: a K cell whose operand bounds
// multiply to at most 2^kcut is left out; kcut <= -192 keeps every cell
// that may be nonzero (the dense loop's bits)
std::atomic<int> g_wsq_kcut{-36};

template <class F> unsigned launch_mode(int mode, F f) {
    switch (mode) {
    case 0: f(std::integral_constant<bool, false>(), std::integral_constant<bool, false>()); break;
    case 1: f(std::integral_constant<bool, false>(), std::integral_constant<bool, true>()); break;
    default: f(std::integral_constant<bool, true>(), std::integral_constant<bool, true>()); break;
    }
    return check_launch();
}

}  // namespace

// A plan holds the device-side scenario and work tables of one batch of
// lattices (built once, outside any graph capture); executing it only
// launches kernels, so it can be captured into a HIP graph.  Lattices may
// need different numbers of squarings (the reference planner gives each
// kernel its own niter, CModel Common.cs:289-307): round r squares only
// the lattices with nsq > r, as the reference's njtodo_r tables do
// (CModel Common.cs:508-545, Device.cs:222-224).
// Work items are kept per tile class (c = 0: 64x64 tiles, c = 1: 128x128):
// each lattice squares with the tile its plan gives it, and a round
// launches one persistent kernel per non-empty class.
struct WsqPlan {
    int n = 0, mode = 0, grid = 0, dmax = 0, rounds = 0;
    std::vector<int> nsq, tiles;
    int n_split = 0;
    std::vector<int> n_round[2];    // items per round and class (per-lattice counts)
    int n_all[2][2] = {{0, 0}, {0, 0}};  // items per class and parity (uniform count)
    void *dev = nullptr;            // one hipMalloc: tables below
    WScen *sc = nullptr, *sc_split = nullptr;
    int *nsq_dev = nullptr;
    WItem *it_split = nullptr, *it_all[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<WItem *> it_round[2];
    // sparse rounds (every lattice d <= 4096): per class, 8 XCD lists of
    // cap live items (16-item chunks in static order) and the chunks'
    // counts (class c: counts + 8 cap[0] / 16 c, XCD x: + x cap[c] / 16)
    bool sparse_ok = false;
    int cap[2] = {0, 0};
    WLive *live[2] = {nullptr, nullptr};
    int *counts = nullptr;
};

// XCD-aware order of a squaring round's work items.  Workgroup b of a
// launch runs on XCD b % 8 and the persistent grid is a multiple of 8, so
// item j runs on XCD j % 8.  Tiles that share operand panels should share
// an XCD (its L2), and the work of a banded W (most of it near the
// diagonal while zero cells are skipped) should spread evenly.  With 8
// lattices or more, each lattice goes whole to the least-loaded XCD: its A
// and B panels are then fetched into one L2 only (16 lattices of d = 2048:
// the group's squarings take 7.33 instead of 7.68 ms per solve).  With
// fewer, a lattice with at least 8 tile rows is cut into 8 blocks of
// consecutive tile rows, one per XCD (each holds its share of the band;
// its A panels stay in its L2), and smaller ones go whole to the
// least-loaded XCD.  The per-XCD lists are then balanced (moving items off
// the longest lists), sorted longest item first and interleaved.
static std::vector<WItem> xcd_order(const std::vector<WItem> &v, const int *ds, int n, int T) {
    constexpr int X = 8;
    if (v.size() < 2 * X) return v;
    std::vector<std::vector<WItem>> per(X);
    std::vector<std::vector<WItem>> by_lat(n);
    for (const WItem &w : v) by_lat[w.s % n].push_back(w);
    std::vector<size_t> load(X, 0);
    for (int i = 0; i < n; i++) {
        if (by_lat[i].empty()) continue;
        const int nt = ds[i] / T;
        if (nt >= X && n < X) {
            for (const WItem &w : by_lat[i]) {
                const int x = w.tm * X / nt;
                per[x].push_back(w);
                load[x]++;
            }
        } else {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            for (const WItem &w : by_lat[i]) per[x].push_back(w);
            load[x] += by_lat[i].size();
        }
    }
    const size_t total = v.size(), target = (total + X - 1) / X;
    for (int x = 0; x < X; x++)
        while (per[x].size() < target) {
            int big = 0;
            for (int y = 1; y < X; y++)
                if (per[y].size() > per[big].size()) big = y;
            if (per[big].size() <= target) break;
            per[x].push_back(per[big].back());
            per[big].pop_back();
        }
    // within an XCD, longest items first (an item's cost is its lattice's d,
    // the K length): the XCD's workgroups take items round-robin, so this is
    // longest-processing-time-first assignment; equal sizes keep their order
    for (int x = 0; x < X; x++)
        std::stable_sort(per[x].begin(), per[x].end(),
                         [&](const WItem &a, const WItem &b) { return ds[a.s % n] > ds[b.s % n]; });
    std::vector<WItem> out;
    out.reserve(total);
    for (size_t k = 0; out.size() < total; k++)
        for (int x = 0; x < X; x++)
            if (k < per[x].size()) out.push_back(per[x][k]);
    return out;
}

// Plan the W-form pipeline for `n` lattices (states[i]: arena handle of a
// state of ophip_wsq_bytes(ds[i], mode) bytes, ds[i] % 128 == 0; nsqs[i]:
// its number of squarings, or null for a count given at exec time).
// tiles[i]: 64 or 128, the tile lattice i squares with; grid: persistent
// workgroups per squaring launch (0 = 2 per CU).  Work items of all
// lattices of a class are interleaved tile-major so uneven sizes balance.
static WsqPlan *make_plan(int n, const unsigned *states, const int *ds, const int *nsqs, int mode,
                          const int *tiles, int grid) {
    Device &D = dev();
    if (n <= 0 || n > kMaxScen || mode < 0 || mode > 2) { set_status(kErrBadArg); return nullptr; }
    for (int i = 0; i < n; i++) {
        if (ds[i] <= 0 || ds[i] % 128 || !ptr(states[i]) || (nsqs && (nsqs[i] < 0 || nsqs[i] > 64)) ||
            !handle_range_ok(states[i], wsq_layout(ds[i], mode).total) ||
            (tiles[i] != 64 && tiles[i] != 128)) {
            set_status(kErrBadArg);
            return nullptr;
        }
    }
    const int ncu = D.n_cu > 0 ? D.n_cu : 256;
    WsqPlan *P = new WsqPlan();
    P->n = n;
    P->mode = mode;
    P->grid = grid > 0 ? grid : 2 * ncu;
    P->tiles.assign(tiles, tiles + n);
    P->nsq.assign(n, 0);
    if (nsqs)
        for (int i = 0; i < n; i++) {
            P->nsq[i] = nsqs[i];
            P->rounds = std::max(P->rounds, nsqs[i]);
        }
    std::vector<WScen> sc(2 * n), sc_split(n);
    for (int i = 0; i < n; i++) {
        sc[i] = make_scen(states[i], ds[i], mode, 0);
        sc[n + i] = make_scen(states[i], ds[i], mode, 1);
        // the first split and diagonal fix work on W0 in place (buffer 0)
        sc_split[i] = make_scen(states[i], ds[i], mode, 1);
        sc_split[i].W = sc[i].W;
        P->dmax = std::max(P->dmax, ds[i]);
    }
    // T: tile of the items; cls: lattices of tile class cls (-1: all)
    auto items_for = [&](int T, int cls, int dir, int round) {
        std::vector<WItem> v;
        const int maxt = P->dmax / T;
        for (int tm = 0; tm < maxt; tm++)
            for (int tn = 0; tn < maxt; tn++)
                for (int i = 0; i < n; i++)
                    if (tm < ds[i] / T && tn < ds[i] / T && (round < 0 || P->nsq[i] > round) &&
                        (cls < 0 || (tiles[i] == 128) == (cls == 1)))
                        v.push_back(WItem{dir * n + i, tm, tn, 0});
        return v;
    };
    std::vector<std::vector<WItem>> lists;
    lists.push_back(items_for(64, -1, 0, -1));
    for (int c = 0; c < 2; c++) {
        const int T = c ? 128 : 64;
        lists.push_back(xcd_order(items_for(T, c, 0, -1), ds, n, T));
        lists.push_back(xcd_order(items_for(T, c, 1, -1), ds, n, T));
        for (int r = 0; r < P->rounds; r++) lists.push_back(xcd_order(items_for(T, c, r & 1, r), ds, n, T));
    }
    size_t nitems = 0;
    for (auto &l : lists) nitems += l.size();
    P->sparse_ok = P->dmax <= 4096;
    // each XCD's list: 16-slot chunks for the 16-item blocks of k_wlive
    for (int c = 0; c < 2; c++)
        P->cap[c] = 16 * (int)(((lists[1 + c * (2 + P->rounds)].size() + 7) / 8 + 15) / 16);
    if (P->cap[0] / 16 > kLiveChunks || P->cap[1] / 16 > kLiveChunks) P->sparse_ok = false;
    size_t bytes = sizeof(WScen) * 3 * n + sizeof(int) * (n + 4) + sizeof(WItem) * nitems;
    bytes = (bytes + 255) & ~(size_t)255;
    const size_t live_off = bytes;
    bytes += sizeof(WLive) * 8 * (size_t)(P->cap[0] + P->cap[1]) + sizeof(int) * (P->cap[0] + P->cap[1]) / 2;
    if (hipMalloc(&P->dev, bytes) != hipSuccess) {
        delete P;
        set_status(kErrOutOfArena);
        return nullptr;
    }
    char *p = (char *)P->dev;
    P->sc = (WScen *)p;
    P->sc_split = P->sc + 2 * n;
    P->nsq_dev = (int *)(P->sc_split + n);
    WItem *cur = (WItem *)(P->nsq_dev + n + 4);
    std::vector<WItem *> where;
    for (auto &l : lists) {
        where.push_back(cur);
        cur += l.size();
    }
    P->live[0] = (WLive *)(p + live_off);
    P->live[1] = P->live[0] + 8 * (size_t)P->cap[0];
    P->counts = (int *)(P->live[1] + 8 * (size_t)P->cap[1]);
    P->it_split = where[0];
    P->n_split = (int)lists[0].size();
    size_t w = 1;
    for (int c = 0; c < 2; c++) {
        for (int src = 0; src < 2; src++, w++) {
            P->it_all[c][src] = where[w];
            P->n_all[c][src] = (int)lists[w].size();
        }
        for (int r = 0; r < P->rounds; r++, w++) {
            P->it_round[c].push_back(where[w]);
            P->n_round[c].push_back((int)lists[w].size());
        }
    }
    hipError_t e = hipMemcpy(P->sc, sc.data(), sizeof(WScen) * 2 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->sc_split, sc_split.data(), sizeof(WScen) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->nsq_dev, P->nsq.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    for (size_t l = 0; l < lists.size() && e == hipSuccess; l++)
        if (!lists[l].empty())
            e = hipMemcpy(where[l], lists[l].data(), sizeof(WItem) * lists[l].size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(P->counts, 0, sizeof(int) * (P->cap[0] + P->cap[1]) / 2);
    if (e != hipSuccess) {
        (void)hipFree(P->dev);
        delete P;
        set_status(hip_status(e));
        return nullptr;
    }
    // both buffers' cell flags: "may be nonzero" everywhere until the
    // split or a round writes that buffer (ordered after earlier work on
    // the states)
    for (int i = 0; i < n; i++)
        for (int b = 0; b < 2; b++) {
            const int nc = ds[i] / 64;
            hipLaunchKernelGGL(k_fill8, dim3(cdiv(nc * nc, 256)), dim3(256), 0, D.stream,
                               reinterpret_cast<uint8_t *>(ptr(states[i] + (unsigned)wsq_layout(ds[i], mode).fl[b])),
                               nc * nc, kFlagAny);
        }
    if (check_launch()) {
        (void)hipFree(P->dev);
        delete P;
        return nullptr;
    }
    set_status(kOk);
    return P;
}

// tile: 128 or 64 for every lattice (0 = 128 when the lattices hold at
// least one 128-tile per CU, else 64)
OPHIP_EXPORT unsigned long ophip_wsq_plan(int n, const unsigned *states, const int *ds,
                                          const int *nsqs, int mode, int tile, int grid) {
    if (n <= 0 || n > kMaxScen || !ds) { set_status(kErrBadArg); return 0; }
    if (tile == 0) {
        const int ncu = dev().n_cu > 0 ? dev().n_cu : 256;
        long t128 = 0;
        for (int i = 0; i < n; i++) t128 += (long)(ds[i] / 128) * (ds[i] / 128);
        tile = (t128 >= (long)ncu) ? 128 : 64;
    }
    std::vector<int> tiles(n, tile);
    return (unsigned long)(uintptr_t)make_plan(n, states, ds, nsqs, mode, tiles.data(), grid);
}

// the same with a tile per lattice (64 or 128 each)
OPHIP_EXPORT unsigned long ophip_wsq_plan_tiles(int n, const unsigned *states, const int *ds,
                                                const int *nsqs, int mode, const int *tiles,
                                                int grid) {
    if (!tiles) { set_status(kErrBadArg); return 0; }
    return (unsigned long)(uintptr_t)make_plan(n, states, ds, nsqs, mode, tiles, grid);
}

OPHIP_EXPORT void ophip_wsq_plan_free(unsigned long plan) {
    WsqPlan *P = (WsqPlan *)(uintptr_t)plan;
    if (!P) return;
    (void)hipStreamSynchronize(dev().stream);
    (void)hipFree(P->dev);
    delete P;
}

// ---------------------------------------------------------------------
// k_rs_wsplit: the FP64 assembly of W0 = dt*L (rs_elem) fused with
// k_wsplit's work (flush rule, partial row sums, cell flags of buffer 0):
// one 64-row stripe of one lattice per workgroup, wave w takes the cells
// w, w+4, ... of the stripe, lane = row (the row's rates evaluated once,
// the wave's cell flags read in one load).  A cell whose entries are all
// structurally zero (rs_block_may_be_nonzero) writes nothing when buffer
// 0's flag already says it holds zeros (the diagonal fix selects its
// partial sums away), else zeros and flag 0.  Every other cell is written
// with the values, partial sums and flag that k_rs_assemble_batch (dt*L)
// followed by k_wsplit leave.
// ---------------------------------------------------------------------
namespace {

struct RSWEntry {
    RSLattice L;
    double dt;
    float *W;      // buffer 0
    double *part;  // partial row sums, [d/64 column blocks][d]
    uint8_t *fl;   // buffer 0's cell flags
};
constexpr int kRSWBatch = 32;
struct RSWBatch {
    RSWEntry e[kRSWBatch];
};

template <bool NM, bool BF, bool LO>
__global__ __launch_bounds__(256) void k_rs_wsplit(const RSWBatch b) {
    using N = Num<BF>;
    const RSWEntry &E = b.e[blockIdx.y];
    const int nx = E.L.nx, nr = E.L.nr, d = nx * nr, nc = d / 64, tm = blockIdx.x;
    if (tm >= nc) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = tm * 64 + lane;
    // the row's node, regime and rates: rs_elem's terms, evaluated once
    const int xi = NM ? i / nr : i % nx, ri = NM ? i % nr : i / nx;
    double am, ap;
    rs_rates(E.L, xi + nx * ri, am, ap);
    const double gd = -(am + ap + (double)(nr - 1) * E.L.q_off);
    uint8_t *frow = E.fl + (size_t)tm * nc;
    for (int q0 = 0; wave + 4 * q0 < nc; q0 += 64) {
        // buffer 0's flags of the wave's next 64 cells, one per lane
        const int tq = wave + 4 * (q0 + lane);
        const int fv = tq < nc ? (int)frow[tq] : 0;
        for (int q = q0; q < q0 + 64; q++) {
            const int tn = wave + 4 * q;
            if (tn >= nc) break;
            const int j0 = tn * 64;
            float *col = E.W + (size_t)d * j0 + i;
            if (!rs_block_may_be_nonzero<NM>(nx, nr, tm * 64, j0, 64)) {
                if (__builtin_amdgcn_readlane(fv, q - q0) != 0) {
#pragma unroll 16
                    for (int c = 0; c < 64; c++) col[(size_t)d * c] = 0.0f;
                    if (lane == 0) frow[tn] = 0;
                }
                continue;
            }
            double rs = 0.0;
            unsigned nz = 0;
            float amax = 0.0f;
            for (int c = 0; c < 64; c++) {
                const int j = j0 + c;
                const int xj = NM ? j / nr : j % nx, rj = NM ? j % nr : j / nx;
                double g = 0.0;
                if (ri == rj) {
                    if (xj == xi - 1) g = am;
                    else if (xj == xi + 1) g = ap;
                    else if (xj == xi) g = gd;
                } else if (xi == xj) {
                    g = E.L.q_off;
                }
                float w = (float)(E.dt * g);  // rs_elem<2, NM>(L, dt, i, j)
                const float ws = w * N::S;
                const uint16_t h = N::enc(ws);
                const uint16_t l = LO ? N::enc((ws - N::dec(h)) * N::G) : (uint16_t)0;
                const unsigned nzv = (h | l) & 0x7fffu;
                if (!nzv) w = 0.0f;  // flush rule: both 16-bit parts +-0 -> FP32 +0
                col[(size_t)d * c] = w;
                rs += (double)w;
                nz |= nzv;
                amax = fmaxf(amax, fabsf(w));
            }
            E.part[(size_t)tn * d + i] = rs;
            const bool any = __ballot(nz != 0) != 0;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o, 64));
            if (lane == 0) frow[tn] = any ? cell_flag(amax) : 0;
        }
    }
}

}  // namespace

// ophip_sek_rs_batch (out_kind 2, node_major: 6) into buffer 0 of each
// lattice's state, fused with the first stage of ophip_wsq_exec (the flush
// rule, partial row sums and buffer 0's cell flags); ophip_wsq_exec_ex with
// presplit = 1 then leaves that stage out.  Needs buffer 0's flags to
// describe its contents, as every plan and exec leaves them.
OPHIP_EXPORT int ophip_wsq_assemble_rs(int n, const int *nx, const int *nr, const unsigned *xval,
                                       const unsigned *drift, const unsigned *vol,
                                       const double *q_off, const double *dt, int node_major,
                                       int mode, const unsigned *states) {
    if (n <= 0 || !nx || !nr || !xval || !drift || !vol || !q_off || !dt || !states || mode < 0 ||
        mode > 2)
        return (int)set_status(kErrBadArg);
    for (int k = 0; k < n; k++) {
        RSLattice L;
        const int d = nx[k] * nr[k];
        if (!rs_args(nx[k], nr[k], xval[k], drift[k], vol[k], L) || d % 64 ||
            !handle_range_ok(states[k], wsq_layout(d, mode).total))
            return (int)set_status(kErrBadArg);
    }
    Device &D = dev();
    for (int k0 = 0; k0 < n; k0 += kRSWBatch) {
        const int m = n - k0 < kRSWBatch ? n - k0 : kRSWBatch;
        RSWBatch b{};
        int most = 0;
        for (int k = 0; k < m; k++) {
            RSWEntry &E = b.e[k];
            rs_args(nx[k0 + k], nr[k0 + k], xval[k0 + k], drift[k0 + k], vol[k0 + k], E.L);
            E.L.q_off = q_off[k0 + k];
            E.dt = dt[k0 + k];
            const int d = nx[k0 + k] * nr[k0 + k];
            const WsqLayout Y = wsq_layout(d, mode);
            E.W = P<float>(states[k0 + k] + (unsigned)Y.W0);
            E.part = P<double>(states[k0 + k] + (unsigned)Y.part);
            E.fl = P<uint8_t>(states[k0 + k] + (unsigned)Y.fl[0]);
            most = d / 64 > most ? d / 64 : most;
        }
        const unsigned e = launch_mode(mode, [&](auto bf, auto lo) {
            constexpr bool BF = decltype(bf)::value, LO = decltype(lo)::value;
            if (node_major)
                hipLaunchKernelGGL((k_rs_wsplit<true, BF, LO>), dim3(most, m), dim3(256), 0, D.stream, b);
            else
                hipLaunchKernelGGL((k_rs_wsplit<false, BF, LO>), dim3(most, m), dim3(256), 0, D.stream, b);
        });
        if (e) return (int)e;
    }
    return 0;
}

// Split + diagonal fix of W0 (buffer 0, written by ophip_sek_rs out_kind
// 2), then the squarings, each followed by the diagonal fix.  nsq >= 0:
// every lattice squares nsq times; nsq < 0: the per-lattice counts given
// to the plan.  Lattice i ends in buffer (its count) & 1.  Kernel launches
// only.  presplit: buffer 0 was written by ophip_wsq_assemble_rs (the
// split is left out).
static int wsq_exec(unsigned long plan, int nsq, bool presplit) {
    Device &D = dev();
    WsqPlan *P = (WsqPlan *)(uintptr_t)plan;
    if (!P) return (int)set_status(kErrBadArg);
    const bool per = nsq < 0;
    const int rounds = per ? P->rounds : nsq;
    const dim3 fixgrid(cdiv(P->dmax, 64), P->n);
    return (int)launch_mode(P->mode, [&](auto bf, auto lo) {
        constexpr bool BF = decltype(bf)::value, LO = decltype(lo)::value;
        // sparse rounds: live-item compaction before each squaring launch;
        // every diagonal fix clears the counters for the next round
        const bool sparse = g_wsq_skip && P->sparse_ok;
        if (!presplit)
            hipLaunchKernelGGL((k_wsplit<BF, LO>), dim3(P->n_split), dim3(64), 0, D.stream, D.base,
                               P->sc_split, P->it_split);
        hipLaunchKernelGGL(k_wfix, fixgrid, dim3(64), 0, D.stream, D.base, P->sc_split,
                           P->n, (const int *)nullptr, 0);
        for (int rr = 0; rr < rounds; rr++) {
            const int src = rr & 1;
            bool any = false;
            for (int c = 1; c >= 0; c--) {
                const int nit = per ? P->n_round[c][rr] : P->n_all[c][src];
                const WItem *its = per ? P->it_round[c][rr] : P->it_all[c][src];
                if (nit == 0) continue;
                any = true;
                // a multiple of 8 keeps item j on XCD j % 8 (xcd_order); the
                // live lists need all 8 XCDs
                int g = nit < P->grid ? nit : P->grid;
                if (g > 8) g -= g % 8;
                if (sparse) g = std::max(g, 8);
                // work items carry s = src*n + i (the full table); the
                // destination's flags cover its nonzero cells (k_wlive)
                const WLive *lv = nullptr;
                const int *cn = nullptr;
                if (sparse) {
                    lv = P->live[c];
                    cn = P->counts + (c ? P->cap[0] / 2 : 0);
                    const int kc = g_wsq_kcut.load(), kthr = kc <= -192 ? 0 : kc + 193;
                    if (c == 1)
                        hipLaunchKernelGGL((k_wlive<2>), dim3(8 * cdiv(cdiv(nit, 8), 16)), dim3(1024), 0, D.stream, D.base,
                                           P->sc, its, nit, P->live[c], P->cap[c], P->counts + (c ? P->cap[0] / 2 : 0), 1, kthr);
                    else
                        hipLaunchKernelGGL((k_wlive<1>), dim3(8 * cdiv(cdiv(nit, 8), 16)), dim3(1024), 0, D.stream, D.base,
                                           P->sc, its, nit, P->live[c], P->cap[c], P->counts + (c ? P->cap[0] / 2 : 0), 1, kthr);
                }
                if (c == 1)
                    hipLaunchKernelGGL((k_wsq<BF, LO, 128, 8>), dim3(g), dim3(512), 0, D.stream, D.base,
                                       P->sc, its, nit, lv, cn, P->cap[c]);
                else
                    hipLaunchKernelGGL((k_wsq<BF, LO, 64, 4>), dim3(g), dim3(256), 0, D.stream, D.base,
                                       P->sc, its, nit, lv, cn, P->cap[c]);
            }
            if (!any) continue;
            // the fix kernel indexes lattices 0..n-1 of the table half for
            // this direction
            const WScen *scn = P->sc + (src ? P->n : 0);
            hipLaunchKernelGGL(k_wfix, fixgrid, dim3(64), 0, D.stream, D.base, scn, P->n,
                               per ? (const int *)P->nsq_dev : (const int *)nullptr, rr);
        }
    });
}

OPHIP_EXPORT int ophip_wsq_exec(unsigned long plan, int nsq) { return wsq_exec(plan, nsq, false); }

// flags: 1 = presplit (buffer 0 written by ophip_wsq_assemble_rs)
OPHIP_EXPORT int ophip_wsq_exec_ex(unsigned long plan, int nsq, int flags) {
    if (flags & ~1) return (int)set_status(kErrBadArg);
    return wsq_exec(plan, nsq, (flags & 1) != 0);
}

// zero-cell skipping of the squaring kernel: 1 on (default), 0 off (every
// K-step visited; the results are the same bits).  Read when kernels are
// launched: a captured graph keeps the setting it was captured with.
// Returns the previous setting.
OPHIP_EXPORT int ophip_wsq_set_skip(int on) {
    const int old = g_wsq_skip;
    g_wsq_skip = on ? 1 : 0;
    return old;
}

// the sparse rounds' K-cell cutoff (log2; -192 or below: off), for rounds
// launched or captured from now on; returns the previous value
OPHIP_EXPORT int ophip_wsq_set_kcut(int log2_bound) {
    return g_wsq_kcut.exchange(log2_bound < -192 ? -192 : log2_bound);
}

// handle of the FP32 W buffer `buf` (0/1) inside a state
OPHIP_EXPORT unsigned ophip_wsq_w(unsigned state, int d, int mode, int buf) {
    WsqLayout L = wsq_layout(d, mode);
    return state + (unsigned)(buf ? L.W1 : L.W0);
}

// handle of one array of a state: which = 0 W of buffer buf (FP32,
// column-major), 2 and 4 the backward layout's fragment-major hi and lo
// parts of W (buf ignored; three-product modes), 5 partial row sums (FP64),
// 6 cell flags of buffer buf (d/64 x d/64 bytes, row-major; 0 = every W
// value of the 64 x 64 cell is +0, otherwise the magnitude bound 96 + E,
// 255 unknown), 7 backward-step tile lists; 0 for 1, 3 and arrays the mode
// does not have
OPHIP_EXPORT unsigned ophip_wsq_array(unsigned state, int d, int mode, int buf, int which) {
    if (d <= 0 || d % 64 || mode < 0 || mode > 2 || buf < 0 || buf > 1 || which < 0 || which > 7)
        return 0;
    WsqLayout L = wsq_layout(d, mode);
    size_t off;
    if (which == 0) off = buf ? L.W1 : L.W0;
    else if (which == 5) off = L.part;
    else if (which == 6) off = L.fl[buf];
    else if (which == 7) off = L.tl;
    else if (which == 2) off = L.Fh;
    else if (which == 4) off = L.Fl;
    else off = (size_t)-1;  // 1, 3: no such array
    if (off == (size_t)-1) return 0;
    return state + (unsigned)off;
}

// ---------------------------------------------------------------------
// Autotuner: a tile per size bucket, then the persistent grid (timed with
// the dense K loop).
// Each bucket (the lattices of one size d) is planned alone with 64- and
// with 128-tiles and the faster one is kept; the combined plan then tries
// 1, 2 and 4 workgroups per CU.  Timing runs the plan's split and two
// squarings on the states as they are (their contents do not matter to
// the timing and are overwritten by the caller's next W0).  Results are
// cached per (mode, d, lattice count) and per full size signature.
// ---------------------------------------------------------------------
namespace {

std::mutex g_tune_mu;
std::map<std::array<int, 3>, int> g_tile_cache;                 // (mode, d, count) -> tile
std::map<std::vector<int>, int> g_grid_cache;                   // (mode, d..., tile...) -> grid

float time_plan(WsqPlan *P, int reps) {
    const unsigned long h = (unsigned long)(uintptr_t)P;
    if (ophip_wsq_exec(h, 2)) return -1.0f;  // warm-up
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0f;
    }
    float ms = -1.0f;
    hipStream_t s = dev().stream;
    if (hipEventRecord(a, s) == hipSuccess) {
        bool ok = true;
        for (int r = 0; r < reps && ok; r++) ok = ophip_wsq_exec(h, 2) == 0;
        if (ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess)
            (void)hipEventElapsedTime(&ms, a, b);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace

// tiles[i] (64/128) and *grid for a plan of these lattices; returns 0 or
// a status.  Not graph-capturable (it synchronizes).
OPHIP_EXPORT int ophip_wsq_autotune(int n, const unsigned *states, const int *ds, const int *nsqs,
                                    int mode, int *tiles, int *grid) {
    if (n <= 0 || n > kMaxScen || !states || !ds || !tiles || !grid || mode < 0 || mode > 2)
        return (int)set_status(kErrBadArg);
    (void)nsqs;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    // time the dense K loop: the states' contents are arbitrary here, and
    // their cell flags would make the timing depend on them
    struct SkipOff {
        int old = ophip_wsq_set_skip(0);
        ~SkipOff() { ophip_wsq_set_skip(old); }
    } skip_off;
    const int ncu = dev().n_cu > 0 ? dev().n_cu : 256;
    std::map<int, std::vector<int>> buckets;
    for (int i = 0; i < n; i++) buckets[ds[i]].push_back(i);
    for (auto &kv : buckets) {
        const int d = kv.first;
        const std::vector<int> &ids = kv.second;
        const std::array<int, 3> key = {mode, d, (int)ids.size()};
        int best = 0;
        auto it = g_tile_cache.find(key);
        if (it != g_tile_cache.end()) {
            best = it->second;
        } else {
            std::vector<unsigned> st;
            std::vector<int> dd;
            for (int i : ids) {
                st.push_back(states[i]);
                dd.push_back(ds[i]);
            }
            float tbest = -1.0f;
            for (int T : {128, 64}) {
                std::vector<int> tt(ids.size(), T);
                WsqPlan *P = make_plan((int)ids.size(), st.data(), dd.data(), nullptr, mode, tt.data(), 0);
                if (!P) return (int)opcuda_get_status();
                const float ms = time_plan(P, 3);
                ophip_wsq_plan_free((unsigned long)(uintptr_t)P);
                if (ms < 0.0f) return (int)set_status(kErrUnsupported);
                if (tbest < 0.0f || ms < tbest) {
                    tbest = ms;
                    best = T;
                }
            }
            g_tile_cache[key] = best;
        }
        for (int i : ids) tiles[i] = best;
    }
    std::vector<int> sig = {mode};
    for (int i = 0; i < n; i++) sig.push_back(ds[i]);
    for (int i = 0; i < n; i++) sig.push_back(tiles[i]);
    auto git = g_grid_cache.find(sig);
    if (git != g_grid_cache.end()) {
        *grid = git->second;
        return (int)set_status(kOk);
    }
    float gbest = -1.0f;
    int g_best = 2 * ncu;
    for (int g : {2 * ncu, ncu, 4 * ncu}) {
        WsqPlan *P = make_plan(n, states, ds, nullptr, mode, tiles, g);
        if (!P) return (int)opcuda_get_status();
        const float ms = time_plan(P, 3);
        ophip_wsq_plan_free((unsigned long)(uintptr_t)P);
        if (ms < 0.0f) return (int)set_status(kErrUnsupported);
        if (gbest < 0.0f || ms < gbest) {
            gbest = ms;
            g_best = g;
        }
    }
    g_grid_cache[sig] = g_best;
    *grid = g_best;
    return (int)set_status(kOk);
}

// ---------------------------------------------------------------------
// W-form backward induction (modes 1 and 2: the layout makes lo parts)
// ---------------------------------------------------------------------
namespace {
template <class F> unsigned launch_bwd(int mode, int ncp, F f) {
    auto nt = [&](auto bf) {
        switch (ncp) {
        case 16: f(bf, std::integral_constant<int, 1>()); break;
        case 32: f(bf, std::integral_constant<int, 2>()); break;
        default: f(bf, std::integral_constant<int, 3>()); break;
        }
    };
    if (mode == 1) nt(std::integral_constant<bool, false>());
    else nt(std::integral_constant<bool, true>());
    return check_launch();
}
bool bwd_args_ok(int d, int ncp, int mode) {
    return d > 0 && d % 128 == 0 && d <= 4096 && (mode == 1 || mode == 2) &&
           (ncp == 16 || ncp == 32 || ncp == 48);
}
}  // namespace

// XT (2 * ncp * d 16-bit values: hi then lo, fragment-major) from X
// (d x ncp row-major FP32): the first step's operand copies
OPHIP_EXPORT int ophip_wsq_bwd_prep(int d, int ncp, int mode, unsigned X, unsigned XT) {
    if (!bwd_args_ok(d, ncp, mode)) return (int)set_status(kErrBadArg);
    const size_t n = (size_t)d * ncp;
    if (!handle_range_ok(X, 4 * n) || !handle_range_ok(XT, 4 * n)) return (int)set_status(kErrBadHandle);
    Device &D = dev();
    const float *x = P<const float>(X);
    uint16_t *xt = P<uint16_t>(XT);
    return (int)launch_bwd(mode, ncp, [&](auto bf, auto nt) {
        constexpr bool BF = decltype(bf)::value;
        constexpr int NT = decltype(nt)::value;
        hipLaunchKernelGGL((k_wbwd_prep<BF, NT>), dim3(cdiv(d / 8 * 16 * NT, 256)), dim3(256), 0,
                           D.stream, d, x, xt, xt + n);
    });
}

namespace {
// one lattice's entry of a batched layout: W of buffer buf, the state's
// fragment arrays
LayScen lay_scen(const WsqLayout &L, unsigned state, int buf) {
    LayScen S;
    S.W = state + (buf ? L.W1 : L.W0);
    S.Hf = state + L.Fh;
    S.Lf = state + L.Fl;
    S.flags = state + L.tf;
    S.tl = state + L.tl;
    S.cf = state + L.fl[buf];
    return S;
}

// the tile lists of m lattices' fragment copies (after their layout), from
// the hi copies; pmode, nx, nr: the fragments' state order (k_wbwd_flags)
void bwd_tile_lists(Device &D, int d, int mode, const LayBatch &bt, int m, int pmode, int nx, int nr) {
    const int waves = d / 16 * (d / 128);
    const int ev = g_cut_visit.load(), el = g_cut_lo.load();
    const uint32_t tvis = hi_bits(mode, ev), tlo = hi_bits(mode, mode == 2 ? el - 3 : el);
    hipLaunchKernelGGL(k_wbwd_flags, dim3(cdiv(waves, 4), m), dim3(256), 0, D.stream, d, D.base, bt, tvis,
                       tlo, pmode, nx, nr);
    hipLaunchKernelGGL(k_wbwd_lists, dim3(cdiv(d / 16 + d / 64 + d / 32, 4), m), dim3(256), 0, D.stream, d,
                       D.base, bt);
}

// the same for one lattice
void bwd_tile_lists(Device &D, const WsqLayout &L, unsigned state, int d, int mode, int buf, int pmode) {
    LayBatch bt{};
    bt.s[0] = lay_scen(L, state, buf);
    bwd_tile_lists(D, d, mode, bt, 1, pmode, 0, 0);
}
}  // namespace

// W of buffer `buf` split into its hi/lo parts (as the squarings split
// it), fragment-major, into the state's fragment arrays; once per solve,
// before the steps
OPHIP_EXPORT int ophip_wsq_bwd_layout(unsigned state, int d, int mode, int buf) {
    if (!bwd_args_ok(d, 16, mode) || buf < 0 || buf > 1) return (int)set_status(kErrBadArg);
    const WsqLayout L = wsq_layout(d, mode);
    if (!handle_range_ok(state, L.total)) return (int)set_status(kErrBadHandle);
    Device &D = dev();
    const float *W = P<const float>(state + (unsigned)(buf ? L.W1 : L.W0));
    uint16_t *Hf = P<uint16_t>(state + (unsigned)L.Fh);
    uint16_t *Lf = P<uint16_t>(state + (unsigned)L.Fl);
    const dim3 grid((unsigned)(((size_t)d * d / 8 + 255) / 256));
    if (mode == 2)
        hipLaunchKernelGGL((k_wfrag<true>), grid, dim3(256), 0, D.stream, d, W, Hf, Lf);
    else
        hipLaunchKernelGGL((k_wfrag<false>), grid, dim3(256), 0, D.stream, d, W, Hf, Lf);
    bwd_tile_lists(D, L, state, d, mode, buf, 1);
    return (int)check_launch();
}

// ophip_wsq_bwd_layout with the states reordered: fragment state i is W's
// state perm[i] (perm: handle of d ints, a permutation of 0..d-1),
// so the steps run in the order of perm's domain.  The step arithmetic is
// the same in every order, its FP32 accumulation is not: a lattice squared
// in a band order steps in its natural order.
OPHIP_EXPORT int ophip_wsq_bwd_layout_perm(unsigned state, int d, int mode, int buf, unsigned perm) {
    if (!bwd_args_ok(d, 16, mode) || buf < 0 || buf > 1) return (int)set_status(kErrBadArg);
    const WsqLayout L = wsq_layout(d, mode);
    if (!handle_range_ok(state, L.total) || !handle_range_ok(perm, 4 * (size_t)d))
        return (int)set_status(kErrBadHandle);
    Device &D = dev();
    const float *W = P<const float>(state + (unsigned)(buf ? L.W1 : L.W0));
    uint16_t *Hf = P<uint16_t>(state + (unsigned)L.Fh);
    uint16_t *Lf = P<uint16_t>(state + (unsigned)L.Fl);
    const dim3 grid((unsigned)(((size_t)d * d / 8 + 255) / 256));
    if (mode == 2)
        hipLaunchKernelGGL((k_wfragp<true>), grid, dim3(256), 0, D.stream, d, W, P<const int>(perm), Hf, Lf);
    else
        hipLaunchKernelGGL((k_wfragp<false>), grid, dim3(256), 0, D.stream, d, W, P<const int>(perm), Hf, Lf);
    bwd_tile_lists(D, L, state, d, mode, buf, 0);
    return (int)check_launch();
}

namespace {
// backward-step kernel: 0 dense k_wbwd1, 1 automatic, 2 k_wbwd_sp (16-row
// tiles), 3 k_wbwd_blk with 64-row blocks, 4 with 32-row blocks
int g_bwd_kernel = 1;
// list entries each wave of the 32-row-block step has in flight
#ifndef OPHIP_BWD_E
#define OPHIP_BWD_E 2
#endif
}  // namespace

OPHIP_EXPORT int ophip_wsq_bwd_step(unsigned state, int d, int mode, int buf, int ncp, float df,
                                    unsigned X, unsigned XT, unsigned Y, unsigned YT);

namespace {
// the kernel a batched step of nb lattices uses (2, 3 or 4; 0 = dense)
int bwd_kernel_for(int nb, int d, int ncp) {
    if (g_bwd_kernel == 0) return 0;
    if (ncp != 48) return 2;
    if (g_bwd_kernel != 1) return g_bwd_kernel;
    return std::min(nb, kBwdBatch) * (d / 32) >= 256 ? 4 : 2;
}
}  // namespace

// the backward-step kernel the current setting gives nb lattices of size d
// with ncp value columns: 0 dense, 2 16-row tiles, 3 64-row blocks, 4
// 32-row blocks
OPHIP_EXPORT int ophip_wsq_bwd_kernel_for(int nb, int d, int ncp) { return bwd_kernel_for(nb, d, ncp); }

OPHIP_EXPORT int ophip_wsq_set_bwd_kernel(int k) {
    const int old = g_bwd_kernel;
    if (k >= 0 && k <= 4) g_bwd_kernel = k;
    return old;
}

// the band-sparse step's cutoffs (exponents, -40 <= visit <= lo <= -1) for
// the tile lists built (or captured) from now on
OPHIP_EXPORT int ophip_wsq_set_bwd_cutoffs(int visit_exp, int lo_exp) {
    if (visit_exp < -40 || visit_exp > lo_exp || lo_exp > -1) return (int)set_status(kErrBadArg);
    g_cut_visit.store(visit_exp);
    g_cut_lo.store(lo_exp);
    return 0;
}

OPHIP_EXPORT int ophip_wsq_get_bwd_cutoffs(int *visit_exp, int *lo_exp) {
    if (!visit_exp || !lo_exp) return (int)set_status(kErrBadArg);
    *visit_exp = g_cut_visit.load();
    *lo_exp = g_cut_lo.load();
    return 0;
}

// One backward step of nb lattices of the same d in one launch per 32:
// lattice i steps Y[i] = df[i] (X[i] + W_i X[i]) with W_i laid out by
// ophip_wsq_bwd_layout(_perm)(states[i], .., buf), visiting the nonzero
// fragments of its tile lists only.
OPHIP_EXPORT int ophip_wsq_bwd_step_batch(int nb, const unsigned *states, int d, int mode, int buf,
                                          int ncp, const float *df, const unsigned *X,
                                          const unsigned *XT, const unsigned *Y, const unsigned *YT) {
    if (!bwd_args_ok(d, ncp, mode) || buf < 0 || buf > 1 || nb <= 0 || !states || !df || !X || !XT ||
        !Y || !YT)
        return (int)set_status(kErrBadArg);
    const WsqLayout L = wsq_layout(d, mode);
    const size_t n = (size_t)d * ncp;
    for (int i = 0; i < nb; i++)
        if (!handle_range_ok(states[i], L.total) || (X[i] && !handle_range_ok(X[i], 4 * n)) ||
            !handle_range_ok(XT[i], 4 * n) || (Y[i] && !handle_range_ok(Y[i], 4 * n)) ||
            !handle_range_ok(YT[i], 4 * n))
            return (int)set_status(kErrBadHandle);
    if (g_bwd_kernel == 0) {  // the dense single-lattice kernel, one launch each
        for (int i = 0; i < nb; i++) {
            const int e = ophip_wsq_bwd_step(states[i], d, mode, buf, ncp, df[i], X[i], XT[i], Y[i], YT[i]);
            if (e) return e;
        }
        return 0;
    }
    Device &D = dev();
    const int nct = ncp / 16;
    // automatic: 32-row blocks once the lattices give a workgroup per CU
    // or more (measured at d = 2048: 1.89 us per lattice and step for 16
    // lattices, against 2.30 us with 64-row blocks and 5.1 us with 16-row
    // tiles; 3 tiles in flight per wave 2.27 us, 4 waves with 4 tiles in
    // flight 1.99 us), 16-row tiles (more, smaller workgroups) below that
    const int kern = bwd_kernel_for(nb, d, ncp);
    for (int i0 = 0; i0 < nb; i0 += kBwdBatch) {
        const int m = std::min(kBwdBatch, nb - i0);
        BwdBatch bt{};
        for (int i = 0; i < m; i++) {
            const unsigned st = states[i0 + i];
            BwdScen &S = bt.s[i];
            S.Hf = st + L.Fh;
            S.Lf = st + L.Fl;
            S.tl = st + L.tl;
            S.X = X[i0 + i];
            S.XT = XT[i0 + i];
            S.Y = Y[i0 + i];
            S.YT = YT[i0 + i];
            S.df = df[i0 + i];
        }
        if (kern == 2) {
            const dim3 grid((unsigned)(d / 16 * nct), (unsigned)m);
            if (mode == 1)
                hipLaunchKernelGGL((k_wbwd_sp<false, 16, 3>), grid, dim3(1024), 0, D.stream, D.base, bt, d, nct);
            else
                hipLaunchKernelGGL((k_wbwd_sp<true, 16, 3>), grid, dim3(1024), 0, D.stream, D.base, bt, d, nct);
        } else {
            const int rows = kern == 3 ? 64 : 32, per_xcd = cdiv(m * (d / rows), 8);
            const dim3 grid((unsigned)(8 * per_xcd));
            if (kern == 3 && mode == 1)
                hipLaunchKernelGGL((k_wbwd_blk<false, 4, 4, 2>), grid, dim3(256), 0, D.stream, D.base, bt, m, d,
                                   per_xcd);
            else if (kern == 3)
                hipLaunchKernelGGL((k_wbwd_blk<true, 4, 4, 2>), grid, dim3(256), 0, D.stream, D.base, bt, m, d,
                                   per_xcd);
            else if (mode == 1)
                hipLaunchKernelGGL((k_wbwd_blk<false, 2, 8, OPHIP_BWD_E>), grid, dim3(512), 0, D.stream, D.base, bt, m, d,
                                   per_xcd);
            else
                hipLaunchKernelGGL((k_wbwd_blk<true, 2, 8, OPHIP_BWD_E>), grid, dim3(512), 0, D.stream, D.base, bt, m, d,
                                   per_xcd);
        }
        const unsigned e = check_launch();
        if (e) return (int)e;
    }
    return 0;
}

// ophip_wsq_bwd_layout_perm for the regime permutation perm[x + nx r] =
// nr x + r (a lattice squared node-major, stepped in the reference order)
// without a permutation array: nr in {1, 2, 4, 8}, nx % 8 == 0, nx nr = d.
OPHIP_EXPORT int ophip_wsq_bwd_layout_regime_batch(int n, const unsigned *states, int d, int mode, int buf,
                                                   int nx, int nr);
OPHIP_EXPORT int ophip_wsq_bwd_layout_regime(unsigned state, int d, int mode, int buf, int nx, int nr) {
    return ophip_wsq_bwd_layout_regime_batch(1, &state, d, mode, buf, nx, nr);
}

// ophip_wsq_bwd_layout_regime for n lattices of the same d, nx and nr in
// one layout launch and one launch each for the flags and the lists (per
// 32 lattices)
OPHIP_EXPORT int ophip_wsq_bwd_layout_regime_batch(int n, const unsigned *states, int d, int mode, int buf,
                                                   int nx, int nr) {
    if (!bwd_args_ok(d, 16, mode) || buf < 0 || buf > 1 || nx <= 0 || nx % 8 || nx * nr != d ||
        (nr != 1 && nr != 2 && nr != 4 && nr != 8) || n <= 0 || !states)
        return (int)set_status(kErrBadArg);
    const WsqLayout L = wsq_layout(d, mode);
    for (int i = 0; i < n; i++)
        if (!handle_range_ok(states[i], L.total)) return (int)set_status(kErrBadHandle);
    Device &D = dev();
    for (int i0 = 0; i0 < n; i0 += kLayBatch) {
        const int m = std::min(kLayBatch, n - i0);
        LayBatch bt{};
        for (int i = 0; i < m; i++) bt.s[i] = lay_scen(L, states[i0 + i], buf);
        const dim3 grid((unsigned)((d / 64) * (d / 64)), m);
        if (mode == 2)
            hipLaunchKernelGGL((k_wfragk<true>), grid, dim3(256), 0, D.stream, d, nx, nr, D.base, bt);
        else
            hipLaunchKernelGGL((k_wfragk<false>), grid, dim3(256), 0, D.stream, d, nx, nr, D.base, bt);
        bwd_tile_lists(D, d, mode, bt, m, 2, nx, nr);
        const unsigned e = check_launch();
        if (e) return (int)e;
    }
    return 0;
}

// Y = df (X + W X) with W laid out by ophip_wsq_bwd_layout(state, .., buf),
// and Y's copies YT for the next step.  X/XT and Y/YT must not overlap.
OPHIP_EXPORT int ophip_wsq_bwd_step(unsigned state, int d, int mode, int buf, int ncp, float df,
                                    unsigned X, unsigned XT, unsigned Y, unsigned YT) {
    if (!bwd_args_ok(d, ncp, mode) || buf < 0 || buf > 1) return (int)set_status(kErrBadArg);
    const WsqLayout L = wsq_layout(d, mode);
    const size_t n = (size_t)d * ncp;
    if (!handle_range_ok(state, L.total) || (X && !handle_range_ok(X, 4 * n)) || !handle_range_ok(XT, 4 * n) ||
        (Y && !handle_range_ok(Y, 4 * n)) || !handle_range_ok(YT, 4 * n))
        return (int)set_status(kErrBadHandle);
    if (g_bwd_kernel) return ophip_wsq_bwd_step_batch(1, &state, d, mode, buf, ncp, &df, &X, &XT, &Y, &YT);
    Device &D = dev();
    const uint16_t *Hf = P<const uint16_t>(state + (unsigned)L.Fh);
    const uint16_t *Lf = P<const uint16_t>(state + (unsigned)L.Fl);
    const uint16_t *xt = P<const uint16_t>(XT);
    const float *x = P<const float>(X);
    float *y = P<float>(Y);
    uint16_t *yt = P<uint16_t>(YT);
    return (int)launch_bwd(mode, ncp, [&](auto bf, auto nt) {
        constexpr bool BF = decltype(bf)::value;
        constexpr int NT = decltype(nt)::value;
        if (d % 256 == 0)
            hipLaunchKernelGGL((k_wbwd1<BF, NT, 8>), dim3(d / 16), dim3(512), 0, D.stream, d, df, Hf, Lf,
                               xt, xt + n, x, y, yt, yt + n);
        else
            hipLaunchKernelGGL((k_wbwd1<BF, NT, 4>), dim3(d / 16), dim3(256), 0, D.stream, d, df, Hf, Lf,
                               xt, xt + n, x, y, yt, yt + n);
    });
}
