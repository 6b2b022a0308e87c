"""GPU parity tests of every drop-in entry point, called through the C-ABI
(libophip.so via ctypes) and checked against the CPU oracle.

Bars: bit-exact for integer work (MT19937 streams, sampled states) and for
kernels whose FP32 operation order is the reference's (sek, sck); for FP32
contractions the reference's own self-check tolerance |a-b|/(1+|a|)
(2e-3 GEMM, gsgemm4/Program.cs:283-291; 1e-3 GEMV, gsgemv4/Program.cs:
384-391) and a tighter scale-relative bound stated per test.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def H(hip):
    return hip


def dev(H, a, order="F"):
    return H.DeviceArray.from_numpy(a, order=order)


def ref_tol(a, b):
    """the reference's self-check metric |a-b|/(1+|a|)"""
    return float(np.max(np.abs(a - b) / (1.0 + np.abs(a))))


# ---- memory ----------------------------------------------------------------

def test_alloc_copy_roundtrip_and_handles(H):
    banner("alloc/copy")
    a = np.arange(1000, dtype=np.float32)
    d = dev(H, a)
    assert d.h != 0 and d.h % 256 == 0
    assert np.array_equal(d.download(), a)
    # handle arithmetic, as the C# layer does (ptr + offset bytes)
    out = np.zeros(10, dtype=np.float32)
    H.check(H.lib.opcuda_memcpy_d2h(d.h + 4 * 100, out.ctypes.data, 40))
    assert np.array_equal(out, a[100:110])
    e = H.DeviceArray.empty(1000)
    H.check(H.lib.opcuda_memcpy_d2d(e.h, d.h, 4000))
    assert np.array_equal(e.download(), a)
    # copying past the allocation is refused, not performed
    st = H.lib.opcuda_memcpy_d2h(d.h, out.ctypes.data, 8000)
    assert st != 0
    d.free()
    e.free()


def test_fill_and_copies_do_not_overrun(H):
    banner("sset/scopy/scopy1")
    n = 1000 + 37
    guard = 256
    buf = dev(H, np.full(n + guard, -7.0, dtype=np.float32))
    H.lib.opcuda_ssetall(buf.h, n, np.float32(2.5), 1)
    H.check()
    H.lib.opcuda_ssetone(buf.h, n, np.float32(9.0), 5)
    H.check()
    got = buf.download()
    want = np.full(n + guard, -7.0, dtype=np.float32)
    want[:n] = 2.5
    want[5] = 9.0
    assert np.array_equal(got, want)
    src = dev(H, np.arange(n, dtype=np.float32))
    dst = dev(H, np.full(n + guard, -1.0, dtype=np.float32))
    H.check(H.lib.opcuda_scopy1(dst.h, src.h, n))
    got = dst.download()
    assert np.array_equal(got[:n], np.arange(n, dtype=np.float32))
    assert np.all(got[n:] == -1.0)  # global_scopy1 overran here (scopy2.cu:91)
    H.lib.opcuda_scopy(100, src.h, 3, dst.h, 2)
    H.check()
    got = dst.download()
    assert np.array_equal(got[0:200:2], np.arange(0, 300, 3, dtype=np.float32))


# ---- first step / generator -------------------------------------------------

def test_sek_bit_exact(H):
    banner("sek")
    rng = np.random.default_rng(10)
    d = 100
    g = rng.standard_normal((d, d)).astype(np.float32)
    gd = dev(H, g)
    out = H.DeviceArray.empty((d, d))
    H.lib.opcuda_sek(d, gd.h, np.float32(1e-3), out.h)
    H.check()
    ref = np.zeros(d * d, dtype=np.float32)
    O.lib.or_sek_f32(d, np.ascontiguousarray(g.reshape(-1, order="F")), np.float32(1e-3), ref)
    assert np.array_equal(out.download(), ref.reshape(d, d, order="F"))
    r = rng.random(d).astype(np.float32) * 0.05
    rd = dev(H, r)
    gf = np.ascontiguousarray(g.reshape(-1, order="F"))
    H.lib.opcuda_sekd(d, gd.h, np.float32(1e-3), out.h, rd.h)
    H.check()
    ref = np.zeros(d * d, dtype=np.float32)
    O.lib.or_sekd_f32(d, gf, np.float32(1e-3), r, None, ref)
    assert np.array_equal(out.download(), ref.reshape(d, d, order="F"))
    want = (np.eye(d) + 1e-3 * g.astype(np.float64)) / (1 + 1e-3 * r[:, None].astype(np.float64))
    assert np.max(np.abs(out.download() - want)) < 1e-6
    lamv = np.array([0.5], dtype=np.float32)
    lam = dev(H, lamv)
    H.lib.opcuda_sekd_lambda(d, gd.h, np.float32(1e-3), out.h, rd.h, lam.h)
    H.check()
    O.lib.or_sekd_f32(d, gf, np.float32(1e-3), r, lamv.ctypes.data, ref)
    assert np.array_equal(out.download(), ref.reshape(d, d, order="F"))


def test_sglvg_vs_fp64(H):
    banner("sglvg")
    d, ni = 128, 2
    x, _ = O.xgrid(d, 0.0, 400.0, 100.0)
    inv = O.invm(x)
    dr, vo = O.lv_params_c1(x, 0.02)
    pars = np.concatenate([dr, 1.3 * dr, vo, 0.9 * vo]).astype(np.float32)
    gen = H.DeviceArray.empty(ni * d * d)
    invd, parsd = dev(H, inv.astype(np.float32)), dev(H, pars)  # keep both alive
    H.lib.opcuda_sglvg(gen.h, ni, d, invd.h, parsd.h)
    H.check()
    got = gen.download().reshape(ni, d * d)
    ref32 = np.zeros(ni * d * d, np.float32)
    O.lib.or_sglvg_f32(ni, d, inv.astype(np.float32), pars, ref32)
    assert np.array_equal(got.reshape(-1), ref32)  # the reference's FP32 steps, bit for bit
    for i, (a, b) in enumerate(((dr, vo), (1.3 * dr, 0.9 * vo))):
        ref = O.lv_generator(x, a, b)
        scale = np.abs(ref).max()
        assert np.abs(got[i].reshape(d, d, order="F") - ref).max() / scale < 1e-6


def test_fp64_assembly_and_first_step(H):
    banner("dgen_rs / sek_rs")
    nx, nr = 48, 3
    x, _ = O.xgrid(nx, 0.0, 400.0, 100.0)
    sig = [0.2, 0.35, 0.5]
    ref = O.rs_generator(x, sig, 0.02, 0.5)
    d = nx * nr
    drift = np.tile(0.02 * x, nr)
    vol = np.concatenate([s * x for s in sig])
    xd, dd, vd = dev(H, x), dev(H, drift), dev(H, vol)
    g = H.DeviceArray.empty((d, d), dtype=np.float64)
    H.check(H.lib.ophip_dgen_rs(nx, nr, xd.h, dd.h, vd.h, 0.5, g.h))
    G = g.download()
    assert np.allclose(G, ref, rtol=1e-13, atol=1e-10)
    assert np.abs(G.sum(axis=1)).max() < 1e-9
    dt = 2.0**-20
    U = H.DeviceArray.empty((d, d), dtype=np.float64)
    H.check(H.lib.ophip_sek_rs(nx, nr, xd.h, dd.h, vd.h, 0.5, dt, 1, U.h))
    assert np.allclose(U.download(), O.first_step(ref, dt), rtol=0, atol=1e-15)
    U32 = H.DeviceArray.empty((d, d), dtype=np.float32)
    H.check(H.lib.ophip_sek_rs(nx, nr, xd.h, dd.h, vd.h, 0.5, dt, 0, U32.h))
    # one rounding of the FP64 value
    want = O.first_step(ref, dt).astype(np.float32)
    got = U32.download()
    assert np.max(np.abs(got.astype(np.float64) - want.astype(np.float64))) <= 6e-8


@pytest.mark.parametrize("out_kind", [0, 3, 6])
def test_sek_rs_batch_equals_single(H, out_kind):
    """ophip_sek_rs_batch: three lattices of different size and parameters
    in one launch, every output the bits of its own ophip_sek_rs"""
    banner(f"sek_rs_batch out_kind={out_kind}")
    import ctypes as C
    lats = [(48, 3, [0.2, 0.35, 0.5], 0.5), (64, 2, [0.3, 0.1], 0.8), (100, 1, [0.25], 0.0)]
    dt, dbl = 2.0**-18, out_kind in (1, 3, 5, 7)
    ty = np.float64 if dbl else np.float32
    keep, args = [], []
    for nx, nr, sig, q in lats:
        x, _ = O.xgrid(nx, 0.0, 400.0, 100.0)
        drift = np.tile(0.02 * x, nr)
        vol = np.concatenate([s * x for s in sig])
        xd, dd, vd = dev(H, x), dev(H, drift), dev(H, vol)
        d = nx * nr
        a, b = H.DeviceArray.empty((d, d), dtype=ty), H.DeviceArray.empty((d, d), dtype=ty)
        keep += [xd, dd, vd, a, b]
        args.append((nx, nr, xd.h, dd.h, vd.h, q, a, b))
    for nx, nr, xh, dh, vh, q, a, _ in args:
        H.check(H.lib.ophip_sek_rs(nx, nr, xh, dh, vh, q, dt, out_kind, a.h))
    n = len(args)

    def arr(t, i):
        return (t * n)(*[r[i] for r in args])
    H.check(H.lib.ophip_sek_rs_batch(n, arr(C.c_int, 0), arr(C.c_int, 1), arr(C.c_uint, 2),
                                     arr(C.c_uint, 3), arr(C.c_uint, 4), arr(C.c_double, 5),
                                     (C.c_double * n)(*[dt] * n), out_kind,
                                     (C.c_uint * n)(*[r[7].h for r in args])))
    H.sync()
    for r in args:
        assert np.array_equal(r[6].download(), r[7].download())


# ---- GEMM ---------------------------------------------------------------------

@pytest.mark.parametrize("m,n,k,lda,ldb,ldc,alpha,beta", [
    (256, 256, 256, 256, 256, 256, 1.0, 0.0),      # fast path, 64-tiles
    (2048, 2048, 2048, 2048, 2048, 2048, 1.0, 0.0),  # fast path, 128-tiles
    (577, 301, 129, 600, 130, 580, 0.5, -1.25),    # generic path, odd shapes
    (64, 1, 64, 64, 64, 64, 1.0, 0.0),
    (1, 33, 7, 1, 7, 1, 2.0, 1.0),
])
def test_sgemm(H, m, n, k, lda, ldb, ldc, alpha, beta):
    banner(f"sgemm {m}x{n}x{k}")
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((lda, k)).astype(np.float32)
    B = rng.standard_normal((ldb, n)).astype(np.float32)
    Cm = rng.standard_normal((ldc, n)).astype(np.float32)
    Ad, Bd, Cd = dev(H, A), dev(H, B), dev(H, Cm)
    H.lib.opcuda_sgemm(m, n, k, np.float32(alpha), Ad.h, lda, Bd.h, ldb, np.float32(beta), Cd.h, ldc)
    H.check()
    got = Cd.download()
    want = alpha * (A[:m, :k].astype(np.float64) @ B[:k, :n].astype(np.float64))
    if beta != 0.0:
        want += beta * Cm[:m, :n].astype(np.float64)
    assert ref_tol(want, got[:m, :n]) < 2e-3
    # FP32 accumulation error bound ~ sqrt(k) ulp of the row norms
    scale = np.sqrt(k) * np.abs(A[:m]).max() * np.abs(B[:k]).max() * abs(alpha) + 1
    assert np.max(np.abs(want - got[:m, :n])) < 1e-5 * scale
    # padding rows of C are untouched
    if ldc > m:
        assert np.array_equal(got[m:, :n], Cm[m:, :n])


def test_ssqmm_pointer_tables(H):
    banner("ssqmm")
    rng = np.random.default_rng(5)
    d, ni = 256, 3
    mats = rng.random((ni, d, d)).astype(np.float32) / d
    buf = H.DeviceArray.empty(3 * ni * d * d)
    size = 4 * d * d
    Ah = [buf.h + size * i for i in range(ni)]
    Bh = [buf.h + size * (ni + (ni - 1 - i)) for i in range(ni)]
    Ch = [buf.h + size * (2 * ni + i) for i in range(ni)]
    for i in range(ni):
        buf.upload(np.ascontiguousarray(mats[i].reshape(-1, order="F")), offset=size * i)
        buf.upload(np.ascontiguousarray(mats[i].T.reshape(-1, order="F")), offset=size * (ni + ni - 1 - i))
    tabs = [dev(H, np.array(t, dtype=np.uint32)) for t in (Ah, Bh, Ch)]
    H.lib.opcuda_ssqmm(d, ni, tabs[0].h, tabs[1].h, tabs[2].h)
    H.check()
    out = buf.download(count=3 * ni * d * d).reshape(3 * ni, d * d)
    for i in range(ni):
        got = out[2 * ni + i].reshape(d, d, order="F")
        want = mats[i].astype(np.float64) @ mats[i].T.astype(np.float64)
        assert ref_tol(want, got) < 2e-3
        assert np.abs(want - got).max() < 1e-6


def test_sgemm4_descriptors(H):
    banner("sgemm4")
    rng = np.random.default_rng(6)
    shapes = [(128, 48, 64), (70, 20, 33), (192, 64, 128)]
    mats, descs = [], []
    for (m, n, k) in shapes:
        A = rng.standard_normal((m, k)).astype(np.float32)
        B = rng.standard_normal((k, n)).astype(np.float32)
        mats.append((dev(H, A), dev(H, B), H.DeviceArray.empty((m, n)), A, B))
    # descriptor block {i,bx,by,m,n,k,A,B,C} per 64x16 tile (gsgemm4/Program.cs:186-205)
    words = []
    for i, ((m, n, k), (Ad, Bd, Cd, _, _)) in enumerate(zip(shapes, mats)):
        for bx in range((m + 63) // 64):
            for by in range((n + 15) // 16):
                words.append([i, bx, by, m, n, k, Ad.h, Bd.h, Cd.h])
    nb = len(words)
    arg = dev(H, np.array(words, dtype=np.uint32).reshape(-1))
    ptrs = dev(H, np.array([arg.h + 36 * b for b in range(nb)], dtype=np.uint32))
    H.lib.opcuda_sgemm4(nb, ptrs.h)
    H.check()
    for (Ad, Bd, Cd, A, B) in mats:
        want = A.astype(np.float64) @ B.astype(np.float64)
        got = Cd.download()
        assert ref_tol(want, got) < 2e-3
        assert np.abs(want - got).max() < 1e-4


def test_strided_batched(H):
    banner("strided batched")
    rng = np.random.default_rng(7)
    b, m, n, k = 4, 128, 256, 64
    A = rng.standard_normal((b, k, m)).astype(np.float32)  # each A_i stored F-order (m x k)
    B = rng.standard_normal((b, n, k)).astype(np.float32)
    Ad = dev(H, A.reshape(-1), order="C")
    Bd = dev(H, B.reshape(-1), order="C")
    Cd = H.DeviceArray.empty(b * m * n)
    H.check(H.lib.ophip_sgemm_strided_batched(m, n, k, np.float32(1), Ad.h, m, m * k, Bd.h, k,
                                               k * n, np.float32(0), Cd.h, m, m * n, b))
    got = Cd.download().reshape(b, n, m)
    for i in range(b):
        want = A[i].T.astype(np.float64) @ B[i].T.astype(np.float64)
        assert np.abs(got[i].T - want).max() < 1e-4


# ---- GEMV -----------------------------------------------------------------------

@pytest.mark.parametrize("m,n,lda,incx,incy,alpha,beta", [
    (2048, 2048, 2048, 1, 1, 1.0, 0.0),
    (300, 77, 301, 2, 3, 0.5, 2.0),
    (21, 256, 21, 256, 1, 0.9, 0.0),
])
def test_sgemv_n_and_t(H, m, n, lda, incx, incy, alpha, beta):
    banner(f"sgemv {m}x{n}")
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((lda, n)).astype(np.float32)
    for trans in (False, True):
        xl = (n if not trans else m)
        yl = (m if not trans else n)
        x = rng.standard_normal(xl * abs(incx)).astype(np.float32)
        y = rng.standard_normal(yl * abs(incy)).astype(np.float32)
        Ad, xd, yd = dev(H, A), dev(H, x), dev(H, y)
        fn = H.lib.opcuda_sgemvT if trans else H.lib.opcuda_sgemv
        fn(m, n, np.float32(alpha), Ad.h, lda, xd.h, incx, np.float32(beta), yd.h, incy)
        H.check()
        Am = A[:m].astype(np.float64)
        xv = x[::incx][:xl].astype(np.float64)
        want = alpha * ((Am.T if trans else Am) @ xv) + beta * y[::incy][:yl]
        got = yd.download()[::incy][:yl]
        assert ref_tol(want, got) < 1e-3
        assert np.abs(want - got).max() < 2e-5 * (np.sqrt(max(m, n)) * 4 + 1)


def test_sgemv3_gather_and_sentinel(H):
    banner("sgemv3")
    rng = np.random.default_rng(8)
    d, nz = 192, 8
    A = (rng.random((d, d)) / d).astype(np.float32)
    Z = rng.random((d, 2 * nz)).astype(np.float32)  # columns 0..nz-1 "from", nz.. "to"
    sentinel = 2 * nz * d
    ncol = 32
    col0 = np.zeros(ncol, dtype=np.int32)
    col1 = np.full(ncol, sentinel, dtype=np.int32)
    # group 0: 5 live columns then a sentinel; group 1: 16 live columns
    live0 = [(3, 8), (0, 9), (7, 10), (3, 11), (1, 12)]
    for c, (a, b) in enumerate(live0):
        col0[c], col1[c] = a * d, b * d
    col1[5] = sentinel
    col0[6], col1[6] = 2 * d, 15 * d  # after a sentinel: must not be written
    col0[16], col1[16] = 4 * d, 13 * d
    col0[17], col1[17] = 5 * d, 14 * d
    col1[18:] = sentinel + 5
    Zd = dev(H, Z)
    Ad, c0d, c1d = dev(H, A), dev(H, col0), dev(H, col1)  # keep alive over the call
    H.lib.opcuda_sgemv3(d, nz, ncol, Ad.h, Zd.h, c0d.h, c1d.h)
    H.check()
    got = Zd.download()
    want = Z.copy().reshape(-1, order="F")
    O.lib.or_sgemv_gather(d, np.ascontiguousarray(A.reshape(-1, order="F")), want, 16,
                          col0[:16].copy(), col1[:16].copy(), sentinel)
    O.lib.or_sgemv_gather(d, np.ascontiguousarray(A.reshape(-1, order="F")), want, 16,
                          col0[16:].copy(), col1[16:].copy(), sentinel)
    want = want.reshape(d, 2 * nz, order="F")
    assert np.abs(got - want).max() < 1e-5
    assert np.array_equal(got[:, 15], Z[:, 15])  # the column after the sentinel


def gsgemv4_tables(H, rng, d, nz, ncols):
    """device buffers in the layout gsgemv4/Program.cs:170-300 builds: Z (2 nz
    vectors of d) then the ni matrices in one float buffer; in one int
    buffer the nblocks descriptor handles, the 10-uint descriptors
    {i, bx, by, d, nz, ncol, A, Z, col0, col1} and per matrix its col0 and
    col1 lists padded to a multiple of 16 with the sentinel 2 nz d; every
    vector index used once (input or output)"""
    ni = len(ncols)
    vals = (0.01 * rng.integers(-1000, 1000, ni * d * d + 2 * d * nz + 1)).astype(np.float32)
    fbuf = dev(H, vals)
    Zh = fbuf.h
    Ah = [fbuf.h + 4 * (2 * nz * d + i * d * d) for i in range(ni)]
    perm = rng.permutation(2 * nz)
    used = 0
    cols = []
    for nc in ncols:
        v0, v1 = perm[used:used + nc], perm[used + nc:used + 2 * nc]
        used += 2 * nc
        pad = 16 * ((nc + 15) // 16)
        c0 = np.full(pad, 2 * nz * d, np.uint32)
        c1 = np.full(pad, 2 * nz * d, np.uint32)
        c0[:nc], c1[:nc] = v0 * d, v1 * d
        cols.append((c0, c1))
    blocks = [(i, bx, by) for i in range(ni) for bx in range((d + 63) // 64)
              for by in range((ncols[i] + 15) // 16)]
    nb = len(blocks)
    isz = 11 * nb + sum(2 * len(c[0]) for c in cols)
    ibuf = H.DeviceArray.empty(isz, dtype=np.uint32)
    host = np.zeros(isz, np.uint32)
    colh, off = [], 11 * nb
    for c0, c1 in cols:
        host[off:off + len(c0)] = c0
        host[off + len(c0):off + 2 * len(c0)] = c1
        colh.append((ibuf.h + 4 * off, ibuf.h + 4 * (off + len(c0))))
        off += 2 * len(c0)
    for b, (i, bx, by) in enumerate(blocks):
        o = nb + 10 * b
        host[b] = ibuf.h + 4 * o
        host[o:o + 10] = [i, bx, by, d, nz, ncols[i], Ah[i], Zh, colh[i][0], colh[i][1]]
    ibuf.upload(host)
    return fbuf, vals, ibuf, host, nb, cols


@pytest.mark.parametrize("ncol_field", [True, False])
def test_sgemv4_descriptors(H, ncol_field):
    """opcuda_sgemv4 on gsgemv4-style descriptor tables (several matrices,
    ragged column counts, sentinel padding; d = 200 leaves a partial 64-row
    tile) against the pinned gather oracle at the reference's threshold
    |a - b|/(1 + |a|) <= 1e-3 (gsgemv4/Program.cs:470-482) and at 1e-5 of
    the result's scale; vectors no column writes stay untouched.  The
    descriptor's ncol field is read by the reference's host code only:
    0 there stops at the sentinels alone"""
    banner(f"sgemv4 descriptors ncol_field={ncol_field}")
    rng = np.random.default_rng(31)
    d, nz, ncols = 200, 48, [5, 16, 23]
    fbuf, vals, ibuf, host, nb, cols = gsgemv4_tables(H, rng, d, nz, ncols)
    if not ncol_field:
        for b in range(nb):
            host[nb + 10 * b + 5] = 0
        ibuf.upload(host)
    H.lib.opcuda_sgemv4(nb, ibuf.h)
    H.check()
    H.sync()
    got = fbuf.download()[:2 * nz * d]
    Z = vals[:2 * nz * d].copy()
    for i, (c0, c1) in enumerate(cols):
        A = np.ascontiguousarray(vals[2 * nz * d + i * d * d: 2 * nz * d + (i + 1) * d * d])
        O.lib.or_sgemv_gather(d, A, Z, len(c0), c0.astype(np.int32), c1.astype(np.int32), 2 * nz * d)
    assert ref_tol(Z, got) <= 1e-3
    assert np.abs(got - Z).max() <= 1e-5 * np.abs(Z).max()
    written = np.concatenate([c1[c1 < 2 * nz * d] // d for _, c1 in cols])
    untouched = np.setdiff1d(np.arange(2 * nz), written)
    g2, v2 = got.reshape(2 * nz, d), vals[:2 * nz * d].reshape(2 * nz, d)
    assert np.array_equal(g2[untouched], v2[untouched])
    assert not np.array_equal(g2[written], v2[written])


def test_skinny_backward_step(H):
    banner("skinny")
    rng = np.random.default_rng(9)
    for m, ncp in ((2048, 48), (256, 32), (384, 48), (128, 16), (320, 16)):
        A = (rng.random((m, m)) / m).astype(np.float32)
        X = rng.random((m, ncp)).astype(np.float32)
        Ad, Xd = dev(H, A), dev(H, X, order="C")
        for plus in (0, 1):
            Y = H.DeviceArray.empty(m * ncp)
            H.check(H.lib.ophip_sgemm_skinny_ex(m, m, ncp, np.float32(0.99), Ad.h, m, Xd.h, ncp,
                                                 Y.h, ncp, plus))
            got = Y.download().reshape(m, ncp)
            want = 0.99 * (A.astype(np.float64) @ X.astype(np.float64) + plus * X)
            assert np.abs(got - want).max() < 2e-6, (m, ncp, plus)
            # the same launch twice gives the same bits (fixed-order reduction)
            H.check(H.lib.ophip_sgemm_skinny_ex(m, m, ncp, np.float32(0.99), Ad.h, m, Xd.h, ncp,
                                                 Y.h, ncp, plus))
            assert np.array_equal(Y.download().reshape(m, ncp), got)


# ---- tables and Monte Carlo ---------------------------------------------------------

def stochastic(rng, nm, d, power=6):
    P = rng.random((nm, d, d)) ** power
    P /= P.sum(axis=2, keepdims=True)
    return P.astype(np.float32)


@pytest.mark.parametrize("d", [200, 2048, 2100])  # one pass up to d = 2048, two passes beyond
def test_sck_bit_exact_both_layouts(H, d):
    banner(f"sck d={d}")
    rng = np.random.default_rng(11)
    nm = 2 if d < 1024 else 1
    P = stochastic(rng, nm, d)
    P[0, 3, 50] = -1e-9  # the reference layout sums it raw; the row layout counts it as zero
    flat = np.ascontiguousarray(np.stack([p.reshape(-1, order="F") for p in P]).reshape(-1))
    t = dev(H, flat)
    c = H.DeviceArray.empty(nm * d * d)
    H.lib.opcuda_sck(d, nm, t.h, c.h)
    H.check()
    want = np.zeros_like(flat)
    O.lib.or_sck_f32(d, nm, flat, want)
    assert np.array_equal(c.download(), want)
    cr = H.DeviceArray.empty(nm * d * d)
    H.check(H.lib.ophip_sck_rows(d, nm, t.h, cr.h))
    assert np.array_equal(cr.download().reshape(nm, d, d), O.sck_rows2_f32(P, 64, 0))
    # the W form's table (1 added to each diagonal entry first)
    Wt = P.copy()
    Wt[:, np.arange(d), np.arange(d)] -= 1.0
    wflat = np.ascontiguousarray(np.stack([p.reshape(-1, order="F") for p in Wt]).reshape(-1))
    tw = dev(H, wflat)
    H.check(H.lib.ophip_sck_rows_ex(d, nm, tw.h, cr.h, 1))
    assert np.array_equal(cr.download().reshape(nm, d, d), O.sck_rows2_f32(Wt, 64, 1))


def test_mt19937_streams_bit_exact(H):
    banner("mt19937")
    nrng = 16
    H.check(H.lib.ophip_mc_set_nrng(nrng))
    try:
        assert H.lib.opcuda_mc_nrng() == nrng
        sz = H.lib.opcuda_mc_status_sz()
        assert sz == 625 * nrng
        mt = H.DeviceArray.empty(sz, dtype=np.uint32)
        seeds = (5489 + np.arange(nrng)).astype(np.uint32)
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()
        per = 64 * 30 + 17  # crosses several twists (624 words each); a partial 64-block
        u = H.DeviceArray.empty(nrng * per)
        for rep in range(2):  # the state carries over between launches, exactly per outputs
            H.check(H.lib.opcuda_mt_benchmark(mt.h, u.h, nrng * per))
            got = u.download().reshape(per, nrng)  # unif[s + nrng*j], smc.cu:712
            for s in (0, 1, 7, 15):
                want = O.mt_uniforms(5489 + s, per * (rep + 1))[per * rep:]
                assert np.array_equal(got[:, s], want), (s, rep)
    finally:
        H.lib.ophip_mc_set_nrng(4096)


@pytest.mark.parametrize("d,nrng,groups,nk", [(50, 4, 2, 7), (256, 8, 1, 16), (700, 2, 3, 5)])
def test_mc_chain_bit_exact(H, d, nrng, groups, nk):
    banner(f"mc1f d={d}")
    rng = np.random.default_rng(d)
    nm = 3
    P = stochastic(rng, nm, d, power=10)
    flat = np.ascontiguousarray(np.stack([p.reshape(-1, order="F") for p in P]).reshape(-1))
    t = dev(H, flat)
    ct = H.DeviceArray.empty(nm * d * d)
    H.lib.opcuda_sck(d, nm, t.h, ct.h)
    H.check()
    ctab = ct.download()
    cr = H.DeviceArray.empty(nm * d * d)
    H.check(H.lib.ophip_sck_rows(d, nm, t.h, cr.h))
    m_k = (np.arange(nk) % nm).astype(np.int32)
    mkd = dev(H, m_k)
    nscen = 64 * nrng * groups
    H.check(H.lib.ophip_mc_set_nrng(nrng))
    try:
        y0 = d // 3
        crows = cr.download()
        want_c = O.mc_chain(nrng, 777, groups, nk, d, y0, ctab, m_k, 0)
        want = O.mc_chain(nrng, 777, groups, nk, d, y0, crows, m_k, 1)
        for rows in (False, True):
            mt = H.DeviceArray.empty(625 * nrng, dtype=np.uint32)
            seeds = (777 + np.arange(nrng)).astype(np.uint32)
            H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
            H.check()
            ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
            host = np.zeros(nscen * nk, dtype=np.int16)
            fn = H.lib.ophip_mc1f_rows if rows else H.lib.opcuda_mc1f
            H.check(fn(mt.h, y0, ysk.h, nscen, nk, d, (cr if rows else ct).h, mkd.h,
                       host.ctypes.data))
            assert np.array_equal(host.reshape(nk, nscen), want if rows else want_c), rows
        # guide tables: the same states from the indexed search
        gd = H.DeviceArray.empty(nm * d * 257, dtype=np.int16)
        H.check(H.lib.ophip_mc_guide(d, nm, cr.h, gd.h))
        crt = cr.download().reshape(nm, d, d)
        g = gd.download().reshape(nm, d, 257)
        for b in (0, 1, 128, 255, 256):  # guide[b] = min{j : b/256 < C[j]}, clamped
            t_ = np.float32(b / 256.0)
            exp = np.where((crt > t_).any(axis=2), (crt > t_).argmax(axis=2), d - 1)
            assert np.array_equal(g[:, :, b], exp), b
        mt = H.DeviceArray.empty(625 * nrng, dtype=np.uint32)
        seeds = (777 + np.arange(nrng)).astype(np.uint32)
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()
        ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
        H.check(H.lib.ophip_mc1f_guided(mt.h, y0, ysk.h, nscen, nk, d, cr.h, gd.h, mkd.h))
        assert np.array_equal(ysk.download().reshape(nk, nscen), want)
    finally:
        H.lib.ophip_mc_set_nrng(4096)


@pytest.mark.parametrize("nscen", [4096, 4096 * 25])
def test_mc1f_reference_batch_sizes(H, nscen):
    """the reference drivers' batch sizes at the default 4096 streams
    (glv/Program.cs:73: 4096; gsv and clv/Program.cs:48: 4096*25): per =
    nscen/4096 paths per stream, bit-exact against the oracle's mapping on
    the reference table layout (opcuda_mc1f) and the row layout (guided);
    each call consumes exactly per*nk outputs of every stream"""
    banner(f"mc1f reference batch nscen={nscen}")
    assert H.lib.opcuda_mc_nrng() == 4096
    nrng, per = 4096, nscen // 4096
    rng = np.random.default_rng(nscen)
    d, nk = 128, 40  # clv's lattice and weekly steps (clv/Program.cs:46-60)
    P = stochastic(rng, 1, d, power=8)
    flat = np.ascontiguousarray(P[0].reshape(-1, order="F"))
    t = dev(H, flat)
    ct = H.DeviceArray.empty(d * d)
    H.lib.opcuda_sck(d, 1, t.h, ct.h)
    H.check()
    cr = H.DeviceArray.empty(d * d)
    H.check(H.lib.ophip_sck_rows(d, 1, t.h, cr.h))
    gd = H.DeviceArray.empty(d * 257, dtype=np.int16)
    H.check(H.lib.ophip_mc_guide(d, 1, cr.h, gd.h))
    mkd = dev(H, np.zeros(nk, dtype=np.int32))
    y0 = 40
    seeds = (5489 + np.arange(nrng)).astype(np.uint32)
    want_c = O.mc_chain(nrng, 5489, 0, nk, d, y0, ct.download(), np.zeros(nk, np.int32), 0, per=per)
    want_r = O.mc_chain(nrng, 5489, 0, nk, d, y0, cr.download(), np.zeros(nk, np.int32), 1, per=per)
    ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
    mt = H.DeviceArray.empty(H.lib.opcuda_mc_status_sz(), dtype=np.uint32)
    H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
    H.check()
    host = np.zeros(nscen * nk, dtype=np.int16)
    H.check(H.lib.opcuda_mc1f(mt.h, y0, ysk.h, nscen, nk, d, ct.h, mkd.h, host.ctypes.data))
    assert np.array_equal(host.reshape(nk, nscen), want_c)
    # the stream continues exactly per*nk outputs later
    u = H.DeviceArray.empty(4096 * 3)
    H.check(H.lib.opcuda_mt_benchmark(mt.h, u.h, 4096 * 3))
    got = u.download().reshape(3, 4096)
    for s in (0, 1, 2049, 4095):
        assert np.array_equal(got[:, s], O.mt_uniforms(5489 + s, per * nk + 3)[per * nk:]), s
    H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
    H.check()
    H.check(H.lib.ophip_mc1f_guided(mt.h, y0, ysk.h, nscen, nk, d, cr.h, gd.h, mkd.h))
    assert np.array_equal(ysk.download().reshape(nk, nscen), want_r)


def test_mc_c1_lv_kernel_bit_exact(H):
    """the 1-D workload's Monte Carlo leg at its size: 2^16 paths x 16 steps
    on its own kernel U(1/16) (plain FP32 solve, reference table layout),
    one MT19937 stream seeded 5489 serving every path (scenario j = path
    j, groups of 64, step-major), bit-exact against the oracle; the paths'
    call prices at T = 1 yr within 5 standard errors of the FP64 oracle's"""
    banner("mc C1: LV d=256, 2^16 x 16, seed 5489")
    import ctypes as C
    import os
    from oplib_amd import opmodel as M
    cfg = M.config_lv1d()
    s = M.Solver(cfg)
    s.solve()
    U = s.kernel()  # U(1/16), column-major
    s.close()
    d, nk, nscen = cfg.d, 16, 1 << 16
    x, x0 = M.grid(cfg)
    Ud = dev(H, U)
    ct = H.DeviceArray.empty(d * d)
    H.lib.opcuda_sck(d, 1, Ud.h, ct.h)
    H.check()
    mk = np.zeros(nk, dtype=np.int32)
    want = O.mc_chain(1, 5489, 0, nk, d, x0, ct.download(), mk, 0, per=nscen)
    H.check(H.lib.ophip_mc_set_nrng(1))
    try:
        mt = H.DeviceArray.empty(625, dtype=np.uint32)
        seed = np.array([5489], dtype=np.uint32)
        H.lib.opcuda_mc_setseed(seed.ctypes.data, mt.h)
        H.check()
        ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
        mkd = dev(H, mk)
        H.check(H.lib.opcuda_mc1f(mt.h, x0, ysk.h, nscen, nk, d, ct.h, mkd.h, 0))
        got = ysk.download().reshape(nk, nscen)
        assert np.array_equal(got, want)
    finally:
        H.lib.ophip_mc_set_nrng(4096)
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "prices_lv1d.npz"))["prices"]
    p64 = gold.reshape(-1, len(cfg.strike_list()))[-1]  # T = 16/16 yr
    xs = x.astype(np.float32)[got[-1]]
    ks = np.array(cfg.strike_list(), dtype=np.float64)
    pay = np.maximum(xs[:, None].astype(np.float64) - ks[None, :], 0.0)
    df = M.df(cfg, 1.0)
    mc, se = df * pay.mean(0), df * pay.std(0) / math.sqrt(nscen)
    z = np.abs(mc - p64) / se
    print(f"[mc C1] max |MC - FP64 operator| = {z.max():.2f} SE", flush=True)
    assert z.max() < 5.0


def test_mc_c2_full_size_bit_exact(H):
    """the regime workload's Monte Carlo leg at its size: 2^20 paths x 252
    daily steps on the 2048-state daily kernel U(1/252) (FP16 x3 solve,
    row-contiguous table), 16384 MT19937 streams, one per wavefront, seeded
    5489 + i, 64 paths each: every sampled state bit-exact against the
    oracle; the guided sampler and the plain bisection agree"""
    banner("mc C2: d=2048, 2^20 x 252, 16384 streams seeded 5489+i")
    from oplib_amd import opmodel as M
    cfg = M.config_rs2048()
    cfg.precision = M.OPM_PREC_F16X3
    s = M.Solver(cfg)
    s.solve()
    U = s.kernel()
    s.close()
    d, nk, nscen, nrng = cfg.d, 252, 1 << 20, 16384
    x, x0 = M.grid(cfg)
    y0 = x0 + cfg.nx * 1  # regime 1's spot
    Ud = dev(H, U)
    cr = H.DeviceArray.empty(d * d)
    H.check(H.lib.ophip_sck_rows(d, 1, Ud.h, cr.h))
    del Ud
    gd = H.DeviceArray.empty(d * 257, dtype=np.int16)
    H.check(H.lib.ophip_mc_guide(d, 1, cr.h, gd.h))
    mk = np.zeros(nk, dtype=np.int32)
    mkd = dev(H, mk)
    seeds = (5489 + np.arange(nrng)).astype(np.uint32)
    H.check(H.lib.ophip_mc_set_nrng(nrng))
    try:
        mt = H.DeviceArray.empty(625 * nrng, dtype=np.uint32)
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()
        ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
        H.check(H.lib.ophip_mc1f_guided(mt.h, y0, ysk.h, nscen, nk, d, cr.h, gd.h, mkd.h))
        got = ysk.download().reshape(nk, nscen)
        print("[mc C2] GPU done, oracle running", flush=True)
        want = O.mc_chain(nrng, 5489, 0, nk, d, y0, cr.download(), mk, 1, per=64)
        bad = np.count_nonzero(got != want)
        print(f"[mc C2] {got.size} states, {bad} differ", flush=True)
        assert bad == 0
        del want
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()
        H.check(H.lib.ophip_mc1f_rows(mt.h, y0, ysk.h, nscen, nk, d, cr.h, mkd.h, 0))
        assert np.array_equal(ysk.download().reshape(nk, nscen), got)
    finally:
        H.lib.ophip_mc_set_nrng(4096)


def payoff_moments(ysk, steps, xval, strikes, df):
    """numpy: discounted mean and second moment of max(x[y] - K, 0) per
    listed step and strike, the payoff in FP32 as the kernel forms it,
    sums in FP64"""
    n = ysk.shape[1]
    m1 = np.zeros((len(steps), len(strikes)))
    m2 = np.zeros_like(m1)
    for i, k in enumerate(steps):
        x = xval[np.clip(ysk[k], 0, len(xval) - 1)].astype(np.float32)
        for c, K in enumerate(strikes):
            v = np.maximum(x - np.float32(K), np.float32(0)).astype(np.float64)
            m1[i, c] = df[i] * v.sum() / n
            m2[i, c] = df[i] ** 2 * (v * v).sum() / n
    return m1, m2


def test_mc_payoff_device(H):
    """ophip_mc_payoff against numpy on random states: discounted means and
    second moments of the call payoffs per listed step and strike (FP64
    sums in another order: 1e-12 relative), out-of-lattice states read the
    lattice edge"""
    banner("mc payoff reduction")
    import ctypes as C
    rng = np.random.default_rng(5)
    d, nk, nscen = 300, 6, 3 * 4096 + 77
    ysk = rng.integers(0, d, size=(nk, nscen)).astype(np.int16)
    ysk[2, :5] = [-3, d, d + 9, 0, d - 1]
    xval = np.linspace(0.0, 400.0, d).astype(np.float32)
    steps = np.array([5, 0, 2], dtype=np.int32)
    strikes = np.array([80.0 + 2.5 * c for c in range(17)], dtype=np.float32)
    df = np.array([0.97, 1.0, 0.99])
    yd = H.DeviceArray.from_numpy(ysk.reshape(-1))
    xd = H.DeviceArray.from_numpy(xval)
    out = H.DeviceArray.empty(2 * len(steps) * len(strikes), dtype=np.float64)
    H.check(H.lib.ophip_mc_payoff(yd.h, nscen, nk, len(steps), steps.ctypes.data_as(C.POINTER(C.c_int)),
                                  df.ctypes.data_as(C.POINTER(C.c_double)), xd.h, d,
                                  strikes.ctypes.data_as(C.POINTER(C.c_float)), len(strikes), out.h))
    got = out.download().reshape(2, len(steps), len(strikes))
    m1, m2 = payoff_moments(ysk, steps, xval, strikes, df)
    assert np.allclose(got[0], m1, rtol=1e-12, atol=0)
    assert np.allclose(got[1], m2, rtol=1e-12, atol=0)
    # a step beyond the states held is refused
    bad = np.array([nk], dtype=np.int32)
    assert H.lib.ophip_mc_payoff(yd.h, nscen, nk, 1, bad.ctypes.data_as(C.POINTER(C.c_int)),
                                 df.ctypes.data_as(C.POINTER(C.c_double)), xd.h, d,
                                 strikes.ctypes.data_as(C.POINTER(C.c_float)), len(strikes), out.h) != 0


def test_mc_prices_match_operator(H):
    """Monte Carlo on the operator's own kernel: 2^18 chains on the
    cumulative table of the plain FP32 U(dt) of a small regime lattice,
    payoffs evaluated on the device, against the backward-induction prices
    of the same kernel (the expectation the sampler estimates), within five
    standard errors"""
    banner("mc vs operator")
    import ctypes as C
    from oplib_amd import opmodel as M
    cfg = M.config_small(nx=64, nr=2, nsteps=8)
    s = M.Solver(cfg)
    s.solve()
    p_bi = s.prices()  # [nr][1][nstrikes]
    U = s.kernel()     # reference order, column-major [y0, y]
    s.close()
    d, nk = cfg.d, cfg.nsteps
    x, x0 = M.grid(cfg)
    Ud = H.DeviceArray.from_numpy(np.ascontiguousarray(U.reshape(-1, order="F")))
    ct = H.DeviceArray.empty(d * d)
    H.check(H.lib.ophip_sck_rows(d, 1, Ud.h, ct.h))
    nrng, G = 64, 64
    nscen = 64 * nrng * G
    H.check(H.lib.ophip_mc_set_nrng(nrng))
    try:
        mt = H.DeviceArray.empty(625 * nrng, dtype=np.uint32)
        seeds = (5489 + np.arange(nrng)).astype(np.uint32)
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()
        mk = H.DeviceArray.from_numpy(np.zeros(nk, dtype=np.int32))
        ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
        xs = H.DeviceArray.from_numpy(np.tile(x, cfg.nr).astype(np.float32))
        ks = np.array(cfg.strike_list(), dtype=np.float32)
        df = np.array([M.df(cfg, cfg.DeltaT * nk)])
        steps = np.array([nk - 1], dtype=np.int32)
        out = H.DeviceArray.empty(2 * len(ks), dtype=np.float64)
        for r in range(cfg.nr):
            y0 = x0 + cfg.nx * r
            H.check(H.lib.ophip_mc1f_rows(mt.h, y0, ysk.h, nscen, nk, d, ct.h, mk.h, 0))
            H.check(H.lib.ophip_mc_payoff(ysk.h, nscen, nk, 1, steps.ctypes.data_as(C.POINTER(C.c_int)),
                                          df.ctypes.data_as(C.POINTER(C.c_double)), xs.h, d,
                                          ks.ctypes.data_as(C.POINTER(C.c_float)), len(ks), out.h))
            m = out.download().reshape(2, len(ks))
            se = np.sqrt(np.maximum(m[1] - m[0] ** 2, 0.0) / nscen)
            dev_ = np.abs(m[0] - p_bi[r, 0])
            print(f"[mc] regime {r}: max |mc - operator| / se = {(dev_ / np.maximum(se, 1e-12)).max():.2f}",
                  flush=True)
            assert np.all(dev_ <= 5.0 * se + 1e-6 * np.abs(p_bi[r, 0]).max())
    finally:
        H.lib.ophip_mc_set_nrng(4096)


def test_mc1f_rejects_bad_shapes(H):
    banner("mc1f args")
    H.check(H.lib.ophip_mc_set_nrng(4))
    try:
        st = H.lib.opcuda_mc1f(256, 0, 256, 102, 4, 8, 256, 256, 0)  # nscen % nrng != 0
        assert st != 0
    finally:
        H.lib.ophip_mc_set_nrng(4096)


@pytest.mark.parametrize("plus", [0, 1])
def test_spot_prices(H, plus):
    """df (plus pay(y0, c) + sum_j U[y0][j] pay(j, c)) for a few rows of a
    column-major U, against FP64 numpy (FP32 accumulation: 1e-5)"""
    import ctypes as C
    banner(f"spot prices plus_identity={plus}")
    rng = np.random.RandomState(21)
    d, ns = 512, 41
    U = rng.uniform(0, 1.0 / d, size=(d, d)).astype(np.float32)
    pay = np.maximum(rng.uniform(0, 400, size=d)[:, None] - (80.0 + np.arange(ns))[None, :], 0).astype(np.float32)
    Ud = dev(H, U)                       # column-major
    Pd = dev(H, pay.T)                   # payT[c + ns j] = pay[j][c]
    rows = np.array([7, 64, 300, 511], dtype=np.int32)
    out = H.DeviceArray.empty(len(rows) * ns)
    H.check(H.lib.ophip_spot_prices(d, Ud.h, len(rows), rows.ctypes.data_as(C.POINTER(C.c_int)), plus,
                                    Pd.h, ns, np.float32(0.97), out.h))
    got = out.download().reshape(len(rows), ns)
    ref = 0.97 * (U[rows].astype(np.float64) @ pay.astype(np.float64) + plus * pay[rows])
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("out_kind", [4, 1, 3, 0, 2])
def test_sv_generator_vs_oracle(H, out_kind):
    """ophip_dgen_sv (FP64 on the device: a row table, then one thread per
    element) against the oracle's restatement of CSVModel.host_d_gen:
    FP64 outputs within 1e-13 of the largest rate (exp and division
    rounding), FP32 outputs one rounding of the FP64 value"""
    banner(f"dgen_sv out_kind={out_kind}")
    from _cases import sv_case
    nx, nr, *arrs = case = sv_case(nx=48, nr=6, seed=3)
    d = nx * nr
    L = O.sv_generator(*case)
    dt = 2.0 ** -16
    hd = [dev(H, np.ascontiguousarray(a, dtype=np.float64)) for a in arrs]
    dbl = out_kind in (1, 3, 4)
    out = H.DeviceArray.empty((d, d), dtype=np.float64 if dbl else np.float32)
    H.check(H.lib.ophip_dgen_sv(nx, nr, *[a.h for a in hd], dt, out_kind, out.h))
    got = out.download()
    want = {4: L, 1: np.eye(d) + dt * L, 3: dt * L, 0: np.eye(d) + dt * L, 2: dt * L}[out_kind]
    scale = np.abs(want - (np.eye(d) if out_kind in (0, 1) else 0)).max()
    if dbl:
        assert np.abs(got - want).max() <= 1e-13 * max(scale, 1.0)
    else:
        # one FP32 rounding of the FP64 value (plus the FP64 rounding slack)
        assert np.all(np.abs(got.astype(np.float64) - want) <= np.abs(want) * 2.0 ** -24 + 1e-13 * scale)


def sv_invm(nx, nr, xv, rv):
    """S2DGrid.pregen (S2DGrid.cs:123-245): per state the 4 x 4 system of
    (x drift, x variance, r drift, r variance) over the neighbours (x-1,
    x+1, r-1, r+1), identity-like rows at the edges, solved; stored
    invm[16 y + i + 4 j] = inverse[i][j]"""
    d = nx * nr
    inv = np.zeros(16 * d)
    for a0 in range(nr):
        for x0 in range(nx):
            y = x0 + nx * a0
            m = np.zeros((4, 4))
            m[0, 0] = xv[y - 1] - xv[y] if x0 > 0 else 0.0
            m[0, 1] = xv[y + 1] - xv[y] if x0 < nx - 1 else 0.0
            if 0 < x0 < nx - 1:
                m[1, 0], m[1, 1] = m[0, 0] ** 2, m[0, 1] ** 2
            elif x0 == 0:
                m[1, 0] = 1.0
            else:
                m[1, 1] = 1.0
            m[2, 2] = rv[y - nx] - rv[y] if a0 > 0 else 0.0
            m[2, 3] = rv[y + nx] - rv[y] if a0 < nr - 1 else 0.0
            if 0 < a0 < nr - 1:
                m[3, 2], m[3, 3] = m[2, 2] ** 2, m[2, 3] ** 2
            elif a0 == 0:
                m[3, 2] = 1.0
            else:
                m[3, 3] = 1.0
            inv[16 * y:16 * y + 16] = np.linalg.inv(m).reshape(-1, order="F")
    return inv


def test_sgsvg_dropin_vs_oracle(H):
    """opcuda_sgsvg with the reference's arguments (CSVModel.device_sgen,
    CSVModel.cs:386-417: the stencil inverses of S2DGrid.pregen and the
    pars block xval | SDrift | SVol | VolDrift | VolVol | Jm | Jp over ni
    generators) against the oracle's FP64 generator of the same (FP32)
    inputs: every row within 4 (nx - 1) 2^-24 of its largest rate (the
    FP32 invm carries 2^-24 per entry into each rate, and the drift fix
    sums rates times distances of up to nx - 1 grid steps and divides by
    one step), one FP32 rounding of the result; the whole block is
    written"""
    banner("sgsvg drop-in")
    from _cases import sv_case
    nx, nr, xv, rv, sd, sv, vv, vd, jm, jp = sv_case(nx=48, nr=6, seed=5)
    d, ni = nx * nr, 2
    gens = [(sd, sv, vd, vv, jm, jp), (0.5 * sd, 1.2 * sv, 0.7 * vd, 1.3 * vv, 0.5 * jm, 2.0 * jp)]
    f32 = lambda a: np.asarray(a, dtype=np.float32)  # noqa: E731
    pars = np.zeros(d + 6 * ni * d, np.float32)
    pars[:d] = xv
    for i, g in enumerate(gens):
        for k, a in enumerate(g):
            pars[d + k * ni * d + i * d: d + k * ni * d + (i + 1) * d] = a
    inv = sv_invm(nx, nr, xv, rv).astype(np.float32)
    gd = H.DeviceArray.from_numpy(np.full(ni * d * d, np.nan, np.float32))
    invd, pd = dev(H, inv), dev(H, pars)
    H.lib.opcuda_sgsvg(gd.h, ni, nx, nr, invd.h, pd.h)
    H.check()
    got = gd.download().reshape(ni, d, d).transpose(0, 2, 1)  # [i][y0][y1]
    assert np.all(np.isfinite(got))
    for i, (s_d, s_v, v_d, v_v, j_m, j_p) in enumerate(gens):
        want = O.sv_generator(nx, nr, f32(xv).astype(np.float64), rv, f32(s_d), f32(s_v), f32(v_v),
                              f32(v_d), f32(j_m), f32(j_p))
        scale = np.abs(want).max(axis=1, keepdims=True)
        err = (np.abs(got[i].astype(np.float64) - want) / scale).max()
        print(f"[sgsvg] generator {i}: max row-relative error {err:.2e}", flush=True)
        assert err <= 4 * (nx - 1) * 2.0**-24, (i, err)
        assert np.all(got[i][want == 0] == 0)  # structural zeros stay zero
    # the reference's entry point refuses what its kernel cannot build
    H.lib.opcuda_sgsvg(gd.h, ni, nx, 1, invd.h, pd.h)
    assert H.lib.opcuda_get_status() != 0


def test_dcmt_lanes_bit_exact(H):
    """the reference's 4096 DCMT-607 lanes from a parameter table in
    MersenneTwister.dat's format: seeding, two calls of uniforms (25 per
    lane: not a whole turn of the 19-word buffer, whose index restarts at 0
    on every call) and the saved states, bit for bit against the oracle"""
    banner("dcmt lanes")
    import ctypes as C
    rng = np.random.default_rng(607)
    n = 4096
    table = np.zeros((n, 4), dtype=np.uint32)
    table[:, 0] = rng.integers(0, 2**31, n, dtype=np.uint32) | np.uint32(0x80000000)  # matrix_a
    table[:, 1] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)       # mask_b
    table[:, 2] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)       # mask_c
    table[:, 3] = np.arange(n, dtype=np.uint32) + 4172
    assert H.lib.ophip_dcmt_load(table.ctypes.data, table.nbytes - 16) == 1
    assert H.lib.ophip_dcmt_load(table.ctypes.data, table.nbytes) == 0
    assert H.lib.ophip_dcmt_status_sz() == 19 * n
    seeds = (5489 + np.arange(n)).astype(np.uint32)
    st = H.DeviceArray.empty(19 * n, dtype=np.uint32)
    H.check(H.lib.ophip_dcmt_setseed(seeds.ctypes.data_as(C.POINTER(C.c_uint)), st.h))
    ost = O.tgfsr_seed(seeds, 19)
    assert np.array_equal(st.download(), ost)
    per = 25
    u = H.DeviceArray.empty(n * per)
    for call in range(2):
        H.check(H.lib.ophip_dcmt_uniforms(st.h, u.h, n * per))
        raw = O.tgfsr_run(O.DCMT607, ost, per, table[:, 0], table[:, 1], table[:, 2])
        want = raw.astype(np.float32) / np.float32(4294967296.0)
        assert np.array_equal(u.download().reshape(per, n), want), call
        assert np.array_equal(st.download(), ost), call


def test_sv_generator_from_term_structure(H):
    """the SV model's generator from its term structure: opm_sv_params
    (CSVModel.mkgen) for the second of two dates, then ophip_dgen_sv on the
    device, against the oracle's mkgen and generator (FP64 L within 1e-13
    of its largest rate)"""
    banner("dgen_sv from term-structure curves")
    from oplib_amd import opmodel as M
    nx, nr = 48, 6
    x, x0 = O.xgrid(nx, 10.0, 400.0, 100.0)
    rv = np.linspace(0.1, 0.6, nr)
    tau = np.array([[0.2, 0.3, 1.0, 0.5], [-0.5, -0.2, 2.0, 0.0], [0.4, 0.1, 0.5, 1.0],
                    [0.3, 0.3, 0.0, 0.0], [2.0, 1.0, 1.5, -0.5], [0.35, 0.3, 1.0, 0.0],
                    [0.1, 0.15, 3.0, 0.2], [0.05, 0.02, 3.0, 0.2]])
    ts, y0 = [0.25, 1.0], x0 + nx * 2
    got = [a[1] for a in M.sv_params(x, rv, y0, ts, tau)]
    want = [a[1] for a in O.sv_params(x, rv, y0, ts, tau)]
    xs, rs = np.tile(x, nr), np.repeat(rv, nx)
    sd = 0.02 * xs
    L = O.sv_generator(nx, nr, xs, rs, sd, *want)
    d = nx * nr
    hd = [dev(H, np.ascontiguousarray(a, dtype=np.float64)) for a in (xs, rs, sd, *got)]
    out = H.DeviceArray.empty((d, d), dtype=np.float64)
    H.check(H.lib.ophip_dgen_sv(nx, nr, *[a.h for a in hd], 2.0 ** -16, 4, out.h))
    scale = np.abs(L).max()
    assert np.abs(out.download() - L).max() <= 1e-13 * scale
    assert np.allclose(L.sum(axis=1), 0.0, atol=1e-12 * scale)
