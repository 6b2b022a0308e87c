"""GPU tests of the U = I + W squaring kernels (ophip_wsq_*): every stage
against a numpy restatement of the same arithmetic, and the batched plan
against single-lattice plans.

Bars: the 16-bit copies are bit-exact (one IEEE round-to-nearest-even of
the FP32 value times the scale); one squaring agrees with the numpy
restatement to FP32 accumulation-order rounding (1e-6 relative to the
largest entry); the diagonal fix leaves FP64 row sums of 0 up to FP32
rounding of the diagonal; a batch gives the same bits as the lattices
squared alone.
"""
import contextlib
import functools
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

S16 = 2.0**15  # FP16 scale of the kernels
G16 = 2048.0
G_BF = 256.0


def banner(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def bf16_round(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def enc(x, mode):
    """16-bit copy (as float32 values) and the low part, as the kernels do"""
    x = np.asarray(x, dtype=np.float32)
    if mode == 2:
        h = bf16_round(x)
        lo = bf16_round((x - h) * np.float32(G_BF))
        return h, lo
    xs = (x * np.float32(S16)).astype(np.float32)
    h = xs.astype(np.float16).astype(np.float32)
    lo = ((xs - h) * np.float32(G16)).astype(np.float16).astype(np.float32)
    return h, lo


def bits16(x, mode):
    if mode == 2:
        return (np.ascontiguousarray(x, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)
    return np.asarray(x, dtype=np.float32).astype(np.float16).view(np.uint16)


class Lattice:
    def __init__(self, H, nx, nr, mode, seed=0):
        from oplib_amd import opmodel as M
        self.H = H
        c = M.config_small(nx=nx, nr=nr, nsq=12)
        self.cfg, self.mode = c, mode
        self.d = nx * nr
        x, _ = M.grid(c)
        dr, vo = M.params(c)
        self.xd = H.DeviceArray.from_numpy(x)
        self.dd = H.DeviceArray.from_numpy(dr)
        self.vd = H.DeviceArray.from_numpy(vo)
        nbytes = H.lib.ophip_wsq_bytes(self.d, mode)
        assert nbytes > 0
        self.state = H.DeviceArray(nbytes, np.uint8)
        self.dt = math.ldexp(c.DeltaT, -c.nsq)
        _, _, g = O.model_generator(c)
        self.W0 = (self.dt * g).astype(np.float32)  # what out_kind 2 writes (one rounding)

    def first_step(self):
        H, c = self.H, self.cfg
        w0 = H.lib.ophip_wsq_array(self.state.h, self.d, self.mode, 0, 0)
        H.check(H.lib.ophip_sek_rs(c.nx, c.nr, self.xd.h, self.dd.h, self.vd.h, c.q_off,
                                   self.dt, 2, w0))

    def arr(self, buf, which, dtype, shape_order="F"):
        H, d = self.H, self.d
        h = H.lib.ophip_wsq_array(self.state.h, d, self.mode, buf, which)
        out = np.empty(d * d, dtype=dtype)
        H.check(H.lib.opcuda_memcpy_d2h(h, out.ctypes.data, out.nbytes))
        return out.reshape(d, d, order=shape_order)


def plan(H, lats, tile, nsqs=None):
    import ctypes as C
    n = len(lats)
    st = (C.c_uint * n)(*[l.state.h for l in lats])
    ds = (C.c_int * n)(*[l.d for l in lats])
    ns = (C.c_int * n)(*nsqs) if nsqs is not None else None
    p = H.lib.ophip_wsq_plan(n, st, ds, ns, lats[0].mode, tile, 0)
    assert p, hex(H.lib.opcuda_get_status())
    return p


def fix_diag(W):
    W = W.astype(np.float32).copy()
    rs = W.astype(np.float64).sum(axis=1)
    idx = np.arange(W.shape[0])
    W[idx, idx] = (W[idx, idx].astype(np.float64) - rs).astype(np.float32)
    return W


def square_ref(W, mode):
    """2W + W*W with the kernels' operand rounding, FP64 accumulation"""
    h, lo = enc(W, mode)
    h64, lo64 = h.astype(np.float64), lo.astype(np.float64)
    if mode == 2:
        P = h64 @ h64 + (h64 @ lo64 + lo64 @ h64) / G_BF
    else:
        P = h64 @ h64
        if mode == 1:
            P = P + (h64 @ lo64 + lo64 @ h64) / G16
        P = P / (S16 * S16)
    return (2.0 * W.astype(np.float64) + P)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("tile", [64, 128])
def test_split_and_one_squaring(hip, mode, tile):
    banner(f"wsq stages mode={mode} tile={tile}")
    H = hip
    L = Lattice(H, 64, 2, mode)
    d = L.d
    p = plan(H, [L], tile)
    try:
        # exec with 0 squarings: split + diagonal fix of W0
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(p, 0))
        H.sync()
        W0 = L.arr(0, 0, np.float32)
        assert np.abs(W0 - L.W0).max() <= 1e-6 * np.abs(L.W0).max() + 1e-12  # diag fix only
        assert np.abs(W0.astype(np.float64).sum(axis=1)).max() < 1e-6 * np.abs(W0).max()
        # flush rule: an element whose 16-bit parts are both +-0 is stored as +0
        want_h, want_l = enc(W0, mode)
        flushed = (want_h == 0) & ((want_l == 0) | (mode == 0))
        assert np.all(W0.view(np.uint32)[flushed] == 0)
        assert np.any(flushed)  # the banded lattice has such elements
        # one squaring
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(p, 1))
        H.sync()
        W1 = L.arr(1, 0, np.float32)
        ref = square_ref(fix_diag(L.W0), mode)
        # the GPU diagonal is then corrected: compare off-diagonal entries
        off = ~np.eye(d, dtype=bool)
        err = np.abs(W1[off] - ref[off]).max()
        print(f"[err] one squaring mode={mode} tile={tile}: {err:.3e} (max |W1| "
              f"{np.abs(ref).max():.3e})", flush=True)
        assert err <= 1e-6 * np.abs(ref).max()
        assert np.abs(W1.astype(np.float64).sum(axis=1)).max() < 1e-6 * np.abs(W1).max()
        # two squarings land back in buffer 0
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(p, 2))
        H.sync()
        W2 = L.arr(0, 0, np.float32)
        ref2 = square_ref(fix_diag(W1), mode)
        err2 = np.abs(W2[off] - ref2[off]).max()
        assert err2 <= 1e-6 * np.abs(ref2).max()
    finally:
        H.lib.ophip_wsq_plan_free(p)


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_equals_single(hip, mode):
    banner(f"wsq batch mode={mode}")
    H = hip
    A = Lattice(H, 64, 2, mode)       # d = 128
    B = Lattice(H, 128, 2, mode)      # d = 256
    C = Lattice(H, 64, 4, mode)       # d = 256
    singles = []
    for L in (A, B, C):
        p = plan(H, [L], 64)
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(p, 5))
        H.sync()
        singles.append(L.arr(1, 0, np.float32))
        H.lib.ophip_wsq_plan_free(p)
    p = plan(H, [A, B, C], 64)
    for L in (A, B, C):
        L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, 5))
    H.sync()
    for L, s in zip((A, B, C), singles):
        assert np.array_equal(L.arr(1, 0, np.float32), s)
    H.lib.ophip_wsq_plan_free(p)
    # per-lattice counts (A squares 3 times, B 5, C 4) give what the
    # lattices give squared alone that many times
    want = {}
    for L, k in ((A, 3), (C, 4)):
        q = plan(H, [L], 64)
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(q, k))
        H.sync()
        want[id(L)] = L.arr(k & 1, 0, np.float32)
        H.lib.ophip_wsq_plan_free(q)
    p = plan(H, [A, B, C], 64, nsqs=[3, 5, 4])
    for L in (A, B, C):
        L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, -1))
    H.sync()
    assert np.array_equal(A.arr(1, 0, np.float32), want[id(A)])
    assert np.array_equal(B.arr(1, 0, np.float32), singles[1])
    assert np.array_equal(C.arr(0, 0, np.float32), want[id(C)])
    H.lib.ophip_wsq_plan_free(p)


@pytest.mark.parametrize("mode", [1])
def test_mixed_tiles_and_autotune(hip, mode):
    """a plan with a tile per lattice gives each lattice the bits of its
    own single-lattice plan with that tile; the autotuner returns valid,
    cached choices"""
    import ctypes as C
    banner(f"wsq mixed tiles + autotune mode={mode}")
    H = hip
    A = Lattice(H, 64, 2, mode)       # d = 128
    B = Lattice(H, 128, 2, mode)      # d = 256
    lats, tl = (A, B), (128, 64)
    singles = []
    for L, t in zip(lats, tl):
        p = plan(H, [L], t)
        L.first_step()
        H.check(H.lib.ophip_wsq_exec(p, 4))
        H.sync()
        singles.append(L.arr(0, 0, np.float32))
        H.lib.ophip_wsq_plan_free(p)
    st = (C.c_uint * 2)(*[L.state.h for L in lats])
    ds = (C.c_int * 2)(*[L.d for L in lats])
    tiles = (C.c_int * 2)(*tl)
    p = H.lib.ophip_wsq_plan_tiles(2, st, ds, None, mode, tiles, 0)
    assert p, hex(H.lib.opcuda_get_status())
    for L in lats:
        L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, 4))
    H.sync()
    for L, s in zip(lats, singles):
        assert np.array_equal(L.arr(0, 0, np.float32), s)
    H.lib.ophip_wsq_plan_free(p)
    # autotuner
    got = (C.c_int * 2)(0, 0)
    grid = C.c_int(0)
    H.check(H.lib.ophip_wsq_autotune(2, st, ds, None, mode, got, C.byref(grid)))
    assert all(t in (64, 128) for t in got) and grid.value > 0
    again = (C.c_int * 2)(0, 0)
    grid2 = C.c_int(0)
    H.check(H.lib.ophip_wsq_autotune(2, st, ds, None, mode, again, C.byref(grid2)))
    assert list(got) == list(again) and grid.value == grid2.value
    print(f"[autotune] tiles {list(got)} grid {grid.value}")


def dec16(u, mode):
    u = np.ascontiguousarray(u, dtype=np.uint16)
    if mode == 2:
        return (u.astype(np.uint32) << 16).view(np.float32)
    return u.view(np.float16).astype(np.float32)


def enc16(x, mode):
    x = np.ascontiguousarray(x, dtype=np.float32)
    if mode == 2:
        return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)
    return x.astype(np.float16).view(np.uint16)


def unfrag(a, nrows, K):
    """fragment-major (16 x 32 tiles, lane l = 16 g + r holds row r, k 8g..8g+7)
    -> [row][k]"""
    a = np.asarray(a).reshape(nrows // 16, K // 32, 4, 16, 8)  # it, kt, g, r, e
    return a.transpose(0, 3, 1, 2, 4).reshape(nrows, K)


BWD_KERNELS = {0: "dense", 2: "16-row tiles", 3: "64-row blocks", 4: "32-row blocks"}


@pytest.mark.parametrize("cut", [None, (-40, -28)])
@pytest.mark.parametrize("kernel", sorted(BWD_KERNELS))
@pytest.mark.parametrize("mode", [1, 2])
def test_wform_backward_step(hip, mode, kernel, cut):
    """Y = df (X + W X) from the state's hi/lo copies (fragment-major): against
    the same arithmetic in numpy (FP64 accumulation, 1e-6 of max|Y|; the
    band-sparse kernels read the hi parts of the fragments reaching
    2^visit and the lo parts of those reaching 2^lo (FP16) or 2^(lo-3)
    (BF16), which leaves out an entry under 2^visit or at most
    max(2^(lo-12), 2^-40) of one), against df (X + W X) with the FP32 W
    (2e-5), and X's / Y's 16-bit copies bit-exact; every step kernel, at
    the default cutoffs and at the tightest"""
    banner(f"wsq backward step mode={mode} kernel={BWD_KERNELS[kernel]} cutoffs={cut}")
    H = hip
    old = H.lib.ophip_wsq_set_bwd_kernel(kernel)
    try:
        with cutoffs(H, cut):
            _backward_step_case(H, mode, kernel == 0)
    finally:
        H.lib.ophip_wsq_set_bwd_kernel(old)


def _backward_step_case(H, mode, dense):
    L = Lattice(H, 128, 2, mode)  # d = 256
    p = plan(H, [L], 128)
    L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, 3))
    H.sync()
    H.lib.ophip_wsq_plan_free(p)
    d, ncp, buf = L.d, 48, 1
    W32 = L.arr(buf, 0, np.float32)
    W = W32.astype(np.float64)
    h32, lo32 = enc(W32, mode)  # the parts the layout splits W into
    Hr, Lr = h32.astype(np.float64), lo32.astype(np.float64)
    S, G = (1.0, G_BF) if mode == 2 else (S16, G16)
    X = np.random.RandomState(3).uniform(0.0, 300.0, size=(d, ncp)).astype(np.float32)
    Xd = H.DeviceArray.from_numpy(X, order="C")
    XT = H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16)
    Yd = H.DeviceArray.empty(d * ncp)
    YT = H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16)
    H.check(H.lib.ophip_wsq_bwd_layout(L.state.h, d, mode, buf))
    H.check(H.lib.ophip_wsq_bwd_prep(d, ncp, mode, Xd.h, XT.h))
    ev, el = get_cutoffs(H)
    vis, lo = (np.kron(m, np.ones((16, 32))) for m in
               tile_flags(bits16(h32, mode), d, mode, (ev, el)))
    df = 0.9999
    H.check(H.lib.ophip_wsq_bwd_step(L.state.h, d, mode, buf, ncp, np.float32(df), Xd.h, XT.h,
                                     Yd.h, YT.h))
    H.sync()
    xt = XT.download().reshape(2, -1)
    xt = np.stack([unfrag(xt[0], ncp, d), unfrag(xt[1], ncp, d)])
    xh = enc16(X.T, mode)
    xl = enc16((X.T - dec16(xh, mode)) * G, mode)
    assert np.array_equal(xt[0], xh) and np.array_equal(xt[1], xl)
    Y = Yd.download().reshape(d, ncp)
    Xh, Xl = dec16(xh, mode).T.astype(np.float64), dec16(xl, mode).T.astype(np.float64)
    full = df * (X.astype(np.float64) + (Hr @ Xh + (Hr @ Xl + Lr @ Xh) / G) / S)
    Hv, Lv = Hr * vis, Lr * lo  # the visited fragments' hi parts, the read lo parts
    ref = df * (X.astype(np.float64) + (Hv @ Xh + (Hv @ Xl + Lv @ Xh) / G) / S)
    # what the band rule leaves out: an entry under 2^ev, or the lo part of
    # an entry under 2^el, at most half an ulp of its hi part
    gone = np.abs(Hr + Lr / G) * (1 - vis) / S
    part = np.abs(Lr / G) * (vis - lo) / S
    assert gone.max() < 2.0**ev * (1 + 2.0**-10)
    assert part.max() <= max(2.0**(el - 12), 2.0**-40) * (1 + 2.0**-10)
    assert np.abs(ref - full).max() <= 1.01 * ((gone + part) @ np.abs(Xh + Xl / G)).max() + 1e-9
    if dense:  # the dense kernel reads every fragment
        ref = full
    assert np.abs(Y - ref).max() <= 1e-6 * np.abs(ref).max()
    exact = df * (X.astype(np.float64) + W @ X.astype(np.float64))
    assert np.abs(Y - exact).max() <= 2e-5 * np.abs(exact).max()
    yt = YT.download().reshape(2, -1)
    yt = np.stack([unfrag(yt[0], ncp, d), unfrag(yt[1], ncp, d)])
    yh = enc16(Y.T, mode)
    assert np.array_equal(yt[0], yh)
    yl = enc16((Y.T - dec16(yh, mode)) * G, mode)
    bad = np.argwhere(yt[1] != yl)
    assert len(bad) == 0, (len(bad), [(float(Y.T[tuple(b)]), int(yh[tuple(b)]), int(yt[1][tuple(b)]),
                                       int(yl[tuple(b)])) for b in bad[:4]])


def get_cutoffs(H):
    import ctypes as C
    ev, el = C.c_int(0), C.c_int(0)
    H.check(H.lib.ophip_wsq_get_bwd_cutoffs(C.byref(ev), C.byref(el)))
    return ev.value, el.value


@contextlib.contextmanager
def cutoffs(H, cut):
    """the band-sparse cutoffs (visit, lo exponents) inside the block; None
    keeps the library's default"""
    old = get_cutoffs(H)
    if cut is not None:
        H.check(H.lib.ophip_wsq_set_bwd_cutoffs(*cut))
    try:
        yield
    finally:
        H.check(H.lib.ophip_wsq_set_bwd_cutoffs(*old))


def tile_flags(hbits, d, mode, cut):
    """numpy: per (16-row tile, 32-column k tile) of the row-major hi copy,
    whether a step visits it (some |W| >= 2^visit) and whether it reads the
    lo part (some |W| >= 2^lo in FP16, whose copies hold W 2^15; 2^(lo-3)
    in BF16, whose copies hold W), on the decoded magnitudes; an FP16
    cutoff below its smallest subnormal counts every nonzero value"""
    ev, el = cut
    m = np.abs(dec16(hbits, mode).astype(np.float64)).reshape(d // 16, 16, d // 32, 32).max(axis=(1, 3))
    if mode == 2:
        return m >= 2.0**ev, m >= 2.0**(el - 3)
    return m >= max(2.0**(ev + 15), 2.0**-24), m >= max(2.0**(el + 15), 2.0**-24)


def tile_lists(hf, d, mode, cut):
    """numpy: per 16-row tile, the visited k tiles and their lo flags"""
    vis, lo = tile_flags(unfrag(hf, d, d), d, mode, cut)
    return [(np.nonzero(vis[i])[0], lo[i]) for i in range(d // 16)]


def backward_lattice(H, nsq, mode=1, nx=128, nr=2):
    L = Lattice(H, nx, nr, mode)
    p = plan(H, [L], 128)
    L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, nsq))
    H.sync()
    H.lib.ophip_wsq_plan_free(p)
    return L


@pytest.mark.parametrize("cut", [(-24, -12), (-40, -28), (-32, -16), (-20, -6)])
@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("nsq", [2, 9])
def test_wform_backward_tile_lists(hip, nsq, mode, cut):
    """the tile lists built with the fragment layout are the 16 x 32
    fragments whose hi part reaches 2^visit, each flagged where it reaches
    2^lo (FP16) or 2^(lo-3) (BF16) (its lo part is read) and, in a block's
    list, with the row tiles that visit it, numpy on the downloaded
    fragments; a banded W after few squarings skips most of them"""
    banner(f"wsq backward tile lists nsq={nsq} mode={mode} cutoffs={cut}")
    H = hip
    L = backward_lattice(H, nsq, mode)
    d, buf = L.d, nsq & 1
    with cutoffs(H, cut):
        H.check(H.lib.ophip_wsq_bwd_layout(L.state.h, d, mode, buf))
        H.sync()
    hf = L.arr(1 - buf, 2, np.uint16, "F").reshape(-1, order="F")
    ref = tile_lists(hf, d, mode, cut)
    h = H.lib.ophip_wsq_array(L.state.h, d, mode, 0, 7)
    nkt = d // 32
    nl = d // 16 + d // 64 + d // 32  # lists per 16-row tile, 64-row block, 32-row block
    tl = np.empty(nl * (nkt + 1), dtype=np.int32)
    H.check(H.lib.opcuda_memcpy_d2h(h, tl.ctypes.data, tl.nbytes))
    tl = tl.reshape(nl, nkt + 1)
    blocks = [[i] for i in range(d // 16)]
    blocks += [list(range(4 * b, 4 * b + 4)) for b in range(d // 64)]
    blocks += [[2 * b, 2 * b + 1] for b in range(d // 32)]
    for i, rows in enumerate(blocks):
        want = functools.reduce(np.union1d, [ref[k][0] for k in rows])
        lo = functools.reduce(np.logical_or, [ref[k][1] for k in rows])
        vm = sum(np.isin(want, ref[k][0]).astype(np.int64) << q for q, k in enumerate(rows))
        lm = sum(ref[k][1][want].astype(np.int64) << q for q, k in enumerate(rows))
        nr = len(rows)
        want = want + 65536 * lo[want] + ((vm << 17) + (lm << (17 + nr)) if nr > 1 else 0)
        assert tl[i, 0] == len(want), (i, tl[i, 0], len(want))
        assert np.array_equal(tl[i, 1:1 + tl[i, 0]], want), i
    frac = sum(len(r[0]) for r in ref) / ((d // 16) * nkt)
    frac_lo = sum(int(r[1].sum()) for r in ref) / ((d // 16) * nkt)
    print(f"visited fragment fraction {frac:.3f}, lo parts read {frac_lo:.3f}")
    if nsq == 2:
        assert frac < 0.5


@pytest.mark.parametrize("kernel", sorted(BWD_KERNELS))
@pytest.mark.parametrize("mode", [1, 2])
def test_wform_backward_values_from_copies(hip, mode, kernel):
    """a step given X = 0 takes X from its copies: the same bits as a step
    given X = hi + lo 2^-G (the copies' reconstruction, exact in FP32) as
    FP32 values; given Y = 0 it writes the same copies YT and leaves Y's
    buffer alone"""
    banner(f"wsq backward step from the copies mode={mode} kernel={BWD_KERNELS[kernel]}")
    H = hip
    old = H.lib.ophip_wsq_set_bwd_kernel(kernel)
    try:
        L = backward_lattice(H, 3, mode)
        d, ncp, buf = L.d, 48, 1
        H.check(H.lib.ophip_wsq_bwd_layout(L.state.h, d, mode, buf))
        S, G = (1.0, G_BF) if mode == 2 else (S16, G16)
        X = np.random.RandomState(4).uniform(0.0, 300.0, size=(d, ncp)).astype(np.float32)
        Xd = H.DeviceArray.from_numpy(X, order="C")
        XT = H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16)
        H.check(H.lib.ophip_wsq_bwd_prep(d, ncp, mode, Xd.h, XT.h))
        xt = XT.download().reshape(2, -1)
        xh, xl = unfrag(xt[0], ncp, d), unfrag(xt[1], ncp, d)  # [col][row]
        xrec = (dec16(xh, mode) + dec16(xl, mode) * np.float32(1.0 / G)).T.astype(np.float32)
        Xr = H.DeviceArray.from_numpy(np.ascontiguousarray(xrec), order="C")
        df = np.float32(0.9999)
        outs = []
        for xh_ in (Xr.h, 0):
            Yd = H.DeviceArray.empty(d * ncp)
            YT = H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16)
            H.check(H.lib.ophip_wsq_bwd_step(L.state.h, d, mode, buf, ncp, df, xh_, XT.h, Yd.h, YT.h))
            H.sync()
            outs.append((Yd.download(), YT.download()))
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        # Y = 0: the copies only
        sentinel = np.full(d * ncp, 7.0, np.float32)
        Yd = H.DeviceArray.from_numpy(sentinel)
        YT = H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16)
        H.check(H.lib.ophip_wsq_bwd_step(L.state.h, d, mode, buf, ncp, df, 0, XT.h, 0, YT.h))
        H.sync()
        assert np.array_equal(YT.download(), outs[1][1])
        assert np.array_equal(Yd.download(), sentinel)
    finally:
        H.lib.ophip_wsq_set_bwd_kernel(old)


@pytest.mark.parametrize("kernel", [2, 3, 4])
def test_wform_backward_batch_equals_single(hip, kernel):
    """one batched launch over three lattices (different squaring counts,
    values and discounts) gives each lattice the bits of its own single
    step, and stays within FP32 accumulation rounding of the dense kernel"""
    banner(f"wsq backward step batch vs single vs dense, kernel={BWD_KERNELS[kernel]}")
    H = hip
    old = H.lib.ophip_wsq_set_bwd_kernel(kernel)
    try:
        _batch_case(H)
    finally:
        H.lib.ophip_wsq_set_bwd_kernel(old)


def _batch_case(H):
    import ctypes as C
    mode, ncp, buf = 1, 48, 1
    nsqs, dfs = (3, 5, 9), (0.9999, 0.999, 0.99)
    lats = [backward_lattice(H, n, mode) for n in nsqs]  # odd counts end in buffer 1
    d, n = lats[0].d, len(lats)
    rs = np.random.RandomState(11)
    Xd, XT, Yd, YT = [], [], [], []
    for L in lats:
        H.check(H.lib.ophip_wsq_bwd_layout(L.state.h, d, mode, buf))
        Xd.append(H.DeviceArray.from_numpy(rs.uniform(0.0, 300.0, size=(d, ncp)).astype(np.float32),
                                           order="C"))
        XT.append(H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16))
        Yd.append(H.DeviceArray.empty(d * ncp))
        YT.append(H.DeviceArray.empty(2 * ncp * d, dtype=np.uint16))
        H.check(H.lib.ophip_wsq_bwd_prep(d, ncp, mode, Xd[-1].h, XT[-1].h))
    single = []
    for i, L in enumerate(lats):
        H.check(H.lib.ophip_wsq_bwd_step(L.state.h, d, mode, buf, ncp, np.float32(dfs[i]), Xd[i].h,
                                         XT[i].h, Yd[i].h, YT[i].h))
        H.sync()
        single.append((Yd[i].download().copy(), YT[i].download().copy()))
        H.lib.opcuda_ssetall(Yd[i].h, d * ncp, np.float32(-1.0), 1)
    H.check()

    def arr(ty, vals):
        return (ty * n)(*vals)
    H.check(H.lib.ophip_wsq_bwd_step_batch(
        n, arr(C.c_uint, [L.state.h for L in lats]), d, mode, buf, ncp, arr(C.c_float, dfs),
        arr(C.c_uint, [a.h for a in Xd]), arr(C.c_uint, [a.h for a in XT]),
        arr(C.c_uint, [a.h for a in Yd]), arr(C.c_uint, [a.h for a in YT])))
    H.sync()
    for i in range(n):
        y1, yt1 = single[i]
        assert np.array_equal(Yd[i].download(), y1), i
        assert np.array_equal(YT[i].download(), yt1), i
    kern = H.lib.ophip_wsq_set_bwd_kernel(0)
    try:
        for i, L in enumerate(lats):
            y1 = single[i][0]
            H.check(H.lib.ophip_wsq_bwd_step(L.state.h, d, mode, buf, ncp, np.float32(dfs[i]),
                                             Xd[i].h, XT[i].h, Yd[i].h, YT[i].h))
            H.sync()
            assert np.abs(Yd[i].download() - y1).max() <= 1e-6 * np.abs(y1).max(), i
    finally:
        H.lib.ophip_wsq_set_bwd_kernel(kern)


def test_wform_backward_layout_perm(hip):
    """the permuted fragment layout (ophip_wsq_bwd_layout_perm) is the plain
    layout of the permuted matrix P W P^T, bit for bit (fragments of the
    other buffer's column-major slots, unpacked)"""
    banner("wsq backward layout with a state permutation")
    H = hip
    mode, buf = 1, 1
    L = Lattice(H, 128, 2, mode)  # d = 256
    p = plan(H, [L], 128)
    L.first_step()
    H.check(H.lib.ophip_wsq_exec(p, 3))
    H.sync()
    H.lib.ophip_wsq_plan_free(p)
    d = L.d
    h32, lo32 = enc(L.arr(buf, 0, np.float32), mode)  # the parts the layout splits W into
    Hr, Lr = bits16(h32, mode), bits16(lo32, mode)
    perm = np.random.RandomState(7).permutation(d).astype(np.int32)
    pd = H.DeviceArray.from_numpy(perm)
    H.check(H.lib.ophip_wsq_bwd_layout_perm(L.state.h, d, mode, buf, pd.h))
    H.sync()
    hf = L.arr(1 - buf, 2, np.uint16, "F").reshape(-1, order="F")
    lf = L.arr(1 - buf, 4, np.uint16, "F").reshape(-1, order="F")
    assert np.array_equal(unfrag(hf, d, d), Hr[np.ix_(perm, perm)])
    assert np.array_equal(unfrag(lf, d, d), Lr[np.ix_(perm, perm)])
    # the identity permutation is the plain layout
    ident = H.DeviceArray.from_numpy(np.arange(d, dtype=np.int32))
    H.check(H.lib.ophip_wsq_bwd_layout_perm(L.state.h, d, mode, buf, ident.h))
    H.sync()
    a = L.arr(1 - buf, 2, np.uint16, "F").copy()
    H.check(H.lib.ophip_wsq_bwd_layout(L.state.h, d, mode, buf))
    H.sync()
    assert np.array_equal(L.arr(1 - buf, 2, np.uint16, "F"), a)


def e4m3_rne(x):
    """round to OCP FP8 e4m3 (bias 7, 3 mantissa bits, subnormals below
    2^-6, max 448), nearest even"""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    e = np.maximum(np.floor(np.log2(np.maximum(a, 1e-300))), -6.0)
    step = 2.0 ** (e - 3)
    return np.sign(x) * np.minimum(np.round(a / step) * step, 448.0)


def e4m3_dec(b):
    b = np.asarray(b, dtype=np.int64)
    s = np.where(b >> 7, -1.0, 1.0)
    ex, man = (b >> 3) & 15, b & 7
    return s * np.where(ex == 0, man * 2.0**-9, (1.0 + man / 8.0) * 2.0 ** (ex - 7.0))


def test_fp8_split_and_one_squaring(hip):
    """FP8 W form: the e4m3 copies are the round-to-nearest-even of the
    scaled FP32 values (bit-exact), scales are the powers of two that put
    each row / column maximum in [64, 128), and one squaring equals the
    same arithmetic in numpy (FP64 accumulation) to 1e-6 of max|W'|"""
    banner("fp8 W form: split + one squaring")
    from oplib_amd import opmodel as M
    H = hip
    c = M.config_small(nx=128, nr=2, nsq=12)
    d = c.nx * c.nr
    x, _ = M.grid(c)
    dr, vo = M.params(c)
    xd, ddv, vd = (H.DeviceArray.from_numpy(a) for a in (x, dr, vo))
    st = H.DeviceArray(H.lib.ophip_w8_bytes(d), np.uint8)
    dt = math.ldexp(c.DeltaT, -c.nsq)
    H.check(H.lib.ophip_sek_rs(c.nx, c.nr, xd.h, ddv.h, vd.h, c.q_off, dt, 2, H.lib.ophip_w8_w(st.h, d, 0)))
    H.check(H.lib.ophip_w8_exec(st.h, d, 1))
    H.sync()

    def rd(h, n, dt_):
        out = np.empty(n, dtype=dt_)
        H.check(H.lib.opcuda_memcpy_d2h(h, out.ctypes.data, out.nbytes))
        return out
    W0 = rd(H.lib.ophip_w8_w(st.h, d, 0), d * d, np.float32).reshape(d, d, order="F").astype(np.float64)
    W1 = rd(H.lib.ophip_w8_w(st.h, d, 1), d * d, np.float32).reshape(d, d, order="F").astype(np.float64)
    rs = rd(H.lib.ophip_w8_array(st.h, d, 4), d, np.float32).astype(np.float64)
    cs = rd(H.lib.ophip_w8_array(st.h, d, 5), d, np.float32).astype(np.float64)
    arrs = [unfrag(rd(H.lib.ophip_w8_array(st.h, d, k), d * d, np.uint8), d, d) for k in range(4)]
    Ah, Al, Bh, Bl = (e4m3_dec(a) for a in arrs)
    # scales: powers of two with max * scale in [64, 128)
    for sc, mx in ((rs, np.abs(W0).max(axis=1)), (cs, np.abs(W0).max(axis=0))):
        assert np.all(np.log2(sc) == np.round(np.log2(sc)))
        assert np.all((mx * sc >= 64) & (mx * sc < 128))
    # copies: hi = rne(W s), lo = rne((W s - hi) 16), bit-exact
    va, vb = W0 * rs[:, None], (W0 * cs[None, :]).T
    for v, hi, lo in ((va, Ah, Al), (vb, Bh, Bl)):
        v32 = v.astype(np.float32).astype(np.float64)
        assert np.array_equal(hi, e4m3_rne(v32))
        assert np.array_equal(lo, e4m3_rne(((v32 - hi).astype(np.float32) * 16.0).astype(np.float64)))
    # one squaring with the same arithmetic, then the FP64 diagonal reset
    P = (Ah @ Bh.T + (Ah @ Bl.T + Al @ Bh.T) / 16.0) / (rs[:, None] * cs[None, :])
    ref = (2.0 * W0 + P).astype(np.float32).astype(np.float64)
    idx = np.arange(d)
    off = ref.copy()
    off[idx, idx] = 0.0
    ref[idx, idx] = -off.sum(axis=1)
    err = np.abs(W1 - ref).max() / np.abs(ref).max()
    print(f"[err] fp8 one squaring vs emulation: {err:.3e}")
    assert err < 1e-6
    # and against the exact square: the e4m3 pair's ~2^-8 of |W|^2
    ex = 2.0 * W0 + W0 @ W0
    rel = np.abs(W1 - ex).max() / np.abs(W0 @ W0).max()
    print(f"[err] fp8 product vs exact: {rel:.3e} of max|W*W|")
    assert rel < 2e-2


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("tile", [64, 128])
def test_zero_cell_skip_same_bits(hip, mode, tile):
    """zero-cell skipping: the flags are exact (flag 0 <=> the cell's FP32
    values are all +0, checked cell by cell on every buffer), the skip is real on a
    banded lattice (most cells are zero after the first squarings), and the
    squarings give the same bits as the dense K loop"""
    banner(f"wsq zero-cell skip mode={mode} tile={tile}")
    H = hip
    lats = [Lattice(H, 512, 1, mode), Lattice(H, 128, 3, mode)]   # d = 512, 384

    def cells_zero(L, buf):
        d, nc = L.d, L.d // 64
        a = L.arr(buf, 0, np.uint32)  # +0 only: the flush stores +0
        return ~(a != 0).reshape(nc, 64, nc, 64).any(axis=(1, 3))

    def flags(L, buf):
        nc = L.d // 64
        h = H.lib.ophip_wsq_array(L.state.h, L.d, mode, buf, 6)
        out = np.empty(nc * nc, dtype=np.uint8)
        H.check(H.lib.opcuda_memcpy_d2h(h, out.ctypes.data, out.nbytes))
        return out.reshape(nc, nc)

    p = plan(H, lats, tile)
    old = H.lib.ophip_wsq_set_skip(1)
    oldk = H.lib.ophip_wsq_set_kcut(-192)  # every K cell that may be nonzero
    try:
        for nsq in (1, 2, 7):
            got = {}
            for skip in (1, 0):
                H.lib.ophip_wsq_set_skip(skip)
                for L in lats:
                    L.first_step()
                H.check(H.lib.ophip_wsq_exec(p, nsq))
                H.sync()
                got[skip] = [[L.arr(b, 0, np.uint32)] for L in lats for b in (0, 1)]
                if skip:
                    for L in lats:
                        for b in (0, 1):
                            z = cells_zero(L, b)
                            f = flags(L, b)
                            # flag 0 only on all-zero cells; all-zero cells flagged 0
                            # except diagonal cells (flagged by the fix)
                            assert not np.any((f == 0) & ~z)
                            off = ~np.eye(L.d // 64, dtype=bool)
                            assert np.array_equal((f == 0)[off], z[off])
                            # a flag 96 + E bounds the cell: max|W| in [2^(E-1), 2^E)
                            nc = L.d // 64
                            wmax = np.abs(L.arr(b, 0, np.float32)).reshape(nc, 64, nc, 64).max(axis=(1, 3))
                            m = (f != 0) & (f != 255)
                            e = f[m].astype(np.int64) - 96
                            assert np.all(wmax[m] < 2.0**e) and np.all(wmax[m] >= 2.0**(e - 1))
                            assert np.all(f[np.eye(nc, dtype=bool)] == 255)
                        zf = (flags(L, nsq & 1) == 0).mean()
                        print(f"[skip] d={L.d} nsq={nsq}: zero cells {zf:.2f}", flush=True)
                        if nsq <= 2 and L.cfg.nr == 1:
                            assert zf > 0.3
            for a, b in zip(got[1], got[0]):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y)
    finally:
        H.lib.ophip_wsq_set_skip(old)
        H.lib.ophip_wsq_set_kcut(oldk)
        H.lib.ophip_wsq_plan_free(p)


@pytest.mark.parametrize("mode", [1, 2])
def test_kcut_leaves_out_tiny_products(hip, mode):
    """the K-cell cutoff: at 2^-36 it moves no entry of W by more than the
    left-out products allow (at most 64 2^-36 per left-out cell and round,
    doubled by every later round: under 2^-20 here) against every K cell;
    a cutoff that bites (2^-8) changes W"""
    banner(f"wsq K-cell cutoff mode={mode}")
    H = hip
    L = Lattice(H, 512, 1, mode)  # d = 512, one band
    p = plan(H, [L], 128)
    oldk = H.lib.ophip_wsq_set_kcut(-36)
    try:
        out = {}
        for kc in (-192, -36, -8):
            H.lib.ophip_wsq_set_kcut(kc)
            L.first_step()
            H.check(H.lib.ophip_wsq_exec(p, 9))
            H.sync()
            out[kc] = L.arr(1, 0, np.float32).astype(np.float64)
        full = out[-192]
        assert np.abs(out[-36] - full).max() <= 2.0**-20
        assert np.abs(out[-8] - full).max() > 0.0  # a cutoff that bites changes W
    finally:
        H.lib.ophip_wsq_set_kcut(oldk)
        H.lib.ophip_wsq_plan_free(p)


@pytest.mark.parametrize("mode", [1, 2])
def test_later_solve_same_bits(hip, mode):
    """a plan's later solve, whose first round writes only where the previous
    solve left values (the destination's flags), gives the bits of a fresh
    state's solve: after a wide-band solve (10 squarings) and a backward
    layout (fragment copies written into the other buffer's column-major
    copy slots), a narrow one (4 squarings): both buffers' W, every bit"""
    banner(f"wsq later solve after a wide one and a layout, mode={mode}")
    H = hip
    A, B = Lattice(H, 128, 2, mode), Lattice(H, 128, 2, mode)
    d = A.d
    pa, pb = plan(H, [A], 128), plan(H, [B], 128)
    try:
        A.first_step()
        H.check(H.lib.ophip_wsq_exec(pa, 10))
        H.check(H.lib.ophip_wsq_bwd_layout(A.state.h, d, mode, 0))
        A.first_step()
        H.check(H.lib.ophip_wsq_exec(pa, 4))
        B.first_step()
        H.check(H.lib.ophip_wsq_exec(pb, 4))
        H.sync()
        nc = d // 64
        for buf in (0, 1):
            fa = np.empty(nc * nc, dtype=np.uint8)
            H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(A.state.h, d, mode, buf, 6),
                                            fa.ctypes.data, fa.nbytes))
            assert np.array_equal(A.arr(buf, 0, np.uint32), B.arr(buf, 0, np.uint32)), buf
            # the flags of both states agree too
            fb = np.empty(nc * nc, dtype=np.uint8)
            H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(B.state.h, d, mode, buf, 6),
                                            fb.ctypes.data, fb.nbytes))
            assert np.array_equal(fa, fb), buf
    finally:
        H.lib.ophip_wsq_plan_free(pa)
        H.lib.ophip_wsq_plan_free(pb)


@pytest.mark.parametrize("mode", [1, 2])
def test_dense_solve_after_layout_same_bits(hip, mode):
    """the dense K loop (ophip_wsq_set_skip(0)) reads every cell of W:
    after an odd-count solve and a backward layout (fragment copies written
    into buffer 0's column-major copy slots), a dense solve gives a fresh
    state's bits in W of both buffers"""
    banner(f"wsq dense solve after a layout, mode={mode}")
    H = hip
    A, B = Lattice(H, 512, 1, mode), Lattice(H, 512, 1, mode)  # one band: zero cells
    d = A.d
    pa, pb = plan(H, [A], 128), plan(H, [B], 128)
    old = H.lib.ophip_wsq_set_skip(1)
    try:
        A.first_step()
        H.check(H.lib.ophip_wsq_exec(pa, 3))
        H.check(H.lib.ophip_wsq_bwd_layout(A.state.h, d, mode, 1))
        H.sync()
        H.lib.ophip_wsq_set_skip(0)
        for nsq in (1, 2):
            A.first_step()
            H.check(H.lib.ophip_wsq_exec(pa, nsq))
            B.first_step()
            H.check(H.lib.ophip_wsq_exec(pb, nsq))
            H.sync()
            for buf in (0, 1):
                assert np.array_equal(A.arr(buf, 0, np.uint32), B.arr(buf, 0, np.uint32)), (nsq, buf)
            H.check(H.lib.ophip_wsq_bwd_layout(A.state.h, d, mode, nsq & 1))
            H.check(H.lib.ophip_wsq_bwd_layout(B.state.h, d, mode, nsq & 1))
    finally:
        H.lib.ophip_wsq_set_skip(old)
        H.lib.ophip_wsq_plan_free(pa)
        H.lib.ophip_wsq_plan_free(pb)


def test_forward_steps_skip_zero_cells(hip):
    """the forward steps p_k = p_{k-1} (I + W) visiting only W's flagged
    cells (ophip_fwd_set_cells) give the bits of the dense steps, on a
    banded W with zero cells"""
    banner("forward steps with zero-cell skipping")
    import ctypes as C
    H, mode, nsq = hip, 1, 2
    L = backward_lattice(H, nsq, mode, nx=512, nr=1)  # one regime: a band, zero cells off it
    d, buf = L.d, nsq & 1
    K = H.lib.ophip_wsq_w(L.state.h, d, mode, buf)
    fl = H.lib.ophip_wsq_array(L.state.h, d, mode, buf, 6)
    flags = np.empty((d // 64) ** 2, dtype=np.uint8)
    H.check(H.lib.opcuda_memcpy_d2h(fl, flags.ctypes.data, flags.nbytes))
    assert (flags == 0).any() and (flags != 0).any()
    nmat, ks = 5, np.array([80.0, 100.0, 120.0], dtype=np.float32)
    x = np.linspace(0.0, 400.0, d).astype(np.float32)
    xd = H.DeviceArray.from_numpy(x)
    outs = []
    for cells in (False, True):
        P0 = np.zeros((nmat + 1) * d, dtype=np.float32)
        P0[300] = 1.0  # x = 235, in the money for every strike
        Pd, od = H.DeviceArray.from_numpy(P0), H.DeviceArray.empty(nmat * len(ks))
        plan = H.lib.ophip_fwd_plan(1, (C.c_int * 1)(d), (C.c_uint * 1)(K), (C.c_int * 1)(1),
                                    (C.c_uint * 1)(Pd.h), (C.c_uint * 1)(xd.h), (C.c_uint * 1)(od.h),
                                    (C.c_float * 1)(0.02), np.float32(0.1), nmat,
                                    (C.c_float * len(ks))(*ks), len(ks))
        assert plan
        if cells:
            H.check(H.lib.ophip_fwd_set_cells(plan, (C.c_uint * 1)(fl)))
        H.check(H.lib.ophip_fwd_exec(plan))
        H.sync()
        outs.append((Pd.download().copy(), od.download().copy()))
        H.lib.ophip_fwd_plan_free(plan)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.all(np.isfinite(outs[0][1])) and outs[0][1].max() > 0


@pytest.mark.parametrize("nx,nr", [(64, 2), (16, 8)])
def test_wform_backward_layout_regime(hip, nx, nr):
    """the regime layout (ophip_wsq_bwd_layout_regime, LDS tiles) is the
    permuted layout with perm[x + nx r] = nr x + r, bit for bit, and gives
    the same tile lists"""
    banner(f"wsq backward layout, regime permutation nx={nx} nr={nr}")
    H, mode, buf = hip, 1, 1
    L = backward_lattice(H, 3, mode, nx=nx, nr=nr)
    d = L.d
    perm = np.array([(i % nx) * nr + i // nx for i in range(d)], dtype=np.int32)
    pd = H.DeviceArray.from_numpy(perm)
    nkt = d // 32
    nl = d // 16 + d // 64 + d // 32
    out = []
    for regime in (False, True):
        if regime:
            H.check(H.lib.ophip_wsq_bwd_layout_regime(L.state.h, d, mode, buf, nx, nr))
        else:
            H.check(H.lib.ophip_wsq_bwd_layout_perm(L.state.h, d, mode, buf, pd.h))
        H.sync()
        tl = np.empty(nl * (nkt + 1), dtype=np.int32)
        H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(L.state.h, d, mode, 0, 7), tl.ctypes.data,
                                        tl.nbytes))
        out.append((L.arr(1 - buf, 2, np.uint16, "F").copy(), L.arr(1 - buf, 4, np.uint16, "F").copy(), tl))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_wform_backward_layout_regime_batch(hip):
    """one batched regime layout over three lattices (different squaring
    counts) gives each the fragment copies and tile lists of its own
    single-lattice layout"""
    banner("wsq backward layout, regime permutation, batch of 3 vs single")
    import ctypes as C
    H, mode, buf, nx, nr = hip, 1, 1, 64, 2
    lats = [backward_lattice(H, n, mode, nx=nx, nr=nr) for n in (3, 5, 9)]  # odd counts: buffer 1
    d = lats[0].d
    nkt = d // 32
    nl = d // 16 + d // 64 + d // 32

    def grab(L):
        tl = np.empty(nl * (nkt + 1), dtype=np.int32)
        H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(L.state.h, d, mode, 0, 7), tl.ctypes.data,
                                        tl.nbytes))
        return (L.arr(1 - buf, 2, np.uint16, "F").copy(), L.arr(1 - buf, 4, np.uint16, "F").copy(), tl)

    st = (C.c_uint * 3)(*[L.state.h for L in lats])
    H.check(H.lib.ophip_wsq_bwd_layout_regime_batch(3, st, d, mode, buf, nx, nr))
    H.sync()
    got = [grab(L) for L in lats]
    for L, g in zip(lats, got):
        H.check(H.lib.ophip_wsq_bwd_layout_regime(L.state.h, d, mode, buf, nx, nr))
        H.sync()
        for a, b in zip(g, grab(L)):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("nm", [0, 1])
@pytest.mark.parametrize("skip", [1, 0])
def test_fused_assembly_matches_two_kernel_path(hip, mode, nm, skip):
    """ophip_wsq_assemble_rs + ophip_wsq_exec_ex(presplit) against
    ophip_sek_rs_batch + ophip_wsq_exec: the same bits in both W buffers and
    both flag arrays after every call of a sequence (the fused kernel skips
    zero cells already flagged zero from the second call on), and the same
    partial sums in every cell flagged nonzero after a call without
    squarings; with the sparse and with the dense K loop"""
    banner(f"wsq fused assembly vs two-kernel path mode={mode} nm={nm} skip={skip}")
    import ctypes as C
    H = hip
    shapes = [(64, 2), (96, 4), (128, 2), (256, 8)]  # d = 128, 384, 256, 2048
    la = [Lattice(H, nx, nr, mode) for nx, nr in shapes]
    lb = [Lattice(H, nx, nr, mode) for nx, nr in shapes]
    pa, pb = plan(H, la, 64), plan(H, lb, 64)
    n = len(shapes)

    def arrs(L):
        return [(L.arr(b, 0, np.uint32).copy(), grab_flags(L, b)) for b in (0, 1)]

    def grab_flags(L, b):
        nc = L.d // 64
        f = np.empty(nc * nc, dtype=np.uint8)
        H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(L.state.h, L.d, mode, b, 6),
                                        f.ctypes.data, f.nbytes))
        return f.reshape(nc, nc)

    def grab_part(L):
        nc = L.d // 64
        p = np.empty(nc * L.d, dtype=np.float64)
        H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(L.state.h, L.d, mode, 0, 5),
                                        p.ctypes.data, p.nbytes))
        return p.reshape(nc, L.d)  # [column block][row]

    nx = (C.c_int * n)(*[L.cfg.nx for L in la])
    nr = (C.c_int * n)(*[L.cfg.nr for L in la])
    xv = lambda ls: (C.c_uint * n)(*[L.xd.h for L in ls])
    dr = lambda ls: (C.c_uint * n)(*[L.dd.h for L in ls])
    vo = lambda ls: (C.c_uint * n)(*[L.vd.h for L in ls])
    q = (C.c_double * n)(*[L.cfg.q_off for L in la])
    old_skip = H.lib.ophip_wsq_set_skip(skip)
    try:
        for k, (nsq, scale) in enumerate([(3, 1.0), (2, 0.5), (0, 0.25), (4, 1.0), (0, 2.0)]):
            dt = (C.c_double * n)(*[L.dt * scale for L in la])
            w0 = (C.c_uint * n)(*[H.lib.ophip_wsq_array(L.state.h, L.d, mode, 0, 0) for L in la])
            H.check(H.lib.ophip_sek_rs_batch(n, nx, nr, xv(la), dr(la), vo(la), q, dt, 6 if nm else 2, w0))
            H.check(H.lib.ophip_wsq_exec(pa, nsq))
            st = (C.c_uint * n)(*[L.state.h for L in lb])
            H.check(H.lib.ophip_wsq_assemble_rs(n, nx, nr, xv(lb), dr(lb), vo(lb), q, dt, nm, mode, st))
            H.check(H.lib.ophip_wsq_exec_ex(pb, nsq, 1))
            H.sync()
            for A, B in zip(la, lb):
                for (wa, fa), (wb, fb) in zip(arrs(A), arrs(B)):
                    assert np.array_equal(wa, wb), (k, A.d)
                    assert np.array_equal(fa, fb), (k, A.d)
                if nsq == 0:
                    f0 = grab_flags(A, 0)
                    sel = np.repeat(f0.T != 0, 64, axis=1)  # [column block][row]
                    assert np.array_equal(grab_part(A)[sel], grab_part(B)[sel]), (k, A.d)
        assert H.lib.ophip_wsq_exec_ex(pb, 1, 2) == 0x10003  # unknown flag bits
        H.lib.ophip_set_status(0)
    finally:
        H.lib.ophip_wsq_set_skip(old_skip)
        H.lib.ophip_wsq_plan_free(pa)
        H.lib.ophip_wsq_plan_free(pb)


def test_fused_assembly_rejects_bad_arguments(hip):
    """ophip_wsq_assemble_rs refuses what it cannot assemble (status
    0x10003, nothing launched): a size that is not a multiple of 64, a
    missing array, a state handle too small for the layout, a bad mode"""
    banner("wsq fused assembly argument checks")
    import ctypes as C
    H = hip
    L = Lattice(H, 64, 2, 1)  # d = 128
    c = L.cfg

    def call(nx=64, nr=2, state=None, mode=1, xv=True):
        n = 1
        a = lambda t, v: (t * n)(v)
        st = a(C.c_uint, L.state.h if state is None else state)
        return H.lib.ophip_wsq_assemble_rs(n, a(C.c_int, nx), a(C.c_int, nr),
                                           a(C.c_uint, L.xd.h) if xv else None, a(C.c_uint, L.dd.h),
                                           a(C.c_uint, L.vd.h), a(C.c_double, c.q_off), a(C.c_double, L.dt),
                                           1, mode, st)

    try:
        assert call() == 0
        H.sync()
        assert call(nx=65, nr=2) == 0x10003  # d = 130
        assert call(xv=False) == 0x10003
        small = H.DeviceArray(1024, np.uint8)
        assert call(state=small.h) == 0x10003
        assert call(mode=3) == 0x10003
        assert H.lib.ophip_wsq_assemble_rs(0, None, None, None, None, None, None, None, 1, 1, None) == 0x10003
    finally:
        H.lib.ophip_set_status(0)
