"""CPU tests that pin the oracle to things other than itself: known
answers, closed forms, library routines and brute force (no GPU)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from _cases import sv_case
from oplib_amd import opmodel as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---- MT19937 -------------------------------------------------------------

def test_mt19937_known_answers():
    # C++ standard [rand.predef]: default-seeded mt19937's 10000th output
    r = O.mt_raw(5489, 10000)
    assert r[0] == 3499211612
    assert r[9999] == 4123659995


def test_mt19937_fixture_and_numpy():
    g = np.load(os.path.join(GOLD, "mt19937.npz"))
    assert np.array_equal(O.mt_raw(5489, 2000), g["s5489"])
    assert np.array_equal(O.mt_raw(5490, 2000), g["s5490"])
    # numpy's legacy RandomState is MT19937 with init_genrand for int seeds
    rs = np.random.RandomState(1)
    assert np.array_equal(rs.randint(0, 2**32, size=2000, dtype=np.uint64).astype(np.uint32),
                          g["s1"])


def test_uniform_mapping_exact():
    raw = O.mt_raw(7, 5000)
    u = O.mt_uniforms(7, 5000)
    assert np.array_equal(u, ((raw >> 8).astype(np.float64) * 2.0**-24).astype(np.float32))
    assert u.min() >= 0.0 and u.max() < 1.0


# ---- lattice / generator -------------------------------------------------

def test_grid_spot_on_node():
    x, x0 = O.xgrid(256, 0.0, 400.0, 100.0)
    assert x0 == 64
    assert x[x0] == 100.0
    assert np.allclose(np.diff(x), 1.5625, rtol=0, atol=1e-12)


def test_stencil_inverse_closed_form():
    x = np.cumsum(np.r_[0.0, np.random.default_rng(0).uniform(0.5, 2.0, 40)])
    inv = O.invm(x).reshape(-1, 2, 2, order="C")  # [node][c][r] since col-major per node
    for i in range(1, len(x) - 1):
        dm, dp = x[i - 1] - x[i], x[i + 1] - x[i]
        mat = np.array([[dm, dp], [dm * dm, dp * dp]])
        mi = inv[i].T  # back to (r, c)
        assert np.allclose(mat @ mi, np.eye(2), atol=1e-12)


def test_lv_rates_uniform_closed_form():
    x, _ = O.xgrid(64, 0.0, 400.0, 100.0)
    h = x[1] - x[0]
    dr, vo = O.lv_params_c1(x, 0.02)
    g = O.lv_generator(x, dr, vo)
    for i in range(1, 63):
        s2, mu = vo[i] ** 2, dr[i]
        a = (s2 / h**2 - mu / h) / 2
        b = (s2 / h**2 + mu / h) / 2
        assert g[i, i - 1] == pytest.approx(a, rel=1e-12)
        assert g[i, i + 1] == pytest.approx(b, rel=1e-12)
    assert np.all(g[0] == 0) and np.all(g[-1] == 0)  # absorbing
    assert np.abs(g.sum(axis=1)).max() < 1e-9
    # exactly tridiagonal
    assert np.count_nonzero(np.triu(g, 2)) == 0 and np.count_nonzero(np.tril(g, -2)) == 0


def test_regime_generator_structure_and_marginal():
    c = M.config_small(nx=12, nr=3, nsq=14)
    x, x0, g = O.model_generator(c)
    nx, nr = 12, 3
    assert np.abs(g.sum(axis=1)).max() < 1e-9
    # off-diagonal regime blocks are q_off * I
    for r1 in range(nr):
        for r2 in range(nr):
            blk = g[r1 * nx:(r1 + 1) * nx, r2 * nx:(r2 + 1) * nx]
            if r1 != r2:
                assert np.array_equal(blk, c.q_off * np.eye(nx))
    # the regime marginal of U is exp(Q t): sum U over the x of each regime
    U = O.kernel_fp64(c)[2]
    Q = c.q_off * (np.ones((nr, nr)) - nr * np.eye(nr))
    import scipy.linalg as sl
    E = sl.expm(Q * c.DeltaT)
    for r1 in range(nr):
        y = 5 + nx * r1
        marg = [U[y, r2 * nx:(r2 + 1) * nx].sum() for r2 in range(nr)]
        # first-order discretisation error only (dt0 = DeltaT 2^-14)
        assert np.allclose(marg, E[r1], atol=1e-6)


# ---- planner -------------------------------------------------------------

def test_planner_formula():
    import ctypes as C
    out = C.c_double()
    n = O.lib.or_plan_niter(1.0 / 16, 1000.0, 2.0**-20, C.byref(out))
    assert n == 16 and out.value == 2.0**-20
    n = O.lib.or_plan_niter(1.0 / 252, 1000.0, 2.0**-20, C.byref(out))
    assert n == math.ceil(math.log2((1 / 252) / 2.0**-20)) == 13
    assert out.value == pytest.approx((1 / 252) * 2.0**-13, rel=1e-15)
    # the 0.5/maxdiag cap binds when the generator is stiff
    n = O.lib.or_plan_niter(1.0, 1e7, 1.0, C.byref(out))
    assert out.value <= 0.5 / 1e7


# ---- first step, squaring ------------------------------------------------

def test_first_step_and_square_chain_vs_numpy():
    c = M.config_small(nx=20, nr=2, nsq=6)
    x, x0, g = O.model_generator(c)
    dt0 = math.ldexp(c.DeltaT, -c.nsq)
    U0 = O.first_step(g, dt0)
    assert np.array_equal(U0, np.eye(g.shape[0]) + dt0 * g)
    assert np.abs(U0.sum(axis=1) - 1).max() < 1e-14
    U = O.square_chain(U0, c.nsq)
    ref = np.linalg.matrix_power(U0, 2**c.nsq)
    assert np.abs(U - ref).max() < 1e-13
    # a Q-matrix (non-negative off-diagonal rates) gives a stochastic kernel
    off = g - np.diag(np.diag(g))
    if off.min() >= 0:
        assert U.min() >= 0


def test_fine_grid_kernel_is_stochastic():
    # zero drift: the central-difference rates are all non-negative (with
    # drift, nodes below r*h/sigma^2 get a negative down-rate, as in the
    # reference's stencil)
    c = M.config_small(nx=64, nr=2, nsq=8)
    c.rate = 0.0
    x, x0, g = O.model_generator(c)
    assert (g - np.diag(np.diag(g))).min() >= 0
    U = O.kernel_fp64(c)[2]
    assert U.min() >= 0
    assert np.abs(U.sum(axis=1) - 1).max() < 1e-12


def test_square_chain_converges_to_expm_first_order():
    f = np.load(os.path.join(GOLD, "expm_small.npz"))
    L, T, E = f["L"], float(f["T"]), f["expm"]
    errs = []
    for nsq in (8, 10, 12):
        U = O.square_chain(O.first_step(L, math.ldexp(T, -nsq)), nsq)
        errs.append(np.abs(U - E).max())
    # first-order method: error shrinks ~4x per two extra squarings
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3
    assert errs[2] < 1e-4


def test_sek_f32_formula():
    rng = np.random.default_rng(1)
    d = 33
    g = rng.standard_normal((d, d)).astype(np.float32)
    out = np.zeros(d * d, dtype=np.float32)
    O.lib.or_sek_f32(d, np.ascontiguousarray(g.reshape(-1, order="F")), np.float32(0.01), out)
    ref = (np.eye(d) + np.float64(np.float32(0.01)) * g.astype(np.float64)).astype(np.float32)
    # single rounding of a fused multiply-add equals rounding the exact value
    assert np.array_equal(out.reshape(d, d, order="F"), ref)


def rnd32(q):
    """a rational rounded to the nearest float32, ties to even (exact)"""
    from fractions import Fraction
    f = np.float32(float(q))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        e = abs(Fraction(float(c)) - q)
        if best is None or e < best[0] or (e == best[0] and int(c.view(np.uint32)) % 2 == 0):
            best = (e, c)
    return best[1]


def test_sekd_f32_each_step_one_rounding():
    # sek.cu:62-103: num = dt gen + id (one FMA), den = dt r + 1 (one FMA;
    # lambda: (dt r) lambda + 1), an IEEE quotient; checked in exact
    # rational arithmetic on a sample of entries
    from fractions import Fraction as Fr
    rng = np.random.default_rng(12)
    d = 40
    g = (rng.standard_normal((d, d)) * 50).astype(np.float32)
    r = (rng.random(d) * 0.05).astype(np.float32)
    dt, lam = np.float32(1e-3), np.float32(0.7)
    gf = np.ascontiguousarray(g.reshape(-1, order="F"))
    out = np.zeros(d * d, np.float32)
    outl = np.zeros(d * d, np.float32)
    O.lib.or_sekd_f32(d, gf, dt, r, None, out)
    lamv = np.array([lam], np.float32)
    O.lib.or_sekd_f32(d, gf, dt, r, lamv.ctypes.data, outl)
    for e in rng.integers(0, d * d, 150).tolist() + [0, d + 1, 5 * d + 5]:
        y1, y2 = e % d, e // d
        num = rnd32(Fr(float(dt)) * Fr(float(gf[e])) + (1 if y1 == y2 else 0))
        den = rnd32(Fr(float(dt)) * Fr(float(r[y1])) + 1)
        assert out[e] == rnd32(Fr(float(num)) / Fr(float(den))), e
        dtr = rnd32(Fr(float(dt)) * Fr(float(r[y1])))
        denl = rnd32(Fr(float(dtr)) * Fr(float(lam)) + 1)
        assert outl[e] == rnd32(Fr(float(num)) / Fr(float(denl))), e


def test_sglvg_f32_each_step_one_rounding():
    # sglvg.cu:42-99: zeroed blocks, boundary rows zero, interior rows
    # down = invm[4y+2] v^2 + (invm[4y] drift) and up likewise (one FMA
    # each after the rounded first product), diagonal -(down + up)
    from fractions import Fraction as Fr
    d, ni = 50, 2
    x, _ = O.xgrid(d, 0.0, 400.0, 100.0)
    inv = O.invm(x).astype(np.float32)
    dr, vo = O.lv_params_c1(x, 0.02)
    pars = np.concatenate([dr, 1.3 * dr, vo, 0.9 * vo]).astype(np.float32)
    gen = np.full(ni * d * d, np.nan, np.float32)
    O.lib.or_sglvg_f32(ni, d, inv, pars, gen)
    G = gen.reshape(ni, d, d).transpose(0, 2, 1)  # [i][y0][y1]
    assert np.all(G[:, 0] == 0) and np.all(G[:, -1] == 0)
    for i in range(ni):
        for y in range(1, d - 1):
            r0, v = pars[y + d * i], pars[ni * d + y + d * i]
            r1 = rnd32(Fr(float(v)) * Fr(float(v)))
            F = lambda a: Fr(float(a))  # noqa: E731
            lo = rnd32(F(inv[4 * y + 2]) * F(r1) + F(rnd32(F(inv[4 * y]) * F(r0))))
            hi = rnd32(F(inv[4 * y + 3]) * F(r1) + F(rnd32(F(inv[4 * y + 1]) * F(r0))))
            assert G[i, y, y - 1] == lo and G[i, y, y + 1] == hi, (i, y)
            assert G[i, y, y] == -rnd32(F(lo) + F(hi))
            row = G[i, y].copy()
            row[[y - 1, y, y + 1]] = 0
            assert np.all(row == 0)
    # and it is the FP64 generator to FP32 accuracy
    ref = O.lv_generator(x, dr, vo)
    assert np.abs(G[0] - ref).max() <= 1e-6 * np.abs(ref).max()


# ---- cumulative tables and sampler ---------------------------------------

def test_sck_properties_and_float64():
    rng = np.random.default_rng(2)
    nm, d = 3, 40
    t = rng.random((nm, d, d)).astype(np.float32)
    t[1, 5, :] = 0
    t[1, 5, 7] = 1.0
    c = O.sck_f32(t).reshape(nm, d, d)  # [m][col][row] (column-major per matrix)
    for m in range(nm):
        C = c[m].T  # [row][y]
        assert np.all(C[:, -1] == 1.0)
        assert np.all(np.diff(C, axis=1) >= 0)
        ref = np.cumsum(t[m].astype(np.float64), axis=1) / t[m].astype(np.float64).sum(1)[:, None]
        assert np.abs(C - ref).max() < 1e-5


def test_sck_raw_sum_lowers_on_negative_entry():
    # sck.cu:64-71 adds the raw entries: a negative one lowers the running
    # sum (hand-worked row 0.5, -0.25, 0.75 -> sums 0.5, 0.25, 1.0)
    d = 3
    t = np.zeros((1, d, d), np.float32)
    t[0, 0] = [0.5, -0.25, 0.75]
    t[0, 1] = [0.0, 1.0, 0.0]
    t[0, 2] = [0.25, 0.25, 0.5]
    c = O.sck_f32(t).reshape(d, d).T  # [row][y]
    assert np.array_equal(c[0], np.float32([0.5, 0.25, 1.0]))
    assert np.array_equal(c[1], np.float32([0.0, 1.0, 1.0]))
    assert np.array_equal(c[2], np.float32([0.25, 0.5, 1.0]))


def test_sck_rows_is_the_transposed_table():
    rng = np.random.default_rng(21)
    nm, d = 2, 77
    t = (rng.random((nm, d, d)) ** 3).astype(np.float32)
    t[1, 4, :] = 0
    t[1, 4, 9] = 2.0
    col = O.sck_f32(t).reshape(nm, d, d)  # [m][y][row]
    rows = np.zeros(nm * d * d, np.float32)
    src = np.ascontiguousarray(np.stack([a.reshape(-1, order="F") for a in t]).reshape(-1))
    O.lib.or_sck_rows_f32(d, nm, src, rows)
    assert np.array_equal(rows.reshape(nm, d, d), np.transpose(col, (0, 2, 1)))


@pytest.mark.parametrize("d,plus", [(200, 0), (130, 1), (64, 0), (33, 1)])
def test_sck_rows_two_level_exact_on_dyadic(d, plus):
    # entries k * 2^-12 with k < 64: every partial sum of a row is exact in
    # FP32, so any summation order gives the same bits as the sequential one
    rng = np.random.default_rng(d)
    t = (rng.integers(0, 64, (2, d, d)) * 2.0**-12).astype(np.float32)
    if plus:
        t[:, np.arange(d), np.arange(d)] -= 1.0  # W = U - I: plus_identity restores U
    t[0, 3, :] = 0
    t[0, 3, 5] = 2.0**-12 if not plus else 2.0**-12
    two = O.sck_rows2_f32(t, 64, plus)
    base = t.astype(np.float64) + (plus * np.eye(d))[None]
    want = np.zeros(2 * d * d, np.float32)
    src = np.ascontiguousarray(np.stack([a.reshape(-1, order="F") for a in base.astype(np.float32)])
                               .reshape(-1))
    O.lib.or_sck_rows_f32(d, 2, src, want)
    assert np.array_equal(two, want.reshape(2, d, d))


def test_sck_rows_two_level_properties():
    rng = np.random.default_rng(23)
    nm, d = 2, 300
    t = (rng.random((nm, d, d)) ** 6).astype(np.float32)
    t[0, 7, 100] = -1e-9  # a rounding-level negative counts as zero
    c = O.sck_rows2_f32(t, 64)
    assert np.all(c[:, :, -1] == 1.0)
    assert np.all(np.diff(c, axis=2) >= 0)
    t64 = np.maximum(t.astype(np.float64), 0.0)
    ref = np.cumsum(t64, axis=2) / t64.sum(2)[:, :, None]
    assert np.abs(c - ref).max() < 4 * d * 2.0**-24
    # chunk = d is the sequential order
    seq = O.sck_rows2_f32(t, d)
    rows = np.zeros(nm * d * d, np.float32)
    src = np.ascontiguousarray(np.stack([a.reshape(-1, order="F") for a in t]).reshape(-1))
    O.lib.or_sck_rows_f32(d, nm, src, rows)
    assert np.array_equal(seq, rows.reshape(nm, d, d))


def test_sgemv_gather_vs_numpy():
    # the gsgemv4 layout (gsgemv4/Program.cs:236-300): Z holds 2*nz vectors
    # of d, column lists are element offsets, padded with the sentinel
    # 2*nz*d; inside a 16-group the stores stop at the first sentinel
    rng = np.random.default_rng(24)
    d, nz = 70, 12
    A = rng.standard_normal((d, d)).astype(np.float32)  # not symmetric
    Z = rng.standard_normal((d, 2 * nz)).astype(np.float32)
    sent = 2 * nz * d
    col0 = np.full(32, sent, np.int32)
    col1 = np.full(32, sent, np.int32)
    pairs = {0: (1, 13), 1: (4, 14), 2: (2, 15), 3: (9, 16),         # group 0: 4 live,
             5: (3, 17),                                             # then a sentinel at 4
             16: (5, 18), 17: (0, 19), 18: (11, 20), 19: (6, 21)}    # group 1: 4 live
    for c, (a, b) in pairs.items():
        col0[c], col1[c] = a * d, b * d
    want = Z.astype(np.float64).copy()
    for c in (0, 1, 2, 3, 16, 17, 18, 19):
        a, b = pairs[c]
        want[:, b] = A.astype(np.float64) @ Z[:, a].astype(np.float64)
    got = np.ascontiguousarray(Z.reshape(-1, order="F"))
    O.lib.or_sgemv_gather(d, np.ascontiguousarray(A.reshape(-1, order="F")), got, 32, col0, col1, sent)
    got = got.reshape(d, 2 * nz, order="F")
    assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()
    assert np.array_equal(got[:, 17], Z[:, 17])  # after the group's first sentinel
    # a transposed operand fails
    bad = A.T.astype(np.float64) @ Z[:, 1].astype(np.float64)
    assert np.abs(got[:, 13] - bad).max() > 1e-2


def test_lv_params_cev_closed_forms():
    e = math.e
    x = np.array([0.0, 100.0 / e, 100.0, 100.0 * e, 400.0])
    s0, beta, r, a, b = 0.3, 0.6, 0.03, 0.2, 0.1
    dr, vo = O.lv_params_cev(x, r, s0, beta, a, b)
    assert np.allclose(dr, r * x, rtol=1e-15, atol=0)
    sig = np.divide(vo, x, out=np.zeros_like(vo), where=x > 0)
    assert vo[0] == 0.0
    assert sig[2] == pytest.approx(s0, rel=1e-15)                                  # sigma(100) = sigma0
    assert sig[3] == pytest.approx(s0 * math.exp(beta - 1) * (1 + a + b), rel=1e-13)  # l = 1
    assert sig[1] == pytest.approx(s0 * math.exp(1 - beta) * (1 - a + b), rel=1e-13)  # l = -1
    l4 = math.log(4.0)
    assert sig[4] == pytest.approx(s0 * 4.0 ** (beta - 1) * (1 + a * l4 + b * l4 * l4), rel=1e-13)
    # a negative factor is floored at 1e-4 (sigma, in price units 1e-4 x)
    dr, vo = O.lv_params_cev(x, r, s0, beta, -2.0, 0.0)
    assert vo[3] == pytest.approx(1e-4 * x[3], rel=1e-15)
    assert vo[1] > 1e-4 * x[1]


@pytest.mark.parametrize("d", [1, 2, 3, 64, 127, 128, 129, 255, 256, 257, 511, 512, 513, 700, 1023])
def test_reference_gpu_bisection_is_the_inverse_cdf(d):
    # smc.cu:224-238 (and the 512/256/128 variants): the fixed-count
    # bisection from lb = 0 returns min{j : u < C[j]} clamped to d-1,
    # state 0 included, the definition the build's sampler implements
    rng = np.random.default_rng(d)
    P = rng.random((d, d)) ** 4
    P[:, rng.integers(0, d, d // 3 + 1)] = 0.0  # zero-mass states
    P[0, :] = 0.0
    P[0, 0] = 1.0  # row 0: all mass on state 0
    P[P.sum(1) == 0, 0] = 1.0
    c = O.sck_f32(P.astype(np.float32)[None])
    us = list(rng.random(400).astype(np.float32)) + [np.float32(0.0), np.float32(1.0 - 2.0**-24)]
    for yp in sorted({0, d // 2, d - 1}):
        row = c[yp::d]
        extra = [row[j] for j in range(0, d, max(1, d // 7))]  # u exactly on a table value
        for u in us + extra:
            got = O.lib.or_sample_ref_bisect(c, d, yp, float(u))
            want = O.lib.or_sample_colmajor(c, d, yp, float(u))
            assert got == want, (d, yp, float(u))
    assert O.lib.or_sample_ref_bisect(c, d, 0, 0.5) == 0  # state 0 is reachable


def test_sampler_linear_is_inverse_cdf():
    rng = np.random.default_rng(3)
    d = 17
    p = rng.random(d)
    p[[0, 4, 5]] = 0
    cdf = np.minimum(np.cumsum(p) / p.sum(), 1).astype(np.float32)
    for u in list(rng.random(200).astype(np.float32)) + [np.float32(0.0), cdf[3], cdf[6]]:
        j = O.lib.or_sample_linear(cdf, d, float(u))
        brute = next((i for i in range(d) if u < cdf[i]), d - 1)
        assert j == brute
        assert p[j] > 0 or j == d - 1


def test_mc_chain_bisect_equals_linear():
    # d <= 64 scans linearly, larger d bisects: both must agree on the same
    # monotone tables (pad a 48-state chain to 96 states with zero mass)
    rng = np.random.default_rng(4)
    d1, d2 = 48, 96
    P = rng.random((d1, d1)) ** 8
    P /= P.sum(1)[:, None]
    big = np.zeros((d2, d2))
    big[:d1, :d1] = P
    big[d1:, d1:] = np.eye(d2 - d1)
    c1 = O.sck_f32(P.astype(np.float32)[None])
    c2 = O.sck_f32(big.astype(np.float32)[None])
    m_k = np.zeros(5, dtype=np.int32)
    y1 = O.mc_chain(3, 5489, 2, 5, d1, 7, c1, m_k)
    y2 = O.mc_chain(3, 5489, 2, 5, d2, 7, c2, m_k)
    assert np.array_equal(y1, y2)


def test_mc_chain_stream_mapping():
    # uniform rows: y = floor(u * 64) exactly, so every draw is visible.
    # Stream s serves `per` paths (scenario s*per + j) in groups of 64; at
    # step k a group's n paths take the stream's next n outputs in order
    d = 64
    P = np.full((d, d), 1.0 / d)
    c = O.sck_f32(P.astype(np.float32)[None])
    for nstreams, per, nk in ((2, 128, 3), (3, 25, 4), (2, 100, 2)):
        y = O.mc_chain(nstreams, 100, 0, nk, d, 0, c, np.zeros(nk, dtype=np.int32), per=per)
        assert y.shape == (nk, nstreams * per)
        for s in range(nstreams):
            u = O.mt_uniforms(100 + s, per * nk)
            pos = 0
            for j0 in range(0, per, 64):
                n = min(64, per - j0)
                for k in range(nk):
                    want = np.floor(u[pos:pos + n] * 64).astype(np.int16)
                    assert np.array_equal(y[k, s * per + j0:s * per + j0 + n], want)
                    pos += n
            assert pos == per * nk


# ---- pricing -------------------------------------------------------------

def test_black_scholes_bstest_value():
    # tests/bstest Program.cs: S=K=100, r=0.05, sigma=0.25, T=1
    assert O.bs_call(100, 100, 0.05, 0.25, 1.0) == pytest.approx(12.336, abs=5e-4)
    # put-call parity through the call formula
    c = O.bs_call(100, 110, 0.03, 0.2, 0.5)
    p = c - 100 + 110 * math.exp(-0.015)
    assert p > 0


def test_operator_method_matches_black_scholes():
    # constant vol, one regime: the lattice price approaches Black-Scholes
    c = M.config_small(nx=256, nr=1, nsq=14, nsteps=52)
    c.sigma_r[0] = 0.25
    c.rate = 0.05
    c.q_off = 0.0
    c.xhi = 400.0
    p = O.solve_fp64(c)[0, 0]
    for i, K in enumerate(c.strike_list()):
        assert p[i] == pytest.approx(O.bs_call(100, K, 0.05, 0.25, 1.0), abs=0.03)


def test_backward_equals_propagation():
    c1 = M.config_small(nx=40, nr=2, nsteps=5)
    c2 = M.config_small(nx=40, nr=2, nsteps=5, price_kind=M.OPM_PRICE_PROPAGATE)
    p1 = O.solve_fp64(c1)[:, 0]
    p2 = O.solve_fp64(c2)[:, -1]
    assert np.allclose(p1, p2, rtol=1e-12, atol=1e-12)


def test_golden_prices_reproduce():
    f = np.load(os.path.join(GOLD, "prices_small_bi.npz"))
    c = M.config_small(nx=64, nr=2)
    assert np.allclose(O.solve_fp64(c), f["prices"], rtol=1e-12, atol=1e-12)


# ---- surrogate pieces ----------------------------------------------------

def test_gelu_grad_finite_difference():
    for x in np.linspace(-4, 4, 17):
        fd = (O.lib.or_gelu(x + 1e-6) - O.lib.or_gelu(x - 1e-6)) / 2e-6
        assert O.lib.or_gelu_grad(x) == pytest.approx(fd, abs=1e-7)


def test_adam_step_closed_form():
    n = 5
    p = np.linspace(-1, 1, n)
    g = np.linspace(0.5, -0.5, n)
    m = np.zeros(n)
    v = np.zeros(n)
    p0 = p.copy()
    O.lib.or_adam_step(n, p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 1)
    # first step: m_hat = g, v_hat = g^2 -> p -= lr * g/(|g| + eps)
    assert np.allclose(p, p0 - 1e-3 * g / (np.abs(g) + 1e-8), atol=1e-15)


def test_mlp_forward_closed_form():
    # one linear layer is W x + b; two layers put an exact-erf GELU between
    rng = np.random.RandomState(3)
    W1, b1 = rng.randn(4, 3), rng.randn(4)
    W2, b2 = rng.randn(2, 4), rng.randn(2)
    x = rng.randn(6, 3)
    out, _ = O.mlp_forward([W1], [b1], x)
    assert np.allclose(out, x @ W1.T + b1, atol=1e-13)
    from scipy.special import erf
    h = x @ W1.T + b1
    h = 0.5 * h * (1.0 + erf(h / math.sqrt(2.0)))
    out, _ = O.mlp_forward([W1, W2], [b1, b2], x)
    assert np.allclose(out, h @ W2.T + b2, atol=1e-12)


def test_mlp_backward_finite_differences():
    # d/dtheta sum(out .* dout) against central differences in FP64
    rng = np.random.RandomState(4)
    dims = [5, 7, 6, 3]
    Ws = [rng.randn(dims[i + 1], dims[i]) * 0.7 for i in range(3)]
    bs = [rng.randn(dims[i + 1]) * 0.3 for i in range(3)]
    x = rng.randn(4, 5)
    dout = rng.randn(4, 3)
    dW, db = O.mlp_backward(Ws, bs, x, dout)

    def f():
        return float(np.sum(O.mlp_forward(Ws, bs, x)[0] * dout))
    h = 1e-6
    for l in range(3):
        for (i, j) in [(0, 0), (dims[l + 1] - 1, dims[l] - 1), (1, 2)]:
            old = Ws[l][i, j]
            Ws[l][i, j] = old + h
            fp = f()
            Ws[l][i, j] = old - h
            fm = f()
            Ws[l][i, j] = old
            assert dW[l][i, j] == pytest.approx((fp - fm) / (2 * h), rel=1e-6, abs=1e-8)
        for i in (0, dims[l + 1] - 1):
            old = bs[l][i]
            bs[l][i] = old + h
            fp = f()
            bs[l][i] = old - h
            fm = f()
            bs[l][i] = old
            assert db[l][i] == pytest.approx((fp - fm) / (2 * h), rel=1e-6, abs=1e-8)


def test_bbf_vol_closed_forms():
    # constant local vol: the implied vol is that vol, and the control
    # variate is the Black-Scholes price (bstest: 12.336)
    p = [0.25, 1.0, 0.05, 0.0, 0.0]
    assert O.bbf_vol(p, 100.0, 120.0) == pytest.approx(0.25, abs=1e-15)
    cv = O.surr_control_variate([p], [100.0], 10, 0.1, 100.0)
    assert cv[0, 9] == pytest.approx(O.bs_call(100, 100, 0.05, 0.25, 1.0), abs=1e-12)
    assert cv[0, 9] == pytest.approx(12.336, abs=5e-4)
    # CEV: 1/sigma = int_0^1 ds / (s0 (K/S)^{s(beta-1)}) = ((K/S)^(1-beta) - 1) / (s0 (1-beta) ln(K/S))
    for beta, K in ((0.5, 130.0), (0.3, 80.0), (0.8, 101.0)):
        p = [0.3, beta, 0.0, 0.0, 0.0]
        inv = ((K / 100.0) ** (1 - beta) - 1.0) / (0.3 * (1 - beta) * math.log(K / 100.0))
        assert O.bbf_vol(p, 100.0, K) == pytest.approx(1.0 / inv, rel=1e-12)


def test_sv_generator_invariants():
    """rows sum to zero, off-diagonal rates are nonnegative, and every row
    with a neighbour to correct through carries the prescribed drift"""
    nx, nr, xv, rv, sd, sv, vv, vd, jm, jp = case = sv_case()
    g = O.sv_generator(*case)
    d = nx * nr
    off = g - np.diag(np.diag(g))
    assert off.min() >= 0.0
    assert np.abs(g.sum(axis=1)).max() <= 1e-12 * np.abs(g).max()
    drift = (g * (xv[None, :] - xv[:, None])).sum(axis=1)
    xi = np.arange(d) % nx
    interior = (xi > 0) & (xi < nx - 1)
    assert np.allclose(drift[interior], sd[interior], rtol=1e-10, atol=1e-12 * np.abs(sd).max())


def test_sv_generator_second_moment_closed_form():
    """without jumps and regime dynamics each interior row solves the 2x2
    system (zero drift, variance svol^2) and then moves the drift onto one
    neighbour: sum L (x1-x0)^2 = svol^2 + sdrift * (dp if sdrift > 0 else dm)"""
    nx, nr, xv, rv, sd, sv, vv, vd, jm, jp = case = sv_case(jumps=False, regimes=False)
    g = O.sv_generator(*case)
    d = nx * nr
    for y in range(d):
        x = y % nx
        if x == 0 or x == nx - 1:
            continue
        dm, dp = xv[y - 1] - xv[y], xv[y + 1] - xv[y]
        m2 = (g[y] * (xv - xv[y]) ** 2).sum()
        want = sv[y] ** 2 + sd[y] * (dp if sd[y] > 0 else dm)
        assert abs(m2 - want) <= 1e-10 * want, (y, m2, want)
        # only the two price neighbours and the diagonal
        assert np.count_nonzero(g[y]) <= 3


def test_sv_generator_regime_rates_and_jumps():
    """interior regimes: the two regime rates carry the regime drift and
    variance; a rate up to regime r+1 is mirrored by jumps into regime r+1
    of the same total, with the exponential profile of the scaled sizes"""
    nx, nr, xv, rv, sd, sv, vv, vd, jm, jp = case = sv_case()
    g = O.sv_generator(*case)
    seen = 0
    for y in range(nx * nr):
        x, r = y % nx, y // nx
        if not (0 < r < nr - 1):
            continue
        bm, bp = g[y, y - nx], g[y, y + nx]
        em, ep = rv[y - nx] - rv[y], rv[y + nx] - rv[y]
        if bm > 0 and bp > 0:
            assert abs(bm * em + bp * ep - vd[y]) <= 1e-12 * (abs(vd[y]) + 1)
            assert abs(bm * em ** 2 + bp * ep ** 2 - vv[y] ** 2) <= 1e-12
        if bp > 0 and 2 < x < nx - 2:
            up = g[y, nx * (r + 1):nx * (r + 2)].copy()
            jumps = up.sum() - up[x] - up[x - 1]
            assert up[x - 1] == 0.0
            assert abs(jumps - bp) <= 1e-12 * bp, (y, jumps, bp)
            h = xv[1] - xv[0]
            assert np.isclose(up[x + 2] / up[x + 1], np.exp(-h / (jp[y] * xv[y])), rtol=1e-12)
            assert np.isclose(up[x - 3] / up[x - 2], np.exp(-h / (jm[y] * xv[y])), rtol=1e-12)
            seen += 1
    assert seen > 10


def test_tgfsr_is_mt19937_with_its_parameters():
    """the generic twisted-GFSR lane code with MT19937's parameters is
    MT19937: the C++-standard known answers and the oracle's own stream"""
    st = O.tgfsr_seed([5489], 624)
    out = O.tgfsr_run(O.MT19937, st, 10000, [0x9908B0DF], [0x9D2C5680], [0xEFC60000])[:, 0]
    assert out[0] == 3499211612 and out[9999] == 4123659995
    st1 = O.tgfsr_seed([1], 624)
    assert np.array_equal(O.tgfsr_run(O.MT19937, st1, 2000, [0x9908B0DF], [0x9D2C5680], [0xEFC60000])[:, 0],
                          O.mt_raw(1, 2000))


def test_mc_hash_sampler_is_the_inverse_cdf():
    """the reference's CPU sampler (bracket tables + bisection) returns the
    inverse-CDF state min{j : u < C[y, j]} of the same uniforms: replayed
    draw by draw with a linear scan on random stochastic tables"""
    rng = np.random.default_rng(11)
    nm, d, nk, nscen, nb = 2, 37, 5, 64, 3
    P = rng.random((nm, d, d)) ** 6
    P /= P.sum(axis=2, keepdims=True)
    m_k = np.array([0, 1, 1, 0, 1], dtype=np.int32)
    y, ct, _ = O.mc_hash(P, nb, nscen, 7, m_k, 1234)
    C = np.cumsum(P, axis=2)
    C = np.minimum(C / C[:, :, -1:], 1.0)
    for b in range(nb):
        st = O.mt_raw(1234 + b, nk * nscen).reshape(nk, nscen)
        yp = np.full(nscen, 7)
        for k in range(nk):
            u = st[k].astype(np.float64) / 4294967296.0
            row = C[m_k[k]][yp]
            want = np.minimum((u[:, None] < row).argmax(axis=1) + 0 * yp, d - 1)
            want = np.where((u[:, None] < row).any(axis=1), want, d - 1)
            assert np.array_equal(y[b, k], want), (b, k)
            yp = want
