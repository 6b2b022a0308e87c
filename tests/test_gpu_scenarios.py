"""End-to-end GPU solves through libopmodel -> libophip, checked against the
FP64 oracle's prices (golden fixtures written by tests/golden/make_golden.py
from the same discretisation: grid, dt0, squaring count, steps).

Error metric: |p - p64| / max(|p64|, floor), floor = 1e-3 * max|p64|
(stated with every result).  The plain FP32 method is the reference's
device method; its error is what FP32 repeated squaring gives and is
bounded here, not hidden.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def banner(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M(hip):
    from oplib_amd import opmodel
    return opmodel


def golden(name):
    return np.load(os.path.join(GOLD, f"prices_{name}.npz"))["prices"]


def report(name, p, g):
    e = O.rel_err(p, g)
    print(f"[err] {name}: max {e.max():.3e} rms {np.sqrt(np.mean(e**2)):.3e} "
          f"(floor 1e-3*max|p64|)", flush=True)
    return e


@pytest.mark.parametrize("name,builder,bound", [
    ("small_bi", lambda M: M.config_small(nx=64, nr=2), 2e-3),
    ("small_prop", lambda M: M.config_small(nx=64, nr=2, price_kind=1, nsteps=6), 2e-3),
    # the reference's FP32 method drifts: rows of U(DeltaT) sum to 1 + e with
    # e ~ 2^16 FP32 roundings, amplified again by the pricing steps; an FP32
    # emulation on the CPU gives ~1e-2 (LV, 16 propagations) and ~1.2e-1
    # (2048 states, 252 backward steps).  Bounds at ~2.5x those.
    ("lv1d", lambda M: M.config_lv1d(), 5e-2),
    ("rs2048", lambda M: M.config_rs2048(), 3e-1),
])
def test_plain_fp32_vs_fp64(M, name, builder, bound):
    banner(f"scenario {name}")
    cfg = builder(M)
    s = M.Solver(cfg)
    s.solve()
    p = s.prices()
    g = golden(name)
    assert p.shape == g.shape
    e = report(name, p, g)
    assert np.all(np.isfinite(p))
    assert e.max() < bound
    # kernel sanity: rows of U(DeltaT) are sub-stochastic up to FP32 drift
    U = s.kernel()
    rs = U.astype(np.float64).sum(axis=1)
    print(f"[rowsum] {name}: max |rowsum-1| {np.abs(rs - 1).max():.3e}", flush=True)
    assert np.abs(rs - 1).max() < 0.05
    s.close()


# max relative error bounds; a CPU emulation of the same arithmetic gives
#              FP16x1   FP16x3   BF16x3
#   rs2048     6.0e-5   3.2e-7   3.0e-6
#   lv1d       3.7e-4   2.8e-7   5.6e-6
#   rs_monthly 1.9e-4   4.3e-7   5.2e-6
# so the single FP16 product meets the 1e-4 bar only for the 252-day
# backward induction; the three-product forms meet it everywhere.
WFORM_BOUNDS = {
    ("rs2048", 1): 1.5e-4, ("rs2048", 2): 1e-5, ("rs2048", 3): 3e-5,
    ("lv1d", 1): 1e-3, ("lv1d", 2): 1e-5, ("lv1d", 3): 3e-5,
    # rs_monthly adds 11 FP32 propagation products: ~2e-5 measured for FP16x3
    ("rs_monthly", 1): 1e-3, ("rs_monthly", 2): 5e-5, ("rs_monthly", 3): 5e-5,
    # FP8 e4m3 hi/lo with power-of-two scales (about 8 bits a pair): measured
    # 8.0e-4 / 2.1e-3 / 1.5e-3, above the 1e-4 bar; bounds at ~2.5x
    ("rs2048", 4): 2e-3, ("lv1d", 4): 5e-3, ("rs_monthly", 4): 4e-3,
}


@pytest.mark.parametrize("name,prec", sorted(WFORM_BOUNDS))
def test_wform_mixed_precision_vs_fp64(M, name, prec):
    """U = I + W squarings on FP16/BF16 matrix cores with row-sum
    renormalisation, against the FP64 oracle (the target bar is 1e-4)."""
    bound = WFORM_BOUNDS[(name, prec)]
    banner(f"wform {name} prec={prec}")
    cfg = {"rs2048": M.config_rs2048, "lv1d": M.config_lv1d,
           "rs_monthly": M.config_rs_monthly}[name]()
    cfg.precision = prec
    s = M.Solver(cfg)
    s.solve()
    p = s.prices()
    e = report(f"{name} prec={prec}", p, golden(name))
    assert np.all(np.isfinite(p))
    assert e.max() < bound
    U = s.kernel().astype(np.float64)
    rs = np.abs(U.sum(axis=1) - 1).max()
    print(f"[rowsum] {name} prec={prec}: {rs:.3e}", flush=True)
    assert rs < 1e-5
    s.close()


def test_wform_graph_replay(M):
    banner("wform graph replay")
    cfg = M.config_small(nx=128, nr=2, nsteps=20)
    cfg.precision = 1
    s = M.Solver(cfg)
    s.solve()
    eager = s.prices().copy()
    s.capture()
    for rep in range(3):
        s.launch()
        assert np.array_equal(s.prices(), eager), rep
        s.solve()
        assert np.array_equal(s.prices(), eager)
    s.close()


def test_kernel_vs_fp64_small(M):
    banner("kernel small")
    cfg = M.config_small(nx=64, nr=2, nsq=10)
    s = M.Solver(cfg)
    s.solve()
    U = s.kernel().astype(np.float64)
    _, _, U64 = O.kernel_fp64(cfg)
    err = np.abs(U - U64).max()
    print(f"[err] kernel small max abs {err:.3e}", flush=True)
    assert err < 1e-4
    s.close()


def test_graph_replay_matches_eager(M):
    banner("graph replay")
    cfg = M.config_small(nx=128, nr=2, nsteps=20)
    s = M.Solver(cfg)
    s.solve()
    eager = s.prices().copy()
    s.capture()
    for rep in range(3):
        s.launch()
        p = s.prices()
        assert np.array_equal(p, eager), rep
        # eager work between replays
        s.solve()
        assert np.array_equal(s.prices(), eager)
    s.close()


def test_propagation_consistent_with_backward(M):
    banner("prop vs backward")
    c1 = M.config_small(nx=64, nr=2, nsteps=5)
    c2 = M.config_small(nx=64, nr=2, nsteps=5, price_kind=1)
    a = M.Solver(c1)
    a.solve()
    b = M.Solver(c2)
    b.solve()
    pa, pb = a.prices()[:, 0], b.prices()[:, -1]
    assert np.max(np.abs(pa - pb)) < 1e-4 * np.max(np.abs(pa))
    a.close()
    b.close()


@pytest.mark.parametrize("prec,bound", [(0, 0.5), (1, 1e-3), (2, 1e-4), (3, 1e-4)])
def test_sweep_batch_vs_fp64(M, prec, bound):
    """64 CEV lattices of different size in one batch (persistent W-form
    squaring launches for prec > 0), every price against the FP64 oracle."""
    banner(f"sweep prec={prec}")
    cfgs = M.config_sweep(precision=prec)
    b = M.Batch(cfgs)
    b.solve()
    p = b.prices()
    g = golden("sweep")
    assert p.shape == g.shape
    e = report(f"sweep prec={prec}", p, g)
    assert np.all(np.isfinite(p))
    assert e.max() < bound
    b.close()


def test_sweep_batch_equals_single(M):
    banner("sweep batch vs single")
    cfgs = M.config_sweep(n=6, precision=2)
    b = M.Batch(cfgs, tile=64)
    b.solve()
    pb = b.prices()
    b.close()
    for i, c in enumerate(cfgs):
        s = M.Batch([c], tile=64)
        s.solve()
        assert np.array_equal(s.prices()[0], pb[i]), i
        s.close()


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_group_equals_single_small(M, prec):
    """opm_group: three scenarios of one small regime lattice (different
    vols, regime rates, rates) squared in one launch per round and stepped
    in one launch per step give each member the bits of its own opm_solve;
    within the method's bound of the FP64 oracle"""
    banner(f"group small prec={prec}")
    cfgs = []
    for i, (lv, q, r) in enumerate([(1.0, 0.5, 0.03), (0.8, 1.0, 0.0), (1.3, 0.2, 0.05)]):
        c = M.config_small(nx=64, nr=2, nsteps=12)
        for k in range(c.nr):
            c.sigma_r[k] *= lv
        c.q_off, c.rate, c.precision = q, r, prec
        cfgs.append(c)
    g = M.Group(cfgs)
    g.solve()
    pg = g.prices()
    for i, c in enumerate(cfgs):
        s = M.Solver(c)
        s.solve()
        assert np.array_equal(s.prices(), pg[i]), i
        s.close()
        e = report(f"group small prec={prec} member {i}", pg[i], O.solve_fp64(c))
        assert e.max() < (1e-3 if prec == 1 else 3e-5)
    g.close()


def bwd_cutoffs(H):
    import ctypes as C
    ev, el = C.c_int(0), C.c_int(0)
    H.check(H.lib.ophip_wsq_get_bwd_cutoffs(C.byref(ev), C.byref(el)))
    return ev.value, el.value


def test_group_rs2048_vs_fp64(M):
    """four scenarios of the 2048-state regime lattice (the workload, then
    seeded vol levels, regime rates and rates) priced together in FP16 x3:
    with the tightest band cutoffs (2^-40, 2^-28) each within 1e-5 of the
    FP64 oracle (fixture); the default cutoffs (2^-24, 2^-12) leave out
    parts of W that move these prices by at most 9.2e-6 relative in exact
    arithmetic (FP64 backward induction with and without them), so the
    default prices stay within 1.5e-5 of the tight ones; the first one the
    bits of the single-scenario solve when both use the same backward-step
    kernel (64-row blocks), and within FP32 accumulation-order rounding of
    the single solve's own default kernel; a captured group replays to the
    same bits"""
    banner("group rs2048 x4 FP16x3")
    from oplib_amd import ophip as H
    gold = np.load(os.path.join(GOLD, "prices_rs2048_scenarios.npz"))["prices"]
    cfgs = M.config_rs2048_scenarios(4)
    dflt = bwd_cutoffs(H)
    assert dflt == (-24, -12)
    H.check(H.lib.ophip_wsq_set_bwd_cutoffs(-40, -28))
    try:
        g = M.Group(cfgs)
        g.solve()
        pt = g.prices()
        g.close()
    finally:
        H.check(H.lib.ophip_wsq_set_bwd_cutoffs(*dflt))
    et = report("group rs2048 x4, cutoffs 2^-40 2^-28", pt, gold[:4])
    assert np.all(np.isfinite(pt)) and et.max() < 1e-5
    g = M.Group(cfgs)
    g.solve()
    pg = g.prices()
    e = report("group rs2048 x4", pg, gold[:4])
    assert np.all(np.isfinite(pg)) and e.max() < 1e-4
    assert O.rel_err(pg, pt).max() < 1.5e-5
    s = M.Solver(cfgs[0])
    s.solve()
    p1 = s.prices()
    assert np.abs(p1 - pg[0]).max() <= 1e-5 * np.abs(p1).max()
    old = H.lib.ophip_wsq_set_bwd_kernel(3)
    try:
        s.solve()
        g.solve()
        assert np.array_equal(s.prices(), g.prices()[0])
    finally:
        H.lib.ophip_wsq_set_bwd_kernel(old)
    s.close()
    g.solve()
    g.capture()
    for rep in range(2):
        g.launch()
        assert np.array_equal(g.prices(), pg), rep
        g.solve()
        assert np.array_equal(g.prices(), pg)
    g.close()


@pytest.mark.parametrize("n", [16, 32])
def test_group_rs2048_headline_sizes_vs_fp64(M, n):
    """the headline method at its group sizes: 16 and 32 scenarios of the
    2048-state regime lattice in one group launch, every one of the
    n x 8 x 41 prices within 1e-4 of the FP64 oracle (fixture of 32 FP64
    solves), replayed from the captured graph"""
    banner(f"group rs2048 x{n} FP16x3 (headline)")
    gold = np.load(os.path.join(GOLD, "prices_rs2048_scenarios.npz"))["prices"]
    assert len(gold) >= n
    g = M.Group(M.config_rs2048_scenarios(n))
    g.capture()
    g.launch()
    pg = g.prices()
    e = report(f"group rs2048 x{n}", pg, gold[:n])
    assert np.all(np.isfinite(pg)) and e.max() <= 1e-4
    g.close()


def test_mcplan_exec_vs_fp64(M):
    """the reference's step plan executed on the device (first steps,
    squaring rounds and chain products through opcuda_ssqmm pointer tables)
    against the FP64 oracle's first steps, squarings and products of the
    same plan: weekly steps over three local-vol generators"""
    banner("mc plan exec")
    from oplib_amd import ophip as H
    nx = 64
    x, _ = O.xgrid(nx, 0.0, 400.0, 100.0)
    gens = []
    for scale in (0.8, 1.0, 1.3):
        dr, vo = O.lv_params_c1(x, 0.02)
        gens.append(O.lv_generator(x, dr, scale * vo))
    maxdiag = [np.abs(np.diag(g)).max() for g in gens]
    p = M.McPlan([7, 14, 21, 28, 35, 42, 49, 56, 63], [30, 60, 90], 1, maxdiag)
    d = nx
    gd = H.DeviceArray.from_numpy(np.concatenate([g.astype(np.float32).reshape(-1, order="F") for g in gens]))
    xt = H.DeviceArray.empty(p.nj * d * d)
    xb = H.DeviceArray.empty(max(p.nj, p.nm) * d * d)
    tp = H.DeviceArray.empty(p.nm * d * d)
    p.exec(d, gd.h, xt.h, xb.h, tp.h)
    got = tp.download().reshape(p.nm, d, d).transpose(0, 2, 1)  # column-major blocks
    Uj = [O.square_chain(O.first_step(gens[p.i_j[j]], p.dt_j[j]), int(p.niter_j[j])) for j in range(p.nj)]
    for m, ch in enumerate(p.chains):
        want = Uj[ch[0]]
        for j in ch[1:]:
            want = want @ Uj[j]
        err = np.abs(got[m] - want).max()
        print(f"[plan] m={m} chain={list(map(int, ch))} max |U_m - FP64| = {err:.2e}", flush=True)
        assert err < 2e-5, (m, err)
    p.close()


def test_create_failures_keep_their_status(M):
    """a refused configuration, and a refused group, leave the reason in the
    library's status (not the status of the cleanup after it)"""
    banner("model-layer failures keep their status")
    from oplib_amd import ophip as H
    c = M.config_small(nx=64, nr=2)
    c.nstrikes = 0
    with pytest.raises(H.OphipError, match="0x10003"):
        M.Solver(c)
    a = M.config_small(nx=128, nr=2)
    a.precision = M.OPM_PREC_F16X3
    b = M.config_small(nx=128, nr=2, nsteps=9)
    b.precision = M.OPM_PREC_F16X3
    with pytest.raises(H.OphipError, match="0x10003"):
        M.Group([a, b])  # different step counts
