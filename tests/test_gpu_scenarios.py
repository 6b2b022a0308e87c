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
    ("lv1d", lambda M: M.config_lv1d(), 2e-2),
    ("rs2048", lambda M: M.config_rs2048(), 2e-2),
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
