"""Shared inputs of the oracle and GPU tests (no tests here)."""
import numpy as np

from oracle import oracle as O


def sv_case(nx=40, nr=5, jumps=True, regimes=True, seed=0):
    """per-state inputs of a stochastic-volatility lattice: prices on
    [0, 400], vol regimes rval 0.1..0.5, x vol rval * x, drift 0.03 x,
    mean-reverting regime drift, relative jump sizes"""
    rng = np.random.default_rng(seed)
    x, _ = O.xgrid(nx, 0.0, 400.0, 100.0)
    xv = np.tile(x, nr)
    rv = np.repeat(np.linspace(0.1, 0.5, nr), nx)
    d = nx * nr
    sd = 0.03 * xv * rng.uniform(0.5, 1.5, d) * rng.choice([-1.0, 1.0], d)
    sv = rv * xv
    vv = (0.3 if regimes else 0.0) * np.ones(d)
    vd = (2.0 * (0.3 - rv) if regimes else 0.0 * rv)
    jm = (0.1 if jumps else 0.0) * np.ones(d)
    jp = (0.05 if jumps else 0.0) * np.ones(d)
    return nx, nr, xv, rv, sd, sv, vv, vd, jm, jp
