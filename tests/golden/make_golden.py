"""Writes the committed golden fixtures under tests/golden/.

    python tests/golden/make_golden.py [surrogate | rs_scenarios]

* prices_*.npz   FP64 prices of the named workloads from the CPU oracle
                 (the reference's FP64 host method restated: generator,
                 I + L dt0, repeated squaring, propagation / backward
                 induction), with the configuration echoed for checking;
* expm_small.npz scipy.linalg.expm(L T) on a tiny lattice, used to bound
                 the first-order discretisation error of the squaring chain;
* mt19937.npz    the first 2000 outputs of std::mt19937-compatible
                 streams (seeds 5489, 5490, 1) from the oracle, checked here
                 against the C++-standard known answers first.

The reference holds no golden vectors (its only automated checks compare
against MKL at run time), so these fixtures are produced by the oracle and
cross-checked against closed forms and library routines before writing.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oplib_amd import opmodel as M  # noqa: E402  (configuration builders only)


def cfg_echo(c):
    return dict(nx=c.nx, nr=c.nr, DeltaT=c.DeltaT, nsq=c.nsq, nsteps=c.nsteps,
                price_kind=c.price_kind, rate=c.rate, q_off=c.q_off,
                strikes=np.array(c.strike_list()), sigma=np.array(c.sigma_list()))


def write_prices(name, c):
    p = O.solve_fp64(c)
    np.savez(os.path.join(HERE, f"prices_{name}.npz"), prices=p, **cfg_echo(c))
    print(name, p.shape, float(p.min()), float(p.max()))


def write_surrogate():
    """FP64 prices of the first 8 held-out surrogate draws (1024-node CEV
    lattice with skew/curvature, 20 squarings to U(0.1), forward chain)"""
    from oplib_amd import surrogate as S  # configuration builders only
    prm = S.draw_params(8, S.SEED_HELDOUT)
    ps = np.stack([O.solve_fp64(S.config(p, M.OPM_PREC_FP32))[0] for p in prm])  # [8][10][21]
    np.savez(os.path.join(HERE, "prices_surrogate.npz"), prices=ps, params=prm,
             strikes=S.STRIKES, nsq=S.NSQ, d=S.D_LATTICE)
    print("surrogate", ps.shape, float(ps.min()), float(ps.max()))


def write_rs_scenarios(n=32):
    """FP64 prices of the 2048-state regime-switching scenario set
    (scenario 0 = the workload, then seeded vol levels / regime rates /
    rates), backward induction over 252 daily steps, 41 strikes"""
    cfgs = M.config_rs2048_scenarios(n)
    ps = np.stack([O.solve_fp64(c) for c in cfgs])  # [n][8][1][41]
    np.savez(os.path.join(HERE, "prices_rs2048_scenarios.npz"), prices=ps,
             sigma=np.array([c.sigma_list() for c in cfgs]), q_off=np.array([c.q_off for c in cfgs]),
             rate=np.array([c.rate for c in cfgs]))
    print("rs2048 scenarios", ps.shape, float(ps.min()), float(ps.max()))


def main():
    if sys.argv[1:] == ["surrogate"]:
        write_surrogate()
        return
    if sys.argv[1:] == ["rs_scenarios"]:
        write_rs_scenarios()
        return
    # MT19937 known answers (C++ [rand.predef]: 10000th output of the
    # default-seeded mt19937 is 4123659995)
    r = O.mt_raw(5489, 10000)
    assert r[0] == 3499211612 and r[9999] == 4123659995
    np.savez(os.path.join(HERE, "mt19937.npz"), s5489=r[:2000], s5490=O.mt_raw(5490, 2000),
             s1=O.mt_raw(1, 2000))

    # expm fixture on a tiny regime lattice
    import scipy.linalg as sl
    c = M.config_small(nx=16, nr=2, nsq=12)
    x, x0, g = O.model_generator(c)
    np.savez(os.path.join(HERE, "expm_small.npz"), L=g, T=c.DeltaT,
             expm=sl.expm(g * c.DeltaT))

    write_prices("lv1d", M.config_lv1d())
    write_prices("rs2048", M.config_rs2048())
    write_prices("rs_monthly", M.config_rs_monthly())
    write_prices("small_bi", M.config_small(nx=64, nr=2))
    cfgs = M.config_sweep()
    ps = np.stack([O.solve_fp64(c)[0] for c in cfgs])  # [64][10][21]
    np.savez(os.path.join(HERE, "prices_sweep.npz"), prices=ps,
             d=np.array([c.nx for c in cfgs]), sigma0=np.array([c.sigma0 for c in cfgs]),
             beta=np.array([c.beta for c in cfgs]), rate=np.array([c.rate for c in cfgs]))
    print("sweep", ps.shape, float(ps.min()), float(ps.max()))
    write_prices("small_prop", M.config_small(nx=64, nr=2, price_kind=M.OPM_PRICE_PROPAGATE,
                                              nsteps=6))
    write_surrogate()
    write_rs_scenarios()


if __name__ == "__main__":
    main()
