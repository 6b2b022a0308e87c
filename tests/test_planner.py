"""The step plan of CModel.make_mc_plan (CModel Common.cs:163-644) in the
host model layer (opm_mcplan_*), on dates worked out by hand from the
reference's rules:

* merged dates t_q: every step and generator date up to the last step;
* i_q: the first generator whose date is >= t_q;
* kernels j: distinct (generator, interval length) pairs, first
  occurrence first (todo_q, Common.cs:214-264);
* niter_j = ceil(log2(DeltaT / dt)), dt = min(0.5 / maxdiag, dt0), dates in
  days / 365.25 (Common.cs:290-308), dt_j = DeltaT 2^-niter;
* step kernels m: a step inside one generator period shares the kernel of
  an earlier step of the same length in that period (todo_k,
  Common.cs:318-415); its chain is the kernels of its intervals in date
  order.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def M():
    from oplib_amd import opmodel
    return opmodel


def test_weekly_steps_monthly_generators(M):
    # steps every 7 days to day 63; generators change on days 30, 60, 90
    p = M.McPlan([7, 14, 21, 28, 35, 42, 49, 56, 63], [30, 60, 90], 1, [100.0, 200.0, 50.0])
    assert (p.nq, p.nj, p.nm, p.nr, p.ns) == (11, 6, 4, 3, 2)
    assert list(p.t_q) == [7, 14, 21, 28, 30, 35, 42, 49, 56, 60, 63]
    assert list(p.i_q) == [0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2]
    # lengths 7,7,7,7,2 | 5,7,7,7,4 | 3 -> (0,7) (0,2) (1,5) (1,7) (1,4) (2,3)
    assert list(p.j_q) == [0, 0, 0, 0, 1, 2, 3, 3, 3, 4, 5]
    assert list(p.i_j) == [0, 0, 1, 1, 1, 2]
    # dt0 = 1 day; generator 1 caps dt at 0.5/200 yr = 0.913 days:
    # 7/1 -> 3, 2/1 -> 1, 5/0.913 -> 3, 7/0.913 -> 3, 4/0.913 -> 3, 3/1 -> 2
    assert list(p.niter_j) == [3, 1, 3, 3, 3, 2]
    days = np.array([7, 2, 5, 7, 4, 3]) / 365.25
    assert np.allclose(p.dt_j, days * 2.0 ** -p.niter_j.astype(float), rtol=0, atol=1e-18)
    # weeks 1-4 share one kernel, week 5 spans day 30, weeks 6-8 share one,
    # week 9 spans day 60
    assert list(p.m_k) == [0, 0, 0, 0, 1, 2, 2, 2, 3]
    assert p.chains == [[0], [1, 2], [3], [4, 5]]
    p.close()


def test_one_step_across_generators(M):
    p = M.McPlan([30], [10, 20, 30], 5, [10.0, 10.0, 10.0])
    assert (p.nq, p.nj, p.nm, p.ns) == (3, 3, 1, 3)
    assert list(p.i_q) == [0, 1, 2] and list(p.j_q) == [0, 1, 2]
    assert p.chains == [[0, 1, 2]]
    # 10-day intervals from a 5-day cap (0.5/10 yr is longer): one squaring each
    assert list(p.niter_j) == [1, 1, 1]
    p.close()


def test_first_step_length_dedup(M):
    # the first interval (today -> day 10) and later 10-day intervals of the
    # same generator share a kernel; a 10-day step after a 5-day one does too
    p = M.McPlan([10, 15, 25, 35], [40], 20, [1.0])
    assert list(p.t_q) == [10, 15, 25, 35]
    assert list(p.j_q) == [0, 1, 0, 0]
    assert list(p.m_k) == [0, 1, 0, 0]
    p.close()


@pytest.mark.parametrize("t_k,t_i", [
    ([7, 7], [30]),          # steps not increasing
    ([0, 7], [30]),          # a step on today
    ([7, 40], [30]),         # past the last generator date
    ([7], [30, 20]),         # generator dates not increasing
])
def test_invalid_dates(M, t_k, t_i):
    from oplib_amd import ophip
    ophip.lib.ophip_set_status(0)
    with pytest.raises(ValueError):
        M.McPlan(t_k, t_i, 1, [1.0] * len(t_i))
    assert ophip.lib.opcuda_get_status() == 0x10003  # OPHIP_ERR_BAD_ARG, kept for the caller
