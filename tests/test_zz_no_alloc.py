"""GPU test of ophip_mc_payoff_noalloc, the payoff entry point that
allocates nothing (its partial sums live in a fixed device array of the
library): called through the C-ABI, its results are the bits of
ophip_mc_payoff on the same inputs, on two calls in a row, the second with
a smaller input; a call beyond the array's size is refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_noalloc_payoff_matches_payoff(hip):
    H = hip
    fn = H.lib.ophip_mc_payoff_noalloc
    rng = np.random.default_rng(41)
    d = 300
    xval = H.DeviceArray.from_numpy(np.linspace(0.0, 400.0, d).astype(np.float32))
    ks = np.array([80.0 + 2.5 * i for i in range(17)], dtype=np.float32)
    ip, dp, fp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
    # (paths, steps held, listed steps): 2^20 paths at 4 listed steps, then
    # a smaller call; the second call leaves the first call's larger
    # partials in the array
    for nscen, nk, steps in (((1 << 20) + 1000, 6, [5, 0, 3, 2]), (70000, 3, [2])):
        banner(f"noalloc payoff nscen={nscen} nk={nk} steps={steps}")
        ysk = H.DeviceArray.from_numpy(rng.integers(-3, d + 3, nscen * nk).astype(np.int16))
        st = np.array(steps, dtype=np.int32)
        df = np.exp(-0.02 * (st + 1) / 12.0)
        outs = []
        for f in (H.lib.ophip_mc_payoff, fn):
            out = H.DeviceArray.from_numpy(np.full(2 * len(st) * len(ks), np.nan))
            H.check(f(ysk.h, nscen, nk, len(st), st.ctypes.data_as(ip), df.ctypes.data_as(dp), xval.h, d,
                      ks.ctypes.data_as(fp), len(ks), out.h))
            H.sync()
            outs.append(out.download())
        assert np.all(np.isfinite(outs[0]))
        assert np.array_equal(outs[0], outs[1])
    # 2^20 paths are 256 chunks: 16 listed steps fit the array, 17 do not
    nscen, nk = 1 << 20, 17
    ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
    out = H.DeviceArray.empty(2 * nk * len(ks), dtype=np.float64)
    st = np.arange(nk, dtype=np.int32)
    df = np.ones(nk)
    bad = fn(ysk.h, nscen, nk, nk, st.ctypes.data_as(ip), df.ctypes.data_as(dp), xval.h, d,
             ks.ctypes.data_as(fp), len(ks), out.h)
    assert bad == H.lib.opcuda_get_status() == 0x10003
