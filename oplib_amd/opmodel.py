"""ctypes binding of libopmodel.so (include/opmodel.h) and the named
workload configurations.

``Solver`` wraps one ``opm_solver``: build the lattice, assemble the FP64
generator and first step on the GPU, square, price.  The configuration
builders below fix every choice the workload texts leave open; each one is
stated where it is made.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import ophip  # loads libophip first (libopmodel links it)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopmodel.so")

OPM_VOL_REGIME, OPM_VOL_C1, OPM_VOL_CEV = 0, 1, 2
OPM_PRICE_BACKWARD, OPM_PRICE_PROPAGATE = 0, 1
OPM_PREC_FP32, OPM_PREC_F16, OPM_PREC_F16X3, OPM_PREC_BF16X3, OPM_PREC_F8X3 = 0, 1, 2, 3, 4


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nr", C.c_int),
        ("xlo", C.c_double), ("xhi", C.c_double), ("spot", C.c_double),
        ("rate", C.c_double), ("q_off", C.c_double),
        ("vol_kind", C.c_int),
        ("sigma_r", C.c_double * 16),
        ("sigma0", C.c_double), ("beta", C.c_double), ("skew", C.c_double), ("curv", C.c_double),
        ("DeltaT", C.c_double),
        ("nsq", C.c_int),
        ("price_kind", C.c_int),
        ("nsteps", C.c_int),
        ("nstrikes", C.c_int),
        ("strikes", C.c_double * 64),
        ("precision", C.c_int),
    ]

    @property
    def d(self):
        return self.nx * self.nr

    def strike_list(self):
        return [self.strikes[i] for i in range(self.nstrikes)]

    def sigma_list(self):
        return [self.sigma_r[i] for i in range(self.nr)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libopmodel.so not found at {LIB_PATH}: build it first")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.opm_grid.argtypes = [P(Config), P(C.c_double), P(C.c_int)]
    lib.opm_grid.restype = C.c_int
    lib.opm_params.argtypes = [P(Config), P(C.c_double), P(C.c_double)]
    lib.opm_params.restype = C.c_int
    lib.opm_create.argtypes = [P(Config)]
    lib.opm_create.restype = C.c_void_p
    lib.opm_destroy.argtypes = [C.c_void_p]
    lib.opm_destroy.restype = None
    for n in ("opm_solve", "opm_capture", "opm_launch"):
        getattr(lib, n).argtypes = [C.c_void_p]
        getattr(lib, n).restype = C.c_int
    lib.opm_prices.argtypes = [C.c_void_p, P(C.c_double)]
    lib.opm_prices.restype = C.c_int
    lib.opm_kernel.argtypes = [C.c_void_p, P(C.c_float)]
    lib.opm_kernel.restype = C.c_int
    lib.opm_dims.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
    lib.opm_dims.restype = C.c_int
    return lib


lib = _load()
lib.opm_max_rate.argtypes = [C.POINTER(Config)]
lib.opm_max_rate.restype = C.c_double
lib.opm_plan_nsq.argtypes = [C.POINTER(Config), C.c_double]
lib.opm_plan_nsq.restype = C.c_int


D64 = C.POINTER(C.c_double)
lib.opm_sv_params.argtypes = [C.c_int, C.c_int, D64, D64, C.c_int, C.c_int, D64, D64, D64,
                              D64, D64, D64, D64, D64]
lib.opm_sv_params.restype = C.c_int


def max_rate(cfg):
    return lib.opm_max_rate(C.byref(cfg))


def sv_params(xval, rval, y0, t_years, tau, rho=None):
    """CSVModel.mkgen (CSVModel.cs:103-172): the SV generator's per-state
    inputs (svol, volvol, voldrift, jm, jp), each ni x d (state y = x + nx r)"""
    xv = np.ascontiguousarray(xval, dtype=np.float64)
    rv = np.ascontiguousarray(rval, dtype=np.float64)
    ts = np.ascontiguousarray(t_years, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(8, 4)
    nx, nr, ni = len(xv), len(rv), len(ts)
    rh = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64).reshape(ni, 8)
    out = [np.zeros((ni, nx * nr)) for _ in range(5)]
    ptr = lambda a: a.ctypes.data_as(D64)  # noqa: E731
    st = lib.opm_sv_params(nx, nr, ptr(xv), ptr(rv), int(y0), ni, ptr(ts), ptr(tau),
                           ptr(rh) if rh is not None else None, *[ptr(a) for a in out])
    if st:
        raise ValueError("opm_sv_params: bad argument or non-finite curve value")
    return out


def plan_nsq(cfg, dt0_cap):
    """niter = ceil(log2(DeltaT / min(0.5/max|L_yy|, dt0_cap)))
    (CModel Common.cs:289-307, in years)"""
    return lib.opm_plan_nsq(C.byref(cfg), dt0_cap)


# ---------------------------------------------------------------------------
# workload configurations
# ---------------------------------------------------------------------------

def _base(nx, nr, rate):
    c = Config()
    c.nx, c.nr = nx, nr
    # uniform on [0,400], rescaled so the node nearest 100 lands on 100
    # (SGrid.cs:540-557): for nx = 256 that is node 64 and h = 1.5625
    c.xlo, c.xhi, c.spot = 0.0, 400.0, 100.0
    c.rate = rate
    c.precision = OPM_PREC_FP32
    return c


def config_lv1d():
    """Plain method, 1-D local vol, d = 256.

    sigma(x) = clip(0.25 (x/100)^-0.5, 0.05, 1), r = 0.02; dt0 = 2^-20 yr,
    16 squarings to U(1/16 yr), then the propagation chain U(k/16) for
    k = 1..16 (15 products; U(1/16) itself is the first) and European calls
    K = 80..120 step 2 at every k (T = 1 yr is k = 16)."""
    c = _base(256, 1, 0.02)
    c.vol_kind = OPM_VOL_C1
    c.DeltaT = 1.0 / 16.0
    c.nsq = 16
    c.price_kind = OPM_PRICE_PROPAGATE
    c.nsteps = 16
    ks = [80.0 + 2.0 * i for i in range(21)]
    c.nstrikes = len(ks)
    for i, k in enumerate(ks):
        c.strikes[i] = k
    return c


def config_rs2048(nsteps=252):
    """Plain method at GPU scale: 256 price nodes x 8 regimes, d = 2048.

    sigma_k = 0.1..0.8, regime rate 0.5/yr off-diagonal, r = 0.02 (not
    stated by the workload; the 1-D workload's rate).  16 squarings of the
    FP32 kernel reach the daily step: dt0 = (1/252) 2^-16 yr (<= 2^-20),
    U(1/252).  Backward induction over 252 daily steps (T = 1 yr) for 41
    strikes K = 80..120 step 1; prices at the spot node in every regime."""
    c = _base(256, 8, 0.02)
    c.vol_kind = OPM_VOL_REGIME
    for k in range(8):
        c.sigma_r[k] = 0.1 * (k + 1)
    c.q_off = 0.5
    c.DeltaT = 1.0 / 252.0
    c.nsq = 16
    c.price_kind = OPM_PRICE_BACKWARD
    c.nsteps = nsteps
    ks = [80.0 + i for i in range(41)]
    c.nstrikes = len(ks)
    for i, k in enumerate(ks):
        c.strikes[i] = k
    return c


def config_rs_monthly():
    """Precision study generator (the 2048-state lattice above): 16
    squarings to U(1/12), dt0 = (1/12) 2^-16, then the propagation chain to
    the 12 monthly maturities, 41 strikes."""
    c = config_rs2048()
    c.DeltaT = 1.0 / 12.0
    c.price_kind = OPM_PRICE_PROPAGATE
    c.nsteps = 12
    return c


def config_small(nx=64, nr=2, price_kind=OPM_PRICE_BACKWARD, nsteps=8, nsq=10):
    """Small lattices for parity tests (same family as the 2048-state one)."""
    c = _base(nx, nr, 0.03)
    c.vol_kind = OPM_VOL_REGIME
    for k in range(nr):
        c.sigma_r[k] = 0.15 + 0.1 * k
    c.q_off = 0.5 if nr > 1 else 0.0
    c.DeltaT = 1.0 / 52.0
    c.nsq = nsq
    c.price_kind = price_kind
    c.nsteps = nsteps
    ks = [90.0 + 2.5 * i for i in range(9)]
    c.nstrikes = len(ks)
    for i, k in enumerate(ks):
        c.strikes[i] = k
    return c


# ---------------------------------------------------------------------------
# solver
# ---------------------------------------------------------------------------

class Solver:
    def __init__(self, cfg: Config):
        ophip.init()
        self.cfg = cfg
        self.p = lib.opm_create(C.byref(cfg))
        if not self.p:
            raise ophip.OphipError(f"opm_create failed: 0x{ophip.lib.opcuda_get_status():x}")
        d, nm = C.c_int(), C.c_int()
        lib.opm_dims(self.p, C.byref(d), C.byref(nm))
        self.d, self.n_mat = d.value, nm.value

    def solve(self):
        ophip.check(lib.opm_solve(self.p), "opm_solve")

    def capture(self):
        ophip.check(lib.opm_capture(self.p), "opm_capture")

    def launch(self):
        ophip.check(lib.opm_launch(self.p), "opm_launch")

    def prices(self) -> np.ndarray:
        c = self.cfg
        out = np.zeros(c.nr * self.n_mat * c.nstrikes)
        ophip.check(lib.opm_prices(self.p, out.ctypes.data_as(C.POINTER(C.c_double))), "prices")
        return out.reshape(c.nr, self.n_mat, c.nstrikes)

    def kernel(self) -> np.ndarray:
        U = np.zeros(self.d * self.d, dtype=np.float32)
        ophip.check(lib.opm_kernel(self.p, U.ctypes.data_as(C.POINTER(C.c_float))), "kernel")
        return U.reshape(self.d, self.d, order="F")

    def close(self):
        if self.p:
            lib.opm_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid(cfg: Config):
    x = np.zeros(cfg.nx)
    x0 = C.c_int()
    lib.opm_grid(C.byref(cfg), x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(x0))
    return x, x0.value


def params(cfg: Config):
    n = cfg.nx * cfg.nr
    dr, vo = np.zeros(n), np.zeros(n)
    lib.opm_params(C.byref(cfg), dr.ctypes.data_as(C.POINTER(C.c_double)),
                   vo.ctypes.data_as(C.POINTER(C.c_double)))
    return dr, vo


def flops(cfg: Config) -> float:
    """Algorithmic FLOPs of one solve: 2d^3 per squaring/propagation GEMM,
    2*d^2*nstrikes per backward step (unpadded strikes)."""
    d = cfg.d
    f = cfg.nsq * 2.0 * d ** 3
    if cfg.price_kind == OPM_PRICE_BACKWARD:
        f += cfg.nsteps * 2.0 * d * d * cfg.nstrikes
    else:
        f += (cfg.nsteps - 1) * 2.0 * d ** 3 + cfg.nsteps * cfg.nr * 2.0 * d * cfg.nstrikes
    return f


def df(cfg: Config, t: float) -> float:
    return math.exp(-cfg.rate * t)


# ---------------------------------------------------------------------------
# calibration sweep
# ---------------------------------------------------------------------------

OPM_PRICE_FORWARD = 2
SWEEP_SIZES = (384, 512, 768, 1024, 1536)


def _mt_uniforms(seed, n):
    """n uniforms x/2^32 from MT19937 init_genrand(seed) (numpy's legacy
    RandomState is that generator; its raw 32-bit outputs are the same
    stream as std::mt19937(seed))."""
    rs = np.random.RandomState(seed)
    return rs.randint(0, 2**32, size=n, dtype=np.uint64).astype(np.float64) / 2.0**32


def config_sweep(n=64, seed=5489, precision=OPM_PREC_F16X3):
    """Calibration sweep: n one-regime CEV lattices.

    Per scenario, four MT19937(seed) uniforms in this order: lattice size
    d = SWEEP_SIZES[floor(5u)], sigma0 = 0.1 + 0.5u, beta = 0.3 + 0.7u,
    r = 0.05u.  sigma(x) = sigma0 (x/100)^(beta-1) on [0,400] (spot on a
    node).  16 squarings to U(0.1 yr), dt0 = 0.1 2^-16; 21 strikes
    K = 80..120 step 2 at the 10 maturities 0.1, 0.2, ..., 1.0 yr."""
    u = _mt_uniforms(seed, 4 * n).reshape(n, 4)
    cfgs = []
    for i in range(n):
        c = _base(SWEEP_SIZES[int(u[i, 0] * len(SWEEP_SIZES))], 1, 0.05 * u[i, 3])
        c.vol_kind = OPM_VOL_CEV
        c.sigma0 = 0.1 + 0.5 * u[i, 1]
        c.beta = 0.3 + 0.7 * u[i, 2]
        c.skew = c.curv = 0.0
        c.DeltaT = 0.1
        # the reference planner with the 2^-20 yr cap on dt0 (at least the
        # 16 squarings of the 2048-state workload): 17 or 18 here
        c.nsq = max(16, plan_nsq(c, 2.0**-20))
        c.price_kind = OPM_PRICE_FORWARD
        c.nsteps = 10
        ks = [80.0 + 2.0 * j for j in range(21)]
        c.nstrikes = len(ks)
        for j, k in enumerate(ks):
            c.strikes[j] = k
        c.precision = precision
        cfgs.append(c)
    return cfgs


lib.opm_batch_create.argtypes = [C.POINTER(Config), C.c_int, C.c_int]
lib.opm_batch_create.restype = C.c_void_p
lib.opm_batch_destroy.argtypes = [C.c_void_p]
lib.opm_batch_destroy.restype = None
for _n in ("opm_batch_solve", "opm_batch_capture", "opm_batch_launch"):
    getattr(lib, _n).argtypes = [C.c_void_p]
    getattr(lib, _n).restype = C.c_int
lib.opm_batch_prices.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
lib.opm_batch_prices.restype = C.c_int


class Batch:
    """Many one-regime lattices solved together (opm_batch)."""

    def __init__(self, cfgs, tile=0):
        ophip.init()
        self.cfgs = list(cfgs)
        arr = (Config * len(self.cfgs))(*self.cfgs)
        self.p = lib.opm_batch_create(arr, len(self.cfgs), tile)
        if not self.p:
            raise ophip.OphipError(
                f"opm_batch_create failed: 0x{ophip.lib.opcuda_get_status():x}")

    def solve(self):
        ophip.check(lib.opm_batch_solve(self.p), "opm_batch_solve")

    def capture(self):
        ophip.check(lib.opm_batch_capture(self.p), "opm_batch_capture")

    def launch(self):
        ophip.check(lib.opm_batch_launch(self.p), "opm_batch_launch")

    def prices(self):
        c = self.cfgs[0]
        out = np.zeros(len(self.cfgs) * c.nsteps * c.nstrikes)
        ophip.check(lib.opm_batch_prices(self.p, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "prices")
        return out.reshape(len(self.cfgs), c.nsteps, c.nstrikes)

    def close(self):
        if self.p:
            lib.opm_batch_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sweep_flops(cfgs):
    """algorithmic FLOPs of a sweep: 2d^3 per squaring, 2d^2 per forward
    step, 2d per strike price"""
    f = 0.0
    for c in cfgs:
        d = c.d
        f += c.nsq * 2.0 * d ** 3 + c.nsteps * 2.0 * d * d + c.nsteps * c.nstrikes * 2.0 * d
    return f


# ---------------------------------------------------------------------------
# groups of scenarios of one lattice size (regime-switching family)
# ---------------------------------------------------------------------------

def config_rs2048_scenarios(n=8, seed=5489, precision=OPM_PREC_F16X3, nsteps=252):
    """n scenarios of the 2048-state regime-switching lattice: scenario 0 is
    the workload itself (sigma_k = 0.1..0.8, regime rate 0.5/yr, r = 0.02);
    scenario i >= 1 draws three MT19937(seed) uniforms in this order (after
    the 3(i-1) before it): a vol level s = 0.8 + 0.4u scaling every sigma_k,
    a regime rate 0.25 + 0.75u and r = 0.05u.  Same grid, dt0, 16 squarings
    to U(1/252), 252 backward steps and 41 strikes."""
    u = _mt_uniforms(seed, 3 * max(n - 1, 0)).reshape(-1, 3)
    cfgs = []
    for i in range(n):
        c = config_rs2048(nsteps)
        if i > 0:
            lv, q, r = u[i - 1]
            for k in range(8):
                c.sigma_r[k] = (0.8 + 0.4 * lv) * 0.1 * (k + 1)
            c.q_off = 0.25 + 0.75 * q
            c.rate = 0.05 * r
        c.precision = precision
        cfgs.append(c)
    return cfgs


lib.opm_group_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
lib.opm_group_create.restype = C.c_void_p
lib.opm_group_destroy.argtypes = [C.c_void_p]
lib.opm_group_destroy.restype = None
for _n in ("opm_group_solve", "opm_group_capture", "opm_group_launch"):
    getattr(lib, _n).argtypes = [C.c_void_p]
    getattr(lib, _n).restype = C.c_int


class Group:
    """Scenarios of one lattice size priced together (opm_group): one
    squaring launch per round for all, one backward-step launch per step
    for all (FP16 x3).  prices() -> [n][nr][1][nstrikes]."""

    def __init__(self, cfgs):
        ophip.init()
        self.solvers = [Solver(c) for c in cfgs]
        arr = (C.c_void_p * len(self.solvers))(*[s.p for s in self.solvers])
        self.p = lib.opm_group_create(arr, len(self.solvers))
        if not self.p:
            err = ophip.lib.opcuda_get_status()  # before the solvers' cleanup resets it
            for s in self.solvers:
                s.close()
            raise ophip.OphipError(f"opm_group_create failed: 0x{err:x}")

    def solve(self):
        ophip.check(lib.opm_group_solve(self.p), "opm_group_solve")

    def capture(self):
        ophip.check(lib.opm_group_capture(self.p), "opm_group_capture")

    def launch(self):
        ophip.check(lib.opm_group_launch(self.p), "opm_group_launch")

    def prices(self):
        return np.stack([s.prices() for s in self.solvers])

    def close(self):
        if self.p:
            lib.opm_group_destroy(self.p)
            self.p = None
        for s in self.solvers:
            s.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# the reference's step plan (opm_mcplan_*)
# ---------------------------------------------------------------------------

_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
lib.opm_mcplan_create.argtypes = [C.c_int, _IP, C.c_int, _IP, C.c_int, _DP]
lib.opm_mcplan_create.restype = C.c_void_p
lib.opm_mcplan_destroy.argtypes = [C.c_void_p]
lib.opm_mcplan_destroy.restype = None
lib.opm_mcplan_sizes.argtypes = [C.c_void_p] + [_IP] * 5
lib.opm_mcplan_sizes.restype = C.c_int
lib.opm_mcplan_arrays.argtypes = [C.c_void_p, _IP, _IP, _IP, _IP, _IP, _DP, _IP, _IP, _IP]
lib.opm_mcplan_arrays.restype = C.c_int
lib.opm_mcplan_exec.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
lib.opm_mcplan_exec.restype = C.c_int


class McPlan:
    """CModel.make_mc_plan in days (host logic only; exec needs the GPU)."""

    def __init__(self, t_k, t_i, dt0_days, maxdiag):
        tk = np.ascontiguousarray(t_k, dtype=np.int32)
        ti = np.ascontiguousarray(t_i, dtype=np.int32)
        md = np.ascontiguousarray(maxdiag, dtype=np.float64)
        self.p = lib.opm_mcplan_create(len(tk), tk.ctypes.data_as(_IP), len(ti), ti.ctypes.data_as(_IP),
                                       int(dt0_days), md.ctypes.data_as(_DP))
        if not self.p:
            raise ValueError("invalid step or generator dates")
        v = [C.c_int() for _ in range(5)]
        lib.opm_mcplan_sizes(self.p, *[C.byref(x) for x in v])
        self.nq, self.nj, self.nm, self.nr, self.ns = [x.value for x in v]
        a = {k: np.zeros(self.nq, np.int32) for k in ("t_q", "i_q", "j_q")}
        a.update({k: np.zeros(self.nj, np.int32) for k in ("i_j", "niter_j")})
        a["dt_j"] = np.zeros(self.nj)
        a["m_k"] = np.zeros(len(tk), np.int32)
        a["chain"] = np.zeros(max(self.nm * self.ns, 1), np.int32)
        a["chain_len"] = np.zeros(self.nm, np.int32)
        order = ("t_q", "i_q", "j_q", "i_j", "niter_j", "dt_j", "m_k", "chain", "chain_len")
        lib.opm_mcplan_arrays(self.p, *[a[k].ctypes.data_as(_DP if k == "dt_j" else _IP) for k in order])
        self.__dict__.update(a)
        self.chains = [list(a["chain"][m * self.ns:m * self.ns + a["chain_len"][m]]) for m in range(self.nm)]

    def exec(self, d, gen, xtpk, xbuf, tpk):
        ophip.check(lib.opm_mcplan_exec(self.p, d, gen, xtpk, xbuf, tpk), "opm_mcplan_exec")

    def close(self):
        if self.p:
            lib.opm_mcplan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
