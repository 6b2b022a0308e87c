"""Learned surrogate with operator-solver fallback.

A five-member ensemble of MLPs 5 -> 256 -> 256 -> 256 -> 210 maps the
local-vol parameters (sigma0, beta, r, a, b) of

    sigma(x) = sigma0 (x/100)^(beta-1) (1 + a ln(x/100) + b ln^2(x/100))

to the 10 x 21 grid of European call prices (maturities 0.1 .. 1.0 yr,
strikes 80 .. 120 step 2, spot 100).  The networks learn the standardized
relative residual of the operator solver's prices over a Black-Scholes
control variate (the short-maturity implied vol of Berestycki-Busca-
Florent), relative to a smooth stand-in for the error metric's
denominator evaluated on the control variate, so that the training loss
weighs errors as the metric does; sigma0 enters on a log scale, and each
input reaches the first layer as a BF16 pair (value and remainder).  The
training draws cover the box widened by 5% on each side.  The control
variate, the ensemble mean and the spread are evaluated on the GPU
(ophip_surr_eval).  Any draw whose relative ensemble spread exceeds the
calibrated threshold is priced by the operator solver instead.

Training labels come from the operator solver itself: the 1024-node
lattice, U = I + W squared 20 times in FP16 with hi/lo split products
(the mixed-precision method, about 1e-6 from FP64), then the forward chain
to the 10 maturities.  All compute is in libophip (forward, backward and
Adam kernels); this module only draws parameters, standardizes targets and
sequences the calls.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import ophip
from . import opmodel as M

PARAM_NAMES = ("sigma0", "beta", "r", "a", "b")
PARAM_LO = np.array([0.1, 0.3, 0.0, -0.5, 0.0])
PARAM_HI = np.array([0.6, 1.0, 0.05, 0.5, 0.5])
STRIKES = np.array([80.0 + 2.0 * j for j in range(21)])
NMAT = 10
DT = 0.1
SPOT = 100.0
D_LATTICE = 1024
# one squaring count for the whole box (the solver's prices are then a
# smooth function of the parameters): the reference planner's count,
# dt0 = min(0.5 / max|L_yy|, 2^-20 yr), at the corner with the largest
# rate (sigma0 0.6, beta 1, a 0.5, b 0.5)
NSQ = 20
NOUT = NMAT * len(STRIKES)
TOL = 1e-4

SEED_TRAIN, SEED_HELDOUT, SEED_INFER, SEED_CAL = 1001, 2002, 3003, 4004


# training draws cover the box widened by this fraction of its width on
# each side (sigma0 on its log scale), so that the box's corners lie inside
# the training distribution
TRAIN_MARGIN = 0.05


def box(margin=0.0):
    """the parameter box, widened by `margin` of its width on each side"""
    lo = PARAM_LO - margin * (PARAM_HI - PARAM_LO)
    hi = PARAM_HI + margin * (PARAM_HI - PARAM_LO)
    ratio = PARAM_HI[0] / PARAM_LO[0]
    lo[0], hi[0] = PARAM_LO[0] * ratio ** -margin, PARAM_HI[0] * ratio ** margin
    return lo, hi


def draw_params(n, seed, margin=0.0):
    """n draws, five MT19937(seed) uniforms x/2^32 each, in the order
    sigma0, beta, r, a, b, mapped onto the box (widened by `margin`)"""
    u = M._mt_uniforms(seed, 5 * n).reshape(n, 5)
    lo, hi = box(margin)
    return lo + (hi - lo) * u


def config(p, precision=M.OPM_PREC_F16X3):
    """the solver configuration of one draw"""
    c = M._base(D_LATTICE, 1, float(p[2]))
    c.vol_kind = M.OPM_VOL_CEV
    c.sigma0, c.beta, c.skew, c.curv = float(p[0]), float(p[1]), float(p[3]), float(p[4])
    c.DeltaT = DT
    c.nsq = NSQ
    c.price_kind = M.OPM_PRICE_FORWARD
    c.nsteps = NMAT
    c.nstrikes = len(STRIKES)
    for j, k in enumerate(STRIKES):
        c.strikes[j] = float(k)
    c.precision = precision
    return c


M.lib.opm_batch_update.argtypes = [C.c_void_p, C.POINTER(M.Config)]
M.lib.opm_batch_update.restype = C.c_int


class PriceSolver:
    """The operator solver over batches of nb draws: one captured graph,
    re-parameterized per batch (opm_batch_update)."""

    def __init__(self, nb=64, precision=M.OPM_PREC_F16X3, tile=0):
        """tile 0: the squaring plan the batch picks on timed solves (its
        K-cell cutoff leaves out products per tile, so the bits depend on
        the plan); 64 or 128: that tile for every lattice"""
        self.nb = nb
        mid = 0.5 * (PARAM_LO + PARAM_HI)
        self.batch = M.Batch([config(mid, precision) for _ in range(nb)], tile=tile)
        self.precision = precision
        self.batch.capture()

    def solve(self, params):
        """prices [n][210] (maturity-major) of the draws params [n][5]"""
        params = np.asarray(params, dtype=np.float64)
        out = np.empty((len(params), NOUT))
        for i in range(0, len(params), self.nb):
            chunk = params[i:i + self.nb]
            padded = np.concatenate([chunk, np.repeat(chunk[-1:], self.nb - len(chunk), 0)])
            arr = (M.Config * self.nb)(*[config(p, self.precision) for p in padded])
            ophip.check(M.lib.opm_batch_update(self.batch.p, arr), "opm_batch_update")
            self.batch.launch()
            out[i:i + len(chunk)] = self.batch.prices()[:len(chunk)].reshape(len(chunk), NOUT)
        return out

    def close(self):
        self.batch.close()


def normalize(params):
    """the box mapped onto [-1, 1], sigma0 on a log scale (the same map as
    the evaluation kernel)"""
    p = np.asarray(params, dtype=np.float64)
    x = 2.0 * (p - PARAM_LO) / (PARAM_HI - PARAM_LO) - 1.0
    x[:, 0] = 2.0 * np.log(p[:, 0] / PARAM_LO[0]) / np.log(PARAM_HI[0] / PARAM_LO[0]) - 1.0
    return x.astype(np.float32)


def residual_scale(base):
    """per-output residual scale: sqrt(base^2 + (1e-3 max|base|)^2) per
    draw, a smooth stand-in for the error metric's denominator
    max(|base|, 1e-3 max|base|) on the control variate (within sqrt(2) of
    it; the kink of the max is a feature the network cannot follow)"""
    base = np.asarray(base, dtype=np.float64)
    fl = 1e-3 * np.abs(base).max(axis=1, keepdims=True)
    return np.sqrt(base * base + fl * fl)


def control_variate(params):
    """the Black-Scholes control variate [n][210], computed on the GPU"""
    params = np.asarray(params, dtype=np.float32)
    n = len(params)
    prm = ophip.DeviceArray.from_numpy(params, order="C")
    base = ophip.DeviceArray.empty(n * NOUT)
    ks = (C.c_float * len(STRIKES))(*STRIKES)
    ophip.check(ophip.lib.ophip_surr_base(n, prm.h, SPOT, DT, NMAT, ks, len(STRIKES), base.h),
                "ophip_surr_base")
    return base.download().reshape(n, NOUT)


class Surrogate:
    """Ensemble of `members` MLPs trained together on the GPU."""

    def __init__(self, members=5, cap=4096, seed=7):
        ophip.init()
        self.E, self.cap = members, cap
        self.net = ophip.lib.ophip_mlp_create(members, cap)
        if not self.net:
            raise ophip.OphipError(f"ophip_mlp_create failed: 0x{ophip.lib.opcuda_get_status():x}")
        self.P = ophip.lib.ophip_mlp_param_count()
        for e in range(members):
            self.set_weights(e, self.init_weights(seed + e))
        self.mu = np.zeros(256, np.float32)
        self.sd = np.ones(256, np.float32)
        self._upload_stats()

    # ---- parameter layout (floats): W1[256][32] b1 W2[256][256] b2 W3 b3 W4[256][256] b4 ----
    LAYERS = ((256, 32, 5, 256), (256, 256, 256, 256), (256, 256, 256, 256), (256, 256, 256, 210))

    def unpack(self, flat):
        """[(W [out][in] live part, b live part)] from a parameter array"""
        out, o = [], 0
        for (po, pi, li, lo) in self.LAYERS:
            W = flat[o:o + po * pi].reshape(po, pi)[:lo, :li]
            o += po * pi
            b = flat[o:o + po][:lo]
            o += po
            out.append((W, b))
        return out

    def pack(self, layers):
        flat = np.zeros(self.P, np.float32)
        o = 0
        for (po, pi, li, lo), (W, b) in zip(self.LAYERS, layers):
            blk = np.zeros((po, pi), np.float32)
            blk[:lo, :li] = W
            flat[o:o + po * pi] = blk.reshape(-1)
            o += po * pi
            flat[o:o + lo] = b
            o += po
        return flat

    def init_weights(self, seed):
        """Glorot-uniform weights, zero biases"""
        rng = np.random.RandomState(seed)
        layers = []
        for (_, _, li, lo) in self.LAYERS:
            a = np.sqrt(6.0 / (li + lo))
            layers.append((rng.uniform(-a, a, size=(lo, li)).astype(np.float32), np.zeros(lo, np.float32)))
        return self.pack(layers)

    def _xfer(self, member, which, arr, to_device):
        ptr = arr.ctypes.data_as(C.POINTER(C.c_float))
        ophip.check(ophip.lib.ophip_mlp_params(self.net, member, which, ptr, 1 if to_device else 0),
                    "ophip_mlp_params")

    def set_weights(self, member, flat):
        self._xfer(member, 0, np.ascontiguousarray(flat, dtype=np.float32), True)

    def get(self, member, which=0):
        a = np.empty(self.P, np.float32)
        self._xfer(member, which, a, False)
        return a

    def _upload_stats(self):
        self._xfer(0, 4, np.ascontiguousarray(self.mu, np.float32), True)
        self._xfer(0, 5, np.ascontiguousarray(self.sd, np.float32), True)

    # ---- training ----
    def fit(self, params, prices, steps, batch=4096, lr=2e-3, lr_end=2e-5, seed=11, log=None,
            log_every=500):
        """Adam on the mean squared standardized residual; each member draws
        its own minibatches (per-epoch permutations); cosine learning-rate
        decay from lr to lr_end.  Returns the per-member loss history."""
        params = np.asarray(params, np.float64)
        n = len(params)
        if n < batch or batch % 64 or batch > self.cap:
            raise ValueError("need batch % 64 == 0, batch <= capacity <= samples")
        base = control_variate(params).astype(np.float64)
        res = (np.asarray(prices, np.float64) - base) / residual_scale(base)
        mu = res.mean(0)
        sd = res.std(0)
        sd = np.maximum(sd, 1e-6 * max(sd.max(), 1e-30))
        self.mu[:] = 0.0
        self.sd[:] = 1.0
        self.mu[:NOUT] = mu
        self.sd[:NOUT] = sd
        self._upload_stats()
        T = ophip.DeviceArray.from_numpy(((res - mu) / sd).astype(np.float32), order="C")
        X = ophip.DeviceArray.from_numpy(normalize(params), order="C")
        per_epoch = n // batch
        idx = ophip.DeviceArray.empty(per_epoch * self.E * batch, dtype=np.int32)
        rng = np.random.RandomState(seed)
        hist = []
        t0 = time.perf_counter()
        lossbuf = np.zeros(self.E)
        for s in range(steps):
            k = s % per_epoch
            if k == 0:
                perms = np.stack([rng.permutation(n)[:per_epoch * batch] for _ in range(self.E)])
                # [step][member][row]
                perms = perms.reshape(self.E, per_epoch, batch).transpose(1, 0, 2)
                idx.upload(np.ascontiguousarray(perms, dtype=np.int32))
            ophip.check(ophip.lib.ophip_mlp_backward(self.net, batch, idx.h + k * self.E * batch * 4,
                                                     X.h, T.h, 0), "ophip_mlp_backward")
            lr_s = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * s / max(steps - 1, 1)))
            ophip.check(ophip.lib.ophip_mlp_adam(self.net, lr_s, 0.9, 0.999, 1e-8), "ophip_mlp_adam")
            if (s + 1) % log_every == 0 or s == steps - 1:
                ophip.check(ophip.lib.ophip_mlp_loss(
                    self.net, batch, lossbuf.ctypes.data_as(C.POINTER(C.c_double))), "loss")
                mse = lossbuf / (batch * NOUT)
                hist.append((s + 1, mse.copy()))
                if log:
                    log(f"[surrogate] step {s + 1}/{steps} lr {lr_s:.2e} mse {mse.mean():.3e} "
                        f"({time.perf_counter() - t0:.1f} s)")
        ophip.sync()
        self.train_seconds = time.perf_counter() - t0
        return hist

    # ---- evaluation ----
    def evaluate(self, params, tol=TOL, prm_dev=None):
        """prices [n][210], relative spread [n] and fallback flags [n] of the
        draws (n % 64 == 0), in chunks of the batch capacity"""
        params = np.asarray(params, np.float32)
        n = len(params)
        if n % 64:
            raise ValueError("draw count must be a multiple of 64")
        prm = prm_dev if prm_dev is not None else ophip.DeviceArray.from_numpy(params, order="C")
        pr = ophip.DeviceArray.empty(n * NOUT)
        sp = ophip.DeviceArray.empty(n)
        fl = ophip.DeviceArray.empty(n, dtype=np.int32)
        self.eval_into(n, prm, pr, sp, fl, tol)
        return (pr.download().reshape(n, NOUT), sp.download(), fl.download())

    def eval_into(self, n, prm, pr, sp, fl, tol=TOL):
        """evaluation on device buffers (no host traffic): the timed path"""
        lo = (C.c_float * 5)(*PARAM_LO)
        hi = (C.c_float * 5)(*PARAM_HI)
        ks = (C.c_float * len(STRIKES))(*STRIKES)
        for i in range(0, n, self.cap):
            b = min(self.cap, n - i)
            ophip.check(ophip.lib.ophip_surr_eval(
                self.net, b, prm.h + i * 5 * 4, lo, hi, SPOT, DT, NMAT, ks, len(STRIKES), tol,
                pr.h + i * NOUT * 4, sp.h + i * 4, fl.h + i * 4), "ophip_surr_eval")

    def calibrate(self, params, prices, tol=TOL):
        """the spread threshold s* for error tolerance tol: the largest
        spread such that every calibration draw with spread <= s* is within
        tol of the solver (0 when none is).  The fallback indicator is then
        spread * tol / s*, compared with tol."""
        n = len(params) // 64 * 64
        pr, sp, _ = self.evaluate(np.asarray(params)[:n], tol)
        e = rel_err(pr, np.asarray(prices)[:n])
        o = np.argsort(sp)
        bad = np.nonzero(e[o] > tol)[0]
        first = bad[0] if len(bad) else n
        return float(sp[o][first - 1]) if first > 0 else 0.0

    def predict(self, params, solver, tol=TOL):
        """surrogate with fallback: (prices [n][210], served mask)"""
        pr, sp, fl = self.evaluate(params, tol)
        bad = np.nonzero(fl)[0]
        out = pr.astype(np.float64)
        if len(bad):
            out[bad] = solver.solve(np.asarray(params)[bad])
        return out, fl == 0, sp

    def close(self):
        if self.net:
            ophip.lib.ophip_mlp_free(self.net)
            self.net = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_err(p, ref):
    """per-draw max relative error with the denominator max(|ref|, 1e-3 max|ref|)"""
    p, ref = np.atleast_2d(p), np.atleast_2d(ref)
    den = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max(axis=1, keepdims=True))
    return (np.abs(p - ref) / den).max(axis=1)
