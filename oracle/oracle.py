"""ctypes binding of the CPU oracle (liboracle.so) plus the FP64 scenario
solve composed from it.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU path,
never by the product (oplib_amd/).

The scenario solve restates the reference's CPU (FP64) path: grid and
stencil (SGrid.cs:492-593, S1DGrid.cs:78-147), generator
(CLVModel.cs:86-136), first step (CModel Host Double.cs:89-121), repeated
squaring with DGEMM (CModel Host Double.cs:132-172, linalg.c:93-101), the
propagation chain (CModel Host Double.cs:200-305) and pricing by payoff
GEMV with discounting (OPPricer.cs:143-146) or backward induction.
Configuration objects are read by attribute (nx, nr, xlo, ...), the same
fields as include/opmodel.h's opm_config.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_sp = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: make -C oracle")
    L = C.CDLL(LIB_PATH)
    I, D, F, U = C.c_int, C.c_double, C.c_float, C.c_uint32
    sig = {
        "or_xgrid_uniform": (None, [I, D, D, D, _dp, C.POINTER(I)]),
        "or_stencil_invm": (None, [I, _dp, _dp]),
        "or_lv_generator": (None, [I, _dp, _dp, _dp, _dp]),
        "or_rs_generator": (None, [I, I, _dp, _dp, _dp, D, D, _dp]),
        "or_sv_generator": (None, [I, I] + [_dp] * 9),
        "or_lv_params_c1": (None, [I, _dp, D, _dp, _dp]),
        "or_lv_params_cev": (None, [I, _dp, D, D, D, D, D, _dp, _dp]),
        "or_plan_niter": (I, [D, D, D, C.POINTER(D)]),
        "or_max_abs_diag": (D, [I, _dp]),
        "or_first_step": (None, [I, _dp, D, _dp]),
        "or_sek_f32": (None, [I, _fp, F, _fp]),
        "or_sekd_f32": (None, [I, _fp, F, _fp, C.c_void_p, _fp]),
        "or_sglvg_f32": (None, [I, I, _fp, _fp, _fp]),
        "or_dgemm": (None, [I, I, I, _dp, I, _dp, I, _dp, I]),
        "or_sgemm_f64acc": (None, [I, I, I, F, _fp, I, _fp, I, F, _fp, I]),
        "or_square_chain": (None, [I, I, _dp, _dp]),
        "or_row_renorm": (None, [I, _dp]),
        "or_set_threads": (I, [I]),
        "or_sgemv_gather": (None, [I, _fp, _fp, I, _ip, _ip, I]),
        "or_backward_induction": (None, [I, _dp, I, D, I, _dp]),
        "or_sck_f32": (None, [I, I, _fp, _fp]),
        "or_sck_rows_f32": (None, [I, I, _fp, _fp]),
        "or_sck_rows2_f32": (None, [I, I, I, I, _fp, _fp]),
        "or_mt19937_init": (None, [U, _up]),
        "or_mt19937_next": (U, [_up]),
        "or_tgfsr_run": (None, [I, I, I, U, U, I, I, _up, _up, _up, _up, I, _up]),
        "or_tgfsr_seed": (None, [I, I, _up, _up]),
        "or_mc_hash_build": (None, [I, I, _dp, _dp, _ip]),
        "or_mc_hash_run": (None, [I, _dp, _ip, I, I, I, I, _ip, U, _sp]),
        "or_u01_from_bits": (F, [U]),
        "or_mt_uniforms": (None, [U, I, _fp]),
        "or_sample_linear": (I, [_fp, I, F]),
        "or_sample_colmajor": (I, [_fp, I, I, F]),
        "or_sample_ref_bisect": (I, [_fp, I, I, F]),
        "or_mc_chain": (None, [I, U, I, I, I, I, _fp, _ip, I, _sp]),
        "or_bs_call": (D, [D, D, D, D, D]),
        "or_ncdf": (D, [D]),
        "or_gelu": (D, [D]),
        "or_gelu_grad": (D, [D]),
        "or_adam_step": (None, [I, _dp, _dp, _dp, _dp, D, D, D, D, I]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    return L


lib = _load()


def set_threads(n: int) -> int:
    return lib.or_set_threads(int(n))


# ---------------------------------------------------------------------------
# building blocks
# ---------------------------------------------------------------------------

def xgrid(nx, xlo, xhi, spot):
    x = np.zeros(nx)
    x0 = C.c_int()
    lib.or_xgrid_uniform(nx, xlo, xhi, spot, x, C.byref(x0))
    return x, x0.value


def invm(x):
    out = np.zeros(4 * len(x))
    lib.or_stencil_invm(len(x), np.ascontiguousarray(x), out)
    return out


def lv_generator(x, drift, vol):
    d = len(x)
    g = np.zeros(d * d)
    lib.or_lv_generator(d, invm(x), np.ascontiguousarray(drift, dtype=np.float64),
                        np.ascontiguousarray(vol, dtype=np.float64), g)
    return g.reshape(d, d, order="F")


def rs_generator(x, sigma_r, rate, q_off):
    nx, nr = len(x), len(sigma_r)
    d = nx * nr
    g = np.zeros(d * d)
    lib.or_rs_generator(nx, nr, invm(x), np.ascontiguousarray(x),
                        np.ascontiguousarray(sigma_r, dtype=np.float64), rate, q_off, g)
    return g.reshape(d, d, order="F")


def sv_generator(nx, nr, xval, rval, sdrift, svol, volvol, voldrift, jm, jp):
    """FP64 stochastic-volatility generator, d x d with d = nx*nr; every
    argument a per-state array of d values (state y = x + nx*r)"""
    d = nx * nr
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (xval, rval, sdrift, svol, volvol, voldrift, jm, jp)]
    for a in arrs:
        assert a.shape == (d,)
    g = np.zeros(d * d)
    lib.or_sv_generator(nx, nr, *arrs, g)
    return g.reshape(d, d, order="F")


def sv_params(xval, rval, y0, t_years, tau, rho=None):
    """CSVModel.mkgen (CSVModel.cs:103-172) in numpy: per date i and state
    y = x + nx r the SV generator inputs (svol, volvol, voldrift, jm, jp),
    each ni x d, from eight curves tau[c] = {A, B, c2, c3} and the dates'
    offsets rho[i][c]"""
    xval, rval = np.asarray(xval, float), np.asarray(rval, float)
    tau = np.asarray(tau, float).reshape(8, 4)
    nx, nr = len(xval), len(rval)
    xs, rs = np.tile(xval, nr), np.repeat(rval, nx)
    sx, smin, smax = np.sqrt(xs), math.sqrt(xs[0]), math.sqrt(xs[nx - 1])
    out = [[], [], [], [], []]
    for i, t in enumerate(float(q) for q in t_years):
        v = []
        for c in range(8):
            a, b, c2, c3 = (float(q) for q in tau[c])
            xi = a if c2 == 0 else a + (b - a) * (1 + c3 * t / c2) * math.exp(-t / c2)
            if rho is not None:
                xi += rho[i][c]
            if not math.isfinite(xi):
                raise ValueError("non-finite curve value")
            v.append(xi)
        beta = (v[2] * (sx - smin) + v[1] * (smax - sx)) / (smax - smin)
        out[0].append(rs * v[0] * (xs / xs[y0]) ** beta)
        out[1].append(np.full(nx * nr, v[3]))
        out[2].append(v[4] * (v[5] - rs))
        out[3].append(np.full(nx * nr, v[6]))
        out[4].append(np.full(nx * nr, v[7]))
    return [np.array(a) for a in out]


def lv_params_c1(x, rate):
    n = len(x)
    dr, vo = np.zeros(n), np.zeros(n)
    lib.or_lv_params_c1(n, np.ascontiguousarray(x), rate, dr, vo)
    return dr, vo


def lv_params_cev(x, rate, sigma0, beta, skew=0.0, curv=0.0):
    n = len(x)
    dr, vo = np.zeros(n), np.zeros(n)
    lib.or_lv_params_cev(n, np.ascontiguousarray(x), rate, sigma0, beta, skew, curv, dr, vo)
    return dr, vo


def first_step(gen, dt):
    d = gen.shape[0]
    U = np.zeros(d * d)
    lib.or_first_step(d, np.ascontiguousarray(gen.reshape(-1, order="F")), dt, U)
    return U.reshape(d, d, order="F")


def dgemm(A, B):
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    Cm = np.zeros(m * n)
    lib.or_dgemm(m, n, k, np.ascontiguousarray(A.reshape(-1, order="F")), m,
                 np.ascontiguousarray(B.reshape(-1, order="F")), k, Cm, m)
    return Cm.reshape(m, n, order="F")


def square_chain(U, nsq):
    d = U.shape[0]
    u = np.ascontiguousarray(U.reshape(-1, order="F")).copy()
    w = np.zeros_like(u)
    lib.or_square_chain(d, nsq, u, w)
    return u.reshape(d, d, order="F")


def sck_f32(tpk):
    """tpk: (nm, d, d) float32 with [m][y0, y] semantics, column-major per
    matrix; returns the same layout."""
    nm, d, _ = tpk.shape
    src = np.ascontiguousarray(np.stack([t.reshape(-1, order="F") for t in tpk]).reshape(-1))
    out = np.zeros_like(src)
    lib.or_sck_f32(d, nm, src, out)
    return out


def sck_rows2_f32(tpk, chunk=64, plus_identity=0):
    """the row-contiguous table in the two-level order; tpk: (nm, d, d)
    float32 with [m][y0, y] semantics; returns (nm, d, d) [m][row][y]"""
    nm, d, _ = tpk.shape
    src = np.ascontiguousarray(np.stack([t.reshape(-1, order="F") for t in tpk]).reshape(-1))
    out = np.zeros_like(src)
    lib.or_sck_rows2_f32(d, nm, chunk, plus_identity, src, out)
    return out.reshape(nm, d, d)


def mt_uniforms(seed, n):
    out = np.zeros(n, dtype=np.float32)
    lib.or_mt_uniforms(seed, n, out)
    return out


def mt_raw(seed, n):
    st = np.zeros(625, dtype=np.uint32)
    lib.or_mt19937_init(seed, st)
    return np.array([lib.or_mt19937_next(st) for _ in range(n)], dtype=np.uint32)


DCMT607 = dict(n=19, m=9, umask=0xFFFFFFFE, lmask=0x1, s0=12, s1=18)
MT19937 = dict(n=624, m=397, umask=0x80000000, lmask=0x7FFFFFFF, s0=11, s1=18)


def tgfsr_seed(seeds, n):
    """lane states (lane-major) from per-lane seeds"""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    st = np.zeros(len(seeds) * n, dtype=np.uint32)
    lib.or_tgfsr_seed(len(seeds), n, seeds, st)
    return st


def tgfsr_run(params, st, count, matrix_a, mask_b, mask_c):
    """count raw outputs per lane, out[c][lane]; st (lane-major) is advanced"""
    nl = len(matrix_a)
    out = np.zeros(nl * count, dtype=np.uint32)
    ma, mb, mc = (np.ascontiguousarray(a, dtype=np.uint32) for a in (matrix_a, mask_b, mask_c))
    lib.or_tgfsr_run(nl, params["n"], params["m"], params["umask"], params["lmask"], params["s0"],
                     params["s1"], ma, mb, mc, st, count, out)
    return out.reshape(count, nl)


def mc_hash(tpk, nblocks, nscen, y0, m_k, seed0):
    """the reference's CPU sampler: tpk (nm, d, d) FP64 with [m][y0, y]
    semantics; returns (states [nblocks][nk][nscen], ctpk, hash)"""
    nm, d, _ = tpk.shape
    t = np.ascontiguousarray(np.stack([a.reshape(-1, order="F") for a in tpk]).reshape(-1), dtype=np.float64)
    ct = np.zeros_like(t)
    hs = np.zeros(nm * d * d, dtype=np.int32)
    lib.or_mc_hash_build(d, nm, t, ct, hs)
    mk = np.ascontiguousarray(m_k, dtype=np.int32)
    y = np.zeros(nblocks * len(mk) * nscen, dtype=np.int16)
    lib.or_mc_hash_run(d, ct, hs, nblocks, nscen, len(mk), y0, mk, seed0, y)
    return y.reshape(nblocks, len(mk), nscen), ct, hs


def mc_chain(nstreams, seed0, groups, nk, d, y0, ctpk_flat, m_k, layout_rows=0, per=None):
    """states [nk][nscen] of the build's stream mapping: stream s serves
    `per` paths (default groups * 64), scenario s*per + j"""
    per = groups * 64 if per is None else per
    nscen = nstreams * per
    out = np.zeros(nscen * nk, dtype=np.int16)
    lib.or_mc_chain(nstreams, seed0, per, nk, d, y0, np.ascontiguousarray(ctpk_flat,
                    dtype=np.float32), np.ascontiguousarray(m_k, dtype=np.int32),
                    layout_rows, out)
    return out.reshape(nk, nscen)


def bs_call(S, K, r, sigma, T):
    return lib.or_bs_call(S, K, r, sigma, T)


# ---------------------------------------------------------------------------
# scenario solve in FP64
# ---------------------------------------------------------------------------

def model_generator(cfg):
    """Grid, FP64 generator and the spot node of a configuration."""
    nx, nr = cfg.nx, cfg.nr
    x, x0 = xgrid(nx, cfg.xlo, cfg.xhi, cfg.spot)
    if cfg.vol_kind == 0:
        sig = [cfg.sigma_r[i] for i in range(nr)]
        g = rs_generator(x, sig, cfg.rate, cfg.q_off)
    elif cfg.vol_kind == 1:
        assert nr == 1
        dr, vo = lv_params_c1(x, cfg.rate)
        g = lv_generator(x, dr, vo)
    else:
        assert nr == 1
        dr, vo = lv_params_cev(x, cfg.rate, cfg.sigma0, cfg.beta, cfg.skew, cfg.curv)
        g = lv_generator(x, dr, vo)
    return x, x0, g


def kernel_fp64(cfg):
    """U(DeltaT) = (I + L dt0)^(2^nsq), dt0 = DeltaT 2^-nsq, in FP64."""
    x, x0, g = model_generator(cfg)
    dt0 = math.ldexp(cfg.DeltaT, -cfg.nsq)
    U = first_step(g, dt0)
    return x, x0, square_chain(U, cfg.nsq)


def payoff(cfg, x):
    ks = np.array([cfg.strikes[i] for i in range(cfg.nstrikes)])
    xs = np.tile(x, cfg.nr)  # state y -> price x[y % nx]
    return np.maximum(xs[:, None] - ks[None, :], 0.0)  # d x nstrikes


def solve_fp64(cfg, U=None):
    """Prices [nr][n_mat][nstrikes] of the FP64 method (same discretisation
    as the GPU method: same grid, dt0, squaring count, steps)."""
    if U is None:
        x, x0, U = kernel_fp64(cfg)
    else:
        x, x0 = xgrid(cfg.nx, cfg.xlo, cfg.xhi, cfg.spot)
    d, nx, nr = U.shape[0], cfg.nx, cfg.nr
    P = payoff(cfg, x)
    rows = [x0 + nx * r for r in range(nr)]
    if cfg.price_kind == 0:  # backward induction
        V = np.ascontiguousarray(P.reshape(-1, order="F")).copy()
        dfs = math.exp(-cfg.rate * cfg.DeltaT)
        lib.or_backward_induction(d, np.ascontiguousarray(U.reshape(-1, order="F")),
                                  cfg.nsteps, dfs, cfg.nstrikes, V)
        V = V.reshape(d, cfg.nstrikes, order="F")
        return np.stack([V[y][None, :] for y in rows])
    out = np.zeros((nr, cfg.nsteps, cfg.nstrikes))
    Uk = U
    for k in range(1, cfg.nsteps + 1):
        if k > 1:
            Uk = dgemm(Uk, U)
        dfk = math.exp(-cfg.rate * cfg.DeltaT * k)
        for r, y in enumerate(rows):
            out[r, k - 1] = dfk * (Uk[y, :] @ P)
    return out


def rel_err(a, b, floor=None):
    """|a-b| / max(|b|, floor); floor defaults to 1e-3 * max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if floor is None:
        floor = 1e-3 * float(np.max(np.abs(b)))
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


# ---------------------------------------------------------------------------
# surrogate control variate (no reference counterpart)
# ---------------------------------------------------------------------------

def local_vol_cev(x, sigma0, beta, a, b):
    """sigma0 (x/100)^(beta-1) (1 + a l + b l^2), l = ln(x/100), floored at
    1e-4 (the lattice's CEV family, lv_params_cev)"""
    l = np.log(np.asarray(x, dtype=np.float64) / 100.0)
    return np.maximum(sigma0 * np.exp((beta - 1.0) * l) * (1.0 + a * l + b * l * l), 1e-4)


def bbf_vol(p, S, K):
    """Berestycki-Busca-Florent short-maturity implied vol: the harmonic
    mean of the local vol along the log path, 1/sigma = int_0^1 ds /
    sigma_loc(S (K/S)^s), by 8-point Gauss-Legendre on [0, 1]"""
    xg, wg = np.polynomial.legendre.leggauss(8)
    s = 0.5 * (1.0 + xg)
    xs = S * np.exp(s * math.log(K / S))
    inv = np.sum(0.5 * wg / local_vol_cev(xs, p[0], p[1], p[3], p[4]))
    return 1.0 / inv


def bs_call_np(S, K, r, sigma, T):
    from scipy.special import erfc
    sq = sigma * math.sqrt(T)
    d1 = (math.log(S / K) + (r + 0.5 * sigma * sigma) * T) / sq
    d2 = d1 - sq
    return S * 0.5 * erfc(-d1 / math.sqrt(2.0)) - K * math.exp(-r * T) * 0.5 * erfc(-d2 / math.sqrt(2.0))


def surr_control_variate(params, strikes, nmat, dT, spot):
    """[n][nmat * nstrikes] Black-Scholes calls at the BBF vol, maturity-major"""
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    out = np.zeros((len(params), nmat * len(strikes)))
    for i, p in enumerate(params):
        for k in range(nmat):
            for c, K in enumerate(strikes):
                out[i, k * len(strikes) + c] = bs_call_np(spot, K, p[2], bbf_vol(p, spot, K), (k + 1) * dT)
    return out


# ---------------------------------------------------------------------------
# MLP (or_mlp_forward / or_mlp_backward): W_l [out][in] row-major
# ---------------------------------------------------------------------------

_PP = C.POINTER(C.POINTER(C.c_double))


def _ptrs(arrs):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    tab = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return arrs, tab


def mlp_forward(Ws, bs, x):
    """outputs [batch][out] and the pre-activations of every layer"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dims = (C.c_int * (len(Ws) + 1))(*([Ws[0].shape[1]] + [W.shape[0] for W in Ws]))
    Wk, Wt = _ptrs(Ws)
    bk, bt = _ptrs(bs)
    n = x.shape[0]
    out = np.zeros((n, Ws[-1].shape[0]))
    acts = np.zeros(n * sum(W.shape[0] for W in Ws))
    lib.or_mlp_forward(len(Ws), dims, Wt, bt, n, x.ctypes.data_as(C.POINTER(C.c_double)),
                       out.ctypes.data_as(C.POINTER(C.c_double)),
                       acts.ctypes.data_as(C.POINTER(C.c_double)))
    return out, acts


def mlp_backward(Ws, bs, x, dout):
    """gradients (dW list, db list) of sum(out .* dout) over the batch"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dout = np.ascontiguousarray(dout, dtype=np.float64)
    _, acts = mlp_forward(Ws, bs, x)
    dims = (C.c_int * (len(Ws) + 1))(*([Ws[0].shape[1]] + [W.shape[0] for W in Ws]))
    Wk, Wt = _ptrs(Ws)
    bk, bt = _ptrs(bs)
    dW = [np.zeros(W.shape) for W in Ws]
    db = [np.zeros(W.shape[0]) for W in Ws]
    dWk, dWt = _ptrs(dW)
    dbk, dbt = _ptrs(db)
    lib.or_mlp_backward(len(Ws), dims, Wt, bt, x.shape[0], x.ctypes.data_as(C.POINTER(C.c_double)),
                        acts.ctypes.data_as(C.POINTER(C.c_double)),
                        dout.ctypes.data_as(C.POINTER(C.c_double)), dWt, dbt)
    return dWk, dbk
