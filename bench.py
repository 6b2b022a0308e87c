#!/usr/bin/env python3
"""Benchmark: priced scenarios/s of the operator-method pricing path on one
MI355X at <= 1e-4 relative price error against FP64.

Workload (one "step" = one priced scenario): the 2048-state regime-
switching local-vol lattice (256 price nodes x 8 vol regimes 0.1..0.8,
regime rate 0.5/yr, r = 0.02), FP64 generator assembly fused with the first
step on the GPU, 16 squarings of the 2048 x 2048 kernel to the daily step,
then 252 backward-induction steps for 41 strikes.  Each method's whole
solve is one HIP-graph launch; inputs are resident in HBM before timing.
`value` is the fastest method whose prices are within 1e-4 of the FP64
oracle (the plain FP32 method, the drop-in, is timed beside it and fails
that bar at this size).

Also measured, in the same run:
  roofline      the headline method's dominant kernel, the band-sparse
                backward step of a scenario group, timed with HIP events on
                the library stream, in GB/s against HBM (and beside the
                measured Infinity-Cache read rate); the FP32 squaring GEMM
                against the FP32 matrix-core peak;
  methods       every method's scenarios/s and error;
  sweep         the 64-lattice calibration sweep, plain vs batched W form;
  mc            the MT19937 chain sampler on the daily table,
                2^20 paths x 252 steps, in transitions/s;
  surrogate     the learned surrogate with solver fallback (training
                labels from the solver, held-out judgement, 2^16 draws);
  cpu_baseline  the CPU oracle's FP64 solve of the same scenario (the
                reference's host method restated in C), on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-surrogate ...]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/counters_*.md: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
    `bench.py --probe`, scaled by the copy kernel of known size), or None"""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "counters_*.md"))):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch") and kernel in line:
                try:
                    return float(line.split()[-1])
                except ValueError:
                    return None
    return None


def band_cutoffs(H):
    """the band-sparse backward step's cutoffs: a 16 x 32 fragment of W is
    read where some |W| >= 2^visit, its lo part where some |W| >= 2^lo"""
    import ctypes as C
    ev, el = C.c_int(0), C.c_int(0)
    H.check(H.lib.ophip_wsq_get_bwd_cutoffs(C.byref(ev), C.byref(el)))
    return {"visit": f"2^{ev.value}", "lo": f"2^{el.value}"}


def roofline_backward(best, methods, wbk, bi_gbs, bi_us, bi_bytes, mall=None):
    """roofline entry of the headline method's dominant kernel: the FP16 x3
    band-sparse backward step (per launch: every lattice's visited hi
    fragments and read lo fragments, and the live value columns), or the FP32
    skinny step; beside HBM's peak, the measured Infinity-Cache read
    bandwidth (the visited fragments of a group fit the 256 MB cache)"""
    if best.startswith("wform_fp16x3"):
        nb = methods[best].get("group", 1)
        w = wbk.get(nb) or wbk[1]
        key = f"{w['kernel'].split('<')[0]}/{nb}"
        r = {"bound": "hbm", "kernel": w["kernel"], "achieved": w["GBps"], "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": w["GBps"] / HBM_PEAK_GBS,
             "traffic": traffic_from_profiles(key),
             "bytes_per_launch": w["bytes_per_launch"], "avg_launch_us": w["avg_launch_us"],
             "dense_bytes_per_launch": w["dense_bytes_per_launch"],
             "visited_fragment_fraction": w["visited_fragment_fraction"],
             "lo_read_fraction": w["lo_read_fraction"]}
        if mall:
            r["mall_read_GBps_measured"] = mall["mall_read_GBps"]
            r["frac_of_mall_read"] = w["GBps"] / mall["mall_read_GBps"]
        return r
    k = "backward step: k_skinny<3> + k_skinny_combine"
    return {"bound": "hbm", "kernel": k, "achieved": bi_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": bi_gbs / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": bi_bytes,
            "avg_launch_us": bi_us}


def bench_mall(H):
    """read bandwidth of the Infinity Cache (MALL) and of HBM: one streaming
    read (ophip_read_probe, 16-byte loads) of a buffer read just before, 20
    times between events; a buffer below the 256 MB cache stays resident,
    1 GiB does not"""
    out = H.DeviceArray.empty(8192, dtype=np.uint32)  # 32 KiB the probe may touch
    res = {}
    for mb in (32, 64, 104, 128, 192, 1024):
        nbytes = mb << 20
        buf = H.DeviceArray(nbytes, np.uint8)
        H.lib.opcuda_ssetall(buf.h, nbytes // 4, np.float32(1.0), 1)
        for _ in range(3):
            H.check(H.lib.ophip_read_probe(buf.h, nbytes, out.h))
        ms = timed_events(H, lambda: H.lib.ophip_read_probe(buf.h, nbytes, out.h), 20)
        H.check()
        res[str(mb)] = nbytes / (ms * 1e-3) / 1e9
        buf.free()
    log(f"[bench] read bandwidth by working set (MB -> GB/s): "
        + ", ".join(f"{k}: {v:.0f}" for k, v in res.items()))
    return {"kernel": "k_read16 (ophip_read_probe)", "read_GBps_by_MB": res,
            "mall_read_GBps": res["104"], "hbm_read_GBps": res["1024"]}


def timed_events(H, fn, reps, a=0, b=1):
    """average ms per call of fn() between two events on the library stream"""
    H.check(H.lib.ophip_event_record(a))
    for _ in range(reps):
        fn()
    H.check(H.lib.ophip_event_record(b))
    ms = H.lib.ophip_event_elapsed_ms(a, b)
    return ms / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte Carlo leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the calibration-sweep leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--no-fp8", action="store_true", help="skip the FP8 W-form method")
    ap.add_argument("--groups", type=int, nargs="*", default=[16, 32],
                    help="group sizes of the FP16 x3 scenario-group method")
    ap.add_argument("--probe-group", type=int, default=32,
                    help="lattices per backward-step launch in --probe")
    ap.add_argument("--surr-train", type=int, default=262144, help="surrogate training draws")
    ap.add_argument("--surr-steps", type=int, default=240000, help="surrogate Adam steps")
    ap.add_argument("--probe", action="store_true",
                    help="only launch the dominant kernel (W-form backward step) 20 times and a "
                         "64 MiB copy kernel 20 times, for counter collection")
    ap.add_argument("--profile-group", type=int, default=0, metavar="N",
                    help="only price one group of N scenarios of the 2048-state lattice: three "
                         "graph launches (for a kernel trace of one launch's kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); a multi-GPU run "
              "would shard the scenarios over the GPUs and gather the prices once; it is "
              "not built or measured here.", file=sys.stderr)
        sys.exit(2)

    from oplib_amd import ophip as H
    from oplib_amd import opmodel as M

    H.init()
    cfg = M.config_rs2048()
    d = cfg.d
    if args.profile_group:
        g = M.Group(M.config_rs2048_scenarios(args.profile_group))
        g.capture()
        for rep in range(3):
            g.launch()
            H.sync()
            log(f"[bench] group launch {rep + 1} of 3 done")
        gold = np.load(os.path.join(ROOT, "tests", "golden", "prices_rs2048_scenarios.npz"))["prices"]
        e = rel_err(g.prices(), gold[:args.profile_group])
        print(json.dumps({"profile_group": args.profile_group, "launches": 3,
                          "max_rel_err": float(e.max())}), flush=True)
        g.close()
        return
    if args.probe:
        nb = args.probe_group
        wb = WformBackward(H, M, M.config_rs2048_scenarios(nb))
        for _ in range(20):
            wb.step()
        sp = M.Solver(M.config_rs2048())  # the plain FP32 solve: 16 k_sgemm squarings, 252 steps
        sp.solve()
        n = 16 << 20  # floats: 64 MiB copied by k_copy4 (opcuda_scopy1)
        A, B = H.DeviceArray.empty(n), H.DeviceArray.empty(n)
        for _ in range(20):
            H.check(H.lib.opcuda_scopy1(B.h, A.h, n), "scopy1")
        H.sync()
        print(json.dumps({"probe": f"FP16 x3 squarings of {nb} lattices (16 rounds each, one lattice per "
                                   f"launch), FP16 x3 backward step ({nb} lattices per launch, kernel "
                                   f"{H.lib.ophip_wsq_bwd_kernel_for(nb, 2048, 48)}) x20, one plain FP32 "
                                   "solve (16 squarings, 252 steps), k_copy4 (64 MiB) x20",
                          "step_bytes_nonzero": wb.bytes, "step_bytes_dense": wb.bytes_dense,
                          "visited_fragment_fraction": wb.nz_frac, "lo_read_fraction": wb.lo_frac,
                          "copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    log(f"[bench] workload rs2048 d={d} nsq={cfg.nsq} steps={cfg.nsteps} strikes={cfg.nstrikes}")
    s = M.Solver(cfg)
    s.solve()
    H.sync()
    eager = s.prices()
    s.capture()
    for _ in range(args.warmup):
        s.launch()
    H.sync()
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.launch()
    H.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.steps / (t1 - t0)
    p = s.prices()
    assert np.array_equal(p, eager), "graph replay differs from the eager solve"
    gold = np.load(os.path.join(ROOT, "tests", "golden", "prices_rs2048.npz"))["prices"]
    den = np.maximum(np.abs(gold), 1e-3 * np.abs(gold).max())
    rel = np.abs(p - gold) / den
    log(f"[bench] {value:.2f} scenarios/s, {ms_per_step:.3f} ms/scenario, "
        f"max rel err {rel.max():.3e}")

    # ---- roofline: the squaring GEMM, events on the library stream ----
    U = H.DeviceArray.empty(d * d)
    W = H.DeviceArray.empty(d * d)
    Ud = s.kernel()  # U(1/252) from the solve: realistic operand values
    U.upload(np.ascontiguousarray(Ud.reshape(-1, order="F")))

    def sq():
        H.lib.ophip_sgemm_strided_batched(d, d, d, np.float32(1), U.h, d, 0, U.h, d, 0,
                                          np.float32(0), W.h, d, 0, 1)
    for _ in range(3):
        sq()
    gemm_ms = timed_events(H, sq, 20)
    gemm_flop = 2.0 * d ** 3
    gemm_tflops = gemm_flop / (gemm_ms * 1e-3) / 1e12
    log(f"[bench] squaring GEMM {gemm_ms * 1e3:.1f} us, {gemm_tflops:.1f} TFLOP/s")

    # ---- backward-induction step: skinny GEMM, GB/s ----
    ncp = 48
    X = H.DeviceArray.empty(d * ncp)
    Y = H.DeviceArray.empty(d * ncp)
    X.upload(np.ones(d * ncp, dtype=np.float32))

    def bi():
        H.lib.ophip_sgemm_skinny(d, d, ncp, np.float32(0.9999), U.h, d, X.h, ncp, Y.h, ncp)
    for _ in range(3):
        bi()
    bi_us = timed_events(H, bi, 50) * 1e3
    bi_bytes = 4.0 * d * d + 4.0 * d * cfg.nstrikes * 2
    bi_gbs = bi_bytes / (bi_us * 1e-6) / 1e9
    log(f"[bench] backward step {bi_us:.2f} us, {bi_gbs:.0f} GB/s algorithmic")

    # ---- the same scenario with the U = I + W mixed-precision squarings ----
    methods = {"plain_fp32": {"scenarios_per_s": value, "ms": ms_per_step,
                              "max_rel_err": float(rel.max()),
                              "meets_1e-4": bool(rel.max() <= 1e-4)}}
    for prec, name in ((1, "wform_fp16x1"), (2, "wform_fp16x3"), (3, "wform_bf16x3"),
                       (4, "wform_fp8x3")):
        if prec == 4 and args.no_fp8:
            continue
        methods[name] = bench_method(H, M, prec, args.steps, args.warmup, gold)
        log(f"[bench] {name}: {methods[name]}")
    # many scenarios of the same lattice priced together (FP16 x3): `steps`
    # timed group launches after `warmup` untimed ones
    for nb in args.groups:
        name = f"wform_fp16x3_group{nb}"
        methods[name] = bench_group(H, M, nb, args.steps, args.warmup)
        log(f"[bench] {name}: {methods[name]}")
    # Infinity-Cache and HBM read bandwidth (the backward step's ceiling)
    mall = bench_mall(H)

    # ---- the FP16 x3 backward step (dominant kernel of the x3 methods), on
    # one lattice and on the largest group, HIP events on the library stream
    wbk = {}
    for nb in sorted({1} | set(args.groups)):
        wb = WformBackward(H, M, M.config_rs2048_scenarios(nb))
        for _ in range(4):
            wb.step()
        us = timed_events(H, wb.step, 100) * 1e3
        H.check()
        kid = H.lib.ophip_wsq_bwd_kernel_for(nb, d, 48)
        kname = {0: "k_wbwd1<FP16, 3, 8> (dense)", 2: "k_wbwd_sp<FP16, 16, 3> (16-row tiles)",
                 3: "k_wbwd_blk<FP16, 4, 4, 2> (64-row blocks)",
                 4: "k_wbwd_blk<FP16, 2, 8, 2> (32-row blocks)"}[kid]
        wbk[nb] = {"kernel": f"{kname}, {nb} lattice{'s' if nb > 1 else ''} per launch",
                   "kernel_id": kid,
                   "avg_launch_us": us, "bytes_per_launch": wb.bytes,
                   "dense_bytes_per_launch": wb.bytes_dense,
                   "visited_fragment_fraction": wb.nz_frac, "lo_read_fraction": wb.lo_frac,
                   "GBps": wb.bytes / (us * 1e-6) / 1e9,
                   "dense_equivalent_GBps": wb.bytes_dense / (us * 1e-6) / 1e9}
        log(f"[bench] FP16x3 backward step x{nb}: {us:.2f} us, {wbk[nb]['GBps']:.0f} GB/s of visited "
            f"fragments ({wb.nz_frac:.3f} of the matrix, lo parts {wb.lo_frac:.3f}), "
            f"{wbk[nb]['dense_equivalent_GBps']:.0f} GB/s dense-equivalent")
        del wb

    # ---- precision study: the same generator to the 12 monthly maturities
    # (41 strikes x 12 maturities x 8 regimes), every method's error ----
    gold_m = np.load(os.path.join(ROOT, "tests", "golden", "prices_rs_monthly.npz"))["prices"]
    precision = {"workload": "2048-state regime lattice, 16 squarings to U(1/12), propagation to "
                             "12 monthly maturities, 41 strikes, prices at every regime's spot node"}
    for prec, name in ((0, "plain_fp32"), (1, "wform_fp16x1"), (2, "wform_fp16x3"),
                       (3, "wform_bf16x3"), (4, "wform_fp8x3")):
        if prec == 4 and args.no_fp8:
            continue
        precision[name] = bench_method(H, M, prec, max(3, args.steps // 4), 1, gold_m,
                                       M.config_rs_monthly)
        log(f"[bench] precision study {name}: {precision[name]}")

    # ---- calibration sweep: 64 lattices, plain FP32 one by one vs batched W form ----
    sweep = None
    if not args.no_sweep:
        sweep = bench_sweep(H, M, max(3, args.steps // 10))

    # ---- Monte Carlo on the daily table ----
    mc = None
    U_host = None
    if not args.no_mc:
        # the daily kernel of the FP16 x3 method (within 3e-6 of FP64), so
        # that the Monte Carlo prices can be held against the FP64 oracle
        cw = M.config_rs2048()
        cw.precision = M.OPM_PREC_F16X3
        sw = M.Solver(cw)
        sw.solve()
        Uw = H.DeviceArray.empty(d * d)
        U_host = sw.kernel()
        Uw.upload(np.ascontiguousarray(U_host.reshape(-1, order="F")))
        sw.close()
        mc = bench_mc(H, Uw, d, cfg, gold[1, 0, :])

    # ---- stochastic-volatility generator assembly (FP64) at 256 x 8 states ----
    sv = bench_sv(H, cfg)

    # ---- learned surrogate with fallback ----
    surrogate = None
    if not args.no_surrogate:
        surrogate = bench_surrogate(H, M, args.surr_train, args.surr_steps, log)

    # ---- CPU baseline: the oracle's FP64 solve of the same scenario ----
    cpu = None
    if not args.no_cpu:
        cpu = bench_cpu(cfg, U_host)

    # the headline: the fastest method on this workload whose prices are
    # within 1e-4 of FP64 (the plain FP32 method is reported beside it)
    ok = {k: v for k, v in methods.items() if v["meets_1e-4"]}
    best = max(ok, key=lambda k: ok[k]["scenarios_per_s"]) if ok else "plain_fp32"
    bm = methods[best]
    names = {"plain_fp32": "plain FP32", "wform_fp16x1": "U = I + W, FP16 x1 squarings",
             "wform_fp16x3": "U = I + W, FP16 x3 (hi/lo) squarings",
             "wform_bf16x3": "U = I + W, BF16 x3 (hi/lo) squarings",
             "wform_fp8x3": "U = I + W, FP8 e4m3 x3 (hi/lo, power-of-two scales) squarings"}
    for nb in args.groups:
        names[f"wform_fp16x3_group{nb}"] = (f"U = I + W, FP16 x3 (hi/lo) squarings, {nb} scenarios "
                                            "per graph launch (one squaring launch per round and one "
                                            "band-sparse backward-step launch per step for all)")
    grp = bm.get("group", 1)
    line = {
        "metric": "priced scenarios/s at <=1e-4 rel err vs FP64",
        "value": bm["scenarios_per_s"],
        "unit": "scenarios/s",
        "n_gpus": 1,
        # a step: one graph launch of the headline method (for a group, all of
        # its scenarios); `steps` launches timed after `warmup` untimed ones
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bm.get("ms_per_group", bm["ms"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if best == "plain_fp32" else best.split("_")[1][:4] + " (MFMA, fp32 accumulate) + fp32",
        "data": "synthetic",
        "config": {"workload": "regime-switching local vol, 256 x 8 = 2048 states, "
                               "16 squarings to U(1/252), 252 backward steps, 41 strikes"
                               + (f"; {grp} scenarios per launch (the workload's parameters, then "
                                  "seeded vol levels, regime rates and rates)" if grp > 1 else ""),
                   "scenarios_per_launch": grp,
                   "d": d, "squarings": cfg.nsq, "backward_steps": cfg.nsteps,
                   "strikes": cfg.nstrikes, "method": names[best],
                   "band_cutoffs": band_cutoffs(H),
                   "plain_fp32_scenarios_per_s": value,
                   "speedup_vs_plain_fp32": bm["scenarios_per_s"] / value},
        "error_vs_fp64": {"max_rel": bm["max_rel_err"],
                          "rms_rel": bm.get("rms_rel_err"),
                          "denominator": "max(|p64|, 1e-3*max|p64|)",
                          "prices": int(bm.get("group", 1) * rel.size),
                          "meets_1e-4": bool(bm["meets_1e-4"])},
        # dominant kernel of the headline method: the backward-induction
        # step (252 per scenario) streams the 2048 x 2048 matrix's hi/lo
        # copies (FP16 x3 W form; the FP32 methods' step is in "gemv")
        "roofline": roofline_backward(best, methods, wbk, bi_gbs, bi_us, bi_bytes, mall),
        "read_bandwidth": mall,
        "roofline_squaring": {"bound": "mfma", "kernel": "k_sgemm<128,128> (plain FP32 squaring)",
                              "achieved": gemm_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": gemm_tflops / FP32_PEAK_TFLOPS,
                              "flop_per_launch": gemm_flop, "avg_launch_us": gemm_ms * 1e3},
        "wform_backward": {str(k): v for k, v in wbk.items()},
        "gemv": {"kernel": "k_skinny<3> + combine (FP32 backward step)", "avg_us": bi_us,
                 "algorithmic_GBps": bi_gbs, "peak_GBps": HBM_PEAK_GBS,
                 "frac": bi_gbs / HBM_PEAK_GBS},
        "methods": methods,
        "precision_study": precision,
        "sweep": sweep,
        "mc": mc,
        "sv_generator": sv,
        "surrogate": surrogate,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


class WformBackward:
    """the FP16 x3 method's dominant kernel on its real operands: each
    lattice (a 2048-state regime scenario) squared 16 times in the W form,
    W's fragment copies and tile lists laid out, then the band-sparse
    backward step (k_wbwd_blk or k_wbwd_sp) of all of them in one launch, alternating two
    value buffers per lattice"""

    def __init__(self, H, M, cfgs, ncp=48):
        import ctypes as C
        import math
        self.H, self.d, self.ncp, self.n = H, cfgs[0].d, ncp, len(cfgs)
        d, mode = self.d, 1
        self.keep, self.states = [], []
        for cfg in cfgs:
            x, _ = M.grid(cfg)
            dr, vo = M.params(cfg)
            arrs = [H.DeviceArray.from_numpy(a) for a in (x, dr, vo)]
            st = H.DeviceArray(H.lib.ophip_wsq_bytes(d, mode), np.uint8)
            self.keep += arrs
            self.states.append(st)
            w0 = H.lib.ophip_wsq_w(st.h, d, mode, 0)
            H.check(H.lib.ophip_sek_rs(cfg.nx, cfg.nr, arrs[0].h, arrs[1].h, arrs[2].h, cfg.q_off,
                                       math.ldexp(cfg.DeltaT, -cfg.nsq), 2, w0))
            sh, ds = (C.c_uint * 1)(st.h), (C.c_int * 1)(d)
            plan = H.lib.ophip_wsq_plan(1, sh, ds, None, mode, 0, 0)
            H.check(H.lib.ophip_wsq_exec(plan, cfg.nsq))
            H.sync()
            H.lib.ophip_wsq_plan_free(plan)
        cfg = cfgs[0]
        self.buf = cfg.nsq & 1
        nkt = d // 32
        self.nz = 0  # visited 16 x 32 fragments over all lattices
        self.nlo = 0  # of which the lo part is read
        for st in self.states:
            H.check(H.lib.ophip_wsq_bwd_layout(st.h, d, mode, self.buf))
            tl = np.empty((d // 16) * (nkt + 1), dtype=np.int32)
            H.check(H.lib.opcuda_memcpy_d2h(H.lib.ophip_wsq_array(st.h, d, mode, 0, 7),
                                            tl.ctypes.data, tl.nbytes))
            tl = tl.reshape(d // 16, nkt + 1)
            self.nz += int(tl[:, 0].sum())
            self.nlo += sum(int((tl[i, 1:1 + tl[i, 0]] >= 65536).sum()) for i in range(d // 16))
        H.sync()
        xs = np.tile(x, cfg.nr)
        pay = np.zeros((d, ncp), np.float32)
        for k in range(cfg.nstrikes):
            pay[:, k] = np.maximum(xs - cfg.strikes[k], 0.0)
        self.V = [[H.DeviceArray.from_numpy(pay, order="C"), H.DeviceArray.empty(d * ncp)]
                  for _ in cfgs]
        self.VT = [[H.DeviceArray.empty(2 * d * ncp, np.uint16) for _ in range(2)] for _ in cfgs]
        for i in range(self.n):
            H.check(H.lib.ophip_wsq_bwd_prep(d, ncp, mode, self.V[i][0].h, self.VT[i][0].h))
        arr = lambda ty, v: (ty * self.n)(*v)  # noqa: E731
        self.st = arr(C.c_uint, [s.h for s in self.states])
        self.df = arr(C.c_float, [0.9999] * self.n)
        self.X = [arr(C.c_uint, [v[p].h for v in self.V]) for p in (0, 1)]
        self.XT = [arr(C.c_uint, [v[p].h for v in self.VT]) for p in (0, 1)]
        # the solves' steps carry the values as their hi/lo copies only (no
        # FP32 values in or out): time the step the same way
        self.none = arr(C.c_uint, [0] * self.n)
        self.t = 0
        # bytes per launch: dense algorithmic (the hi/lo copies of every W,
        # 4 d^2 each, and the live value columns read and written) and the
        # part the band-sparse kernel must read (1 KB per visited fragment's
        # hi part and per read lo part)
        self.bytes_dense = self.n * (4.0 * d * d + 4.0 * d * cfg.nstrikes * 2)
        self.bytes = 1024.0 * (self.nz + self.nlo) + self.n * 4.0 * d * cfg.nstrikes * 2
        self.nz_frac = self.nz / (self.n * (d // 16) * nkt)
        self.lo_frac = self.nlo / (self.n * (d // 16) * nkt)

    def step(self):
        a, b = self.t & 1, (self.t + 1) & 1
        self.H.lib.ophip_wsq_bwd_step_batch(self.n, self.st, self.d, 1, self.buf, self.ncp, self.df,
                                            self.none, self.XT[a], self.none, self.XT[b])
        self.t += 1


def rel_err(p, gold):
    den = np.maximum(np.abs(gold), 1e-3 * np.abs(gold).max())
    return np.abs(p - gold) / den


def bench_method(H, M, prec, steps, warmup, gold, make=None):
    cfg = (make or M.config_rs2048)()
    cfg.precision = prec
    s = M.Solver(cfg)
    s.solve()
    s.capture()
    for _ in range(warmup):
        s.launch()
    H.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.launch()
    H.sync()
    t = (time.perf_counter() - t0) / steps
    e = rel_err(s.prices(), gold)
    s.close()
    return {"scenarios_per_s": 1.0 / t, "ms": t * 1e3, "max_rel_err": float(e.max()),
            "rms_rel_err": float(np.sqrt(np.mean(e ** 2))), "meets_1e-4": bool(e.max() <= 1e-4)}


def bench_group(H, M, nb, steps, warmup):
    """nb scenarios of the 2048-state lattice (the workload, then seeded vol
    levels, regime rates and rates) priced together in FP16 x3: one
    squaring launch per round and one backward-step launch per step for
    all of them, the whole group one graph launch"""
    gold = np.load(os.path.join(ROOT, "tests", "golden", "prices_rs2048_scenarios.npz"))["prices"]
    nb = min(nb, len(gold))
    g = M.Group(M.config_rs2048_scenarios(nb))
    g.solve()
    g.capture()
    for _ in range(warmup):
        g.launch()
    H.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
    H.sync()
    t = (time.perf_counter() - t0) / steps
    e = rel_err(g.prices(), gold[:nb])
    g.close()
    return {"scenarios_per_s": nb / t, "ms": t * 1e3 / nb, "ms_per_group": t * 1e3, "group": nb,
            "max_rel_err": float(e.max()), "rms_rel_err": float(np.sqrt(np.mean(e ** 2))),
            "meets_1e-4": bool(e.max() <= 1e-4)}


def bench_sweep(H, M, reps):
    gold_path = os.path.join(ROOT, "tests", "golden", "prices_sweep.npz")
    gold = np.load(gold_path)["prices"] if os.path.exists(gold_path) else None
    out = {"workload": "64 CEV lattices, d in {384..1536}, 17-18 squarings to U(0.1), "
                       "21 strikes x 10 maturities", "scenarios": 64}
    for prec, name in ((0, "plain_fp32_unbatched"), (2, "wform_fp16x3_batched"),
                       (1, "wform_fp16x1_batched"), (3, "wform_bf16x3_batched")):
        cfgs = M.config_sweep(precision=prec)
        b = M.Batch(cfgs)
        b.solve()
        b.capture()
        b.launch()
        H.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.launch()
        H.sync()
        t = (time.perf_counter() - t0) / reps
        p = b.prices()
        r = {"sweeps_per_s": 1.0 / t, "scenarios_per_s": 64.0 / t, "ms_per_sweep": t * 1e3,
             "tflops_algorithmic": M.sweep_flops(cfgs) / t / 1e12}
        if gold is not None:
            e = rel_err(p, gold)
            r.update({"max_rel_err": float(e.max()), "meets_1e-4": bool(e.max() <= 1e-4)})
        out[name] = r
        log(f"[bench] sweep {name}: {r}")
        b.close()
    base = out["plain_fp32_unbatched"]["scenarios_per_s"]
    for k in list(out):
        if isinstance(out[k], dict) and k != "plain_fp32_unbatched":
            out[k]["speedup_vs_plain"] = out[k]["scenarios_per_s"] / base
    return out


def bench_mc(H, U, d, cfg, op_prices):
    # cumulative table of the daily kernel, row-contiguous
    ct = H.DeviceArray.empty(d * d)
    H.check(H.lib.ophip_sck_rows(d, 1, U.h, ct.h))
    sck_ms = timed_events(H, lambda: H.lib.ophip_sck_rows(d, 1, U.h, ct.h), 5)
    H.check()
    sck_gbs = 8.0 * d * d / (sck_ms * 1e-3) / 1e9  # 4 B read + 4 B written per entry
    log(f"[bench] cumulative table {sck_ms * 1e3:.1f} us, {sck_gbs:.0f} GB/s")
    nk = 252
    nscen = 1 << 20
    mk = H.DeviceArray.from_numpy(np.zeros(nk, dtype=np.int32))
    ysk = H.DeviceArray.empty(nscen * nk, dtype=np.int16)
    y0 = 64 + 256 * 1

    gd = H.DeviceArray.empty(d * 257, dtype=np.int16)
    H.check(H.lib.ophip_mc_guide(d, 1, ct.h, gd.h))
    guide_us = timed_events(H, lambda: H.lib.ophip_mc_guide(d, 1, ct.h, gd.h), 5) * 1e3
    log(f"[bench] guide table {guide_us:.1f} us")
    # one stream per wavefront of 64 paths (16384 streams seeded 5489 + i,
    # the workload's mapping), and 4096 streams of 256 paths beside it
    by_streams = {}
    for nrng in (4096, 16384):
        H.check(H.lib.ophip_mc_set_nrng(nrng))
        mt = H.DeviceArray.empty(625 * nrng, dtype=np.uint32)
        seeds = (5489 + np.arange(nrng)).astype(np.uint32)
        H.lib.opcuda_mc_setseed(seeds.ctypes.data, mt.h)
        H.check()

        def run_bisect():
            H.lib.ophip_mc1f_rows(mt.h, y0, ysk.h, nscen, nk, d, ct.h, mk.h, 0)
        run_bisect()
        H.sync()
        ms_bisect = timed_events(H, run_bisect, 3)

        def run():
            H.lib.ophip_mc1f_guided(mt.h, y0, ysk.h, nscen, nk, d, ct.h, gd.h, mk.h)
        run()
        H.sync()
        ms = timed_events(H, run, 3)
        by_streams[nrng] = {"ms_guided": ms, "ms_full_bisection": ms_bisect}
        log(f"[bench] MC {nrng} streams: guided {ms:.2f} ms, bisection {ms_bisect:.2f} ms per 2^20 x 252")
    H.check(H.lib.ophip_mc_set_nrng(4096))
    nrng = 16384
    ms, ms_bisect = by_streams[nrng]["ms_guided"], by_streams[nrng]["ms_full_bisection"]
    trans = float(nscen) * nk
    log(f"[bench] MC {ms:.2f} ms per 2^20 x 252, {trans / ms / 1e6:.2f} G transitions/s")
    # the paths' call payoffs at T = 1 yr evaluated on the device, against the
    # backward-induction prices of the same FP32 kernel (regime 1's spot)
    import ctypes as C
    from oplib_amd import opmodel as M
    x, _ = M.grid(cfg)
    xs = H.DeviceArray.from_numpy(np.tile(x, cfg.nr).astype(np.float32))
    ks = np.array(cfg.strike_list(), dtype=np.float32)
    dfs = np.array([math.exp(-cfg.rate * cfg.DeltaT * nk)])
    steps = np.array([nk - 1], dtype=np.int32)
    out = H.DeviceArray.empty(2 * len(ks), dtype=np.float64)

    def pay():
        H.lib.ophip_mc_payoff(ysk.h, nscen, nk, 1, steps.ctypes.data_as(C.POINTER(C.c_int)),
                              dfs.ctypes.data_as(C.POINTER(C.c_double)), xs.h, d,
                              ks.ctypes.data_as(C.POINTER(C.c_float)), len(ks), out.h)
    pay()
    H.check()
    pay_us = timed_events(H, pay, 10) * 1e3
    m = out.download().reshape(2, len(ks))
    se = np.sqrt(np.maximum(m[1] - m[0] ** 2, 0.0) / nscen)
    z = np.abs(m[0] - op_prices) / np.maximum(se, 1e-300)
    log(f"[bench] MC payoffs on the device {pay_us:.1f} us; max |MC - FP64 operator| = {z.max():.2f} SE")
    ysk.free()
    log(f"[bench] MC full bisection {ms_bisect:.2f} ms")
    return {"kernel": "k_mc1f_guided (MT19937 per wavefront + guide table + bisection)",
            "paths": nscen, "ms_full_bisection": ms_bisect, "guide_table_us": guide_us,
            "payoff_on_device_us": pay_us,
            "price_vs_operator": {"strikes": len(ks), "max_abs_diff_in_standard_errors": float(z.max()),
                                  "max_rel_diff": float((np.abs(m[0] - op_prices) /
                                                         np.maximum(np.abs(op_prices), 1e-3)).max()),
                                  "operator": "FP64 oracle backward induction, regime 1 spot (the chains "
                                              "sample the FP16 x3 method's U(1/252))"},
            "cumulative_table_us": sck_ms * 1e3, "cumulative_table_GBps": sck_gbs,
            "steps": nk, "ms": ms, "transitions_per_s": trans / (ms * 1e-3),
            "streams": nrng, "paths_per_stream": nscen // nrng, "d": d,
            "by_streams": {str(k): v for k, v in by_streams.items()}}


def bench_sv(H, cfg):
    """FP64 stochastic-volatility generator (ophip_dgen_sv) on the 256 x 8
    lattice: prices on [0, 400], regimes 0.1..0.8, mean-reverting regime
    drift, vol of vol 0.5, jumps of 10% / 5% of the price; bytes = the d x d
    FP64 output"""
    nx, nr = cfg.nx, cfg.nr
    from oplib_amd import opmodel as M
    x, _ = M.grid(cfg)
    d = nx * nr
    xv = np.tile(x, nr)
    rv = np.repeat(np.linspace(0.1, 0.8, nr), nx)
    arrs = [xv, rv, cfg.rate * xv, rv * xv, 0.5 * np.ones(d), 2.0 * (0.45 - rv), 0.1 * np.ones(d),
            0.05 * np.ones(d)]
    hd = [H.DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float64)) for a in arrs]
    out = H.DeviceArray.empty(d * d, dtype=np.float64)

    def run():
        H.lib.ophip_dgen_sv(nx, nr, *[a.h for a in hd], 0.0, 4, out.h)
    run()
    H.check()
    us = timed_events(H, run, 10) * 1e3
    gbs = 8.0 * d * d / (us * 1e-6) / 1e9
    log(f"[bench] SV generator d={d}: {us:.1f} us, {gbs:.0f} GB/s written")
    return {"kernel": "k_sv_rows + k_sv_fill<0, double>", "d": d, "us": us, "GBps_written": gbs}


def bench_surrogate(H, M, ntrain, steps, log_fn):
    """Learned surrogate with fallback on the CEV-with-skew family: labels
    from the mixed-precision batched solver, a 5-member ensemble trained on
    the GPU, the spread threshold calibrated on separate draws, judged on
    4096 held-out draws, then 2^16 inference draws priced with fallback."""
    from oplib_amd import surrogate as S
    out = {"workload": "1024-node CEV lattice with skew/curvature, 20 squarings to U(0.1), "
                       "21 strikes x 10 maturities; params (sigma0, beta, r, a, b) uniform on the box",
           "members": 5, "train_draws": ntrain, "train_steps": steps}
    sol = S.PriceSolver(nb=64)
    t0 = time.perf_counter()
    ptr = S.draw_params(ntrain, S.SEED_TRAIN, S.TRAIN_MARGIN)
    ytr = sol.solve(ptr)
    t_lab = time.perf_counter() - t0
    out["label_solves_per_s"] = ntrain / t_lab
    log_fn(f"[bench] surrogate labels: {ntrain} solves in {t_lab:.1f} s")
    pho = S.draw_params(4096, S.SEED_HELDOUT)
    yho = sol.solve(pho)
    pcal = S.draw_params(8192, S.SEED_CAL)
    ycal = sol.solve(pcal)
    net = S.Surrogate(members=5, cap=4096, seed=7)
    net.fit(ptr, ytr, steps=steps, batch=4096, lr=3e-3,
            lr_end=1.5e-5)
    out["train_seconds"] = net.train_seconds
    log_fn(f"[bench] surrogate trained in {net.train_seconds:.1f} s")
    sstar = net.calibrate(pcal, ycal, S.TOL)
    out["spread_threshold"] = sstar
    # held-out judgement against the solver
    pr, sp, fl = net.evaluate(pho, tol=sstar)
    e = S.rel_err(pr, yho)
    served = fl == 0
    e_cv = S.rel_err(S.control_variate(pho), yho)
    out["heldout"] = {
        "draws": len(pho), "served_fraction": float(served.mean()),
        "served_max_rel_err": float(e[served].max()) if served.any() else None,
        "served_over_1e-4": int((e[served] > S.TOL).sum()),
        "network_median_rel_err": float(np.median(e)), "network_max_rel_err": float(e.max()),
        "control_variate_median_rel_err": float(np.median(e_cv)),
        "judge": "the FP16x3 W-form solver (within 1e-4 of FP64 on the committed fixture)"}
    # inference: 2^16 resident draws, surrogate + fallback solves
    n = 1 << 16
    pin = S.draw_params(n, S.SEED_INFER)
    prm = H.DeviceArray.from_numpy(pin.astype(np.float32), order="C")
    prd, spd = H.DeviceArray.empty(n * S.NOUT), H.DeviceArray.empty(n)
    fld = H.DeviceArray.empty(n, dtype=np.int32)
    net.eval_into(n, prm, prd, spd, fld, sstar)
    H.sync()
    t0 = time.perf_counter()
    net.eval_into(n, prm, prd, spd, fld, sstar)
    H.sync()
    t_eval = time.perf_counter() - t0
    flags = fld.download()
    bad = np.nonzero(flags)[0]
    t0 = time.perf_counter()
    if len(bad):
        sol.solve(pin[bad])
    t_fb = time.perf_counter() - t0
    rate = n / (t_eval + t_fb)
    out["inference"] = {"draws": n, "eval_ms": t_eval * 1e3, "network_draws_per_s": n / t_eval,
                        "fallback_draws": int(len(bad)), "fallback_seconds": t_fb,
                        "scenarios_per_s": rate}
    # the methods without the network on the same family
    mixed = len(bad) / t_fb if len(bad) >= 64 else None
    sol.close()
    plain = S.PriceSolver(nb=64, precision=M.OPM_PREC_FP32)
    plain.solve(pin[:64])
    t0 = time.perf_counter()
    pp = plain.solve(pin[:256])
    t_pl = time.perf_counter() - t0
    plain.close()
    out["plain_fp32_scenarios_per_s"] = 256 / t_pl
    out["mixed_batched_scenarios_per_s"] = mixed if mixed else out["label_solves_per_s"]
    out["speedup_vs_plain"] = rate / out["plain_fp32_scenarios_per_s"]
    net.close()
    log_fn(f"[bench] surrogate: {out}")
    return out


def cpu_info():
    """the host's CPU model and thread counts"""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"model": model, "os_cpu_count": os.cpu_count(), "threads_allowed": aff,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def bench_cpu(cfg, U_daily=None):
    from oracle import oracle as O
    # the host threads this process may use: OMP_NUM_THREADS where it is set
    # (the GPU's share of the host's cores), else every thread of the
    # affinity mask (the reference's Environment.ProcessorCount,
    # cmt/Program.cs:55)
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or aff
    O.set_threads(nth)
    t0 = time.perf_counter()
    k = 0
    while True:  # whole solves until about 10 s (at most 8)
        O.solve_fp64(cfg)
        k += 1
        t = time.perf_counter() - t0
        if t >= 10.0 or k >= 8:
            break
    log(f"[bench] CPU FP64 oracle: {k} solves in {t:.2f} s on {nth} threads")
    out = {"value": k / t, "unit": "scenarios/s", "cores": nth, "kind": "port",
           "sample": f"{k} full FP64 solves of the same scenario (assembly, first step, "
                     "16 DGEMM squarings at d=2048, 252 backward steps x 41 strikes) by the "
                     "C oracle with OpenMP", "seconds": t, "host": cpu_info()}
    O.set_threads(1)
    t0 = time.perf_counter()
    O.solve_fp64(cfg)
    t1 = time.perf_counter() - t0
    O.set_threads(nth)
    out["single_thread"] = {"value": 1.0 / t1, "unit": "scenarios/s", "cores": 1,
                            "sample": "one full FP64 solve of the same scenario on one thread"}
    log(f"[bench] CPU FP64 oracle, one thread: {t1:.2f} s per solve")
    out["components"] = cpu_components(O, nth, U_daily)
    out["components"]["clv"] = cpu_clv(O, nth)
    out["components"]["csv"] = cpu_csv(O, nth)
    return out


def plan_days(step_days, gen_days):
    """the reference's step plan (CModel.make_mc_plan, CModel Common.cs:163-
    644) for steps ending at step_days (days) and generators valid up to
    gen_days: each step cut at the generator dates into pieces (generator,
    days), the pieces deduplicated; returns the pieces and each step's chain
    of piece indices"""
    pieces, chains, index = [], [], {}
    prev = 0
    for t in step_days:
        chain, a = [], prev
        while a < t:
            i = next(j for j, g in enumerate(gen_days) if g > a)
            b = min(t, gen_days[i])
            key = (i, b - a)
            if key not in index:
                index[key] = len(pieces)
                pieces.append(key)
            chain.append(index[key])
            a = b
        chains.append(chain)
        prev = t
    return pieces, chains


def cpu_plan_exec(O, gens, pieces, chains):
    """the plan in FP64 on the host (CModel Host Double.cs:89-305): per piece
    dt = min(0.5/max|L_yy|, 1 day), niter = ceil(log2(days/dt)) (the
    oracle's planner formula), U0 = I + L dt_j, niter DGEMM squarings; per
    step the product of its chain; returns the step kernels and the DGEMM
    flop count (2 d^3 per product)"""
    import ctypes as C
    d = gens[0].shape[0]
    flop, Uj = 0.0, []
    for i, days in pieces:
        L = gens[i]
        dtj = C.c_double()
        niter = O.lib.or_plan_niter(days / 365.25, float(np.abs(np.diag(L)).max()), 1.0 / 365.25,
                                    C.byref(dtj))
        Uj.append(O.square_chain(O.first_step(L, dtj.value), niter))
        flop += niter * 2.0 * d ** 3
    steps = []
    for ch in chains:
        U = Uj[ch[0]]
        for j in ch[1:]:
            U = O.dgemm(U, Uj[j])
            flop += 2.0 * d ** 3
        steps.append(U)
    return steps, flop


def cpu_mc_hash(O, steps, y0, nblocks, nscen, chunk=8):
    """the reference's CPU sampler (hash-bracket tables, CModel Host
    Double.cs:403-458; rangen.cpp:96-153) over the step kernels: table build
    and nblocks x nscen paths x len(steps) steps, in chunks of blocks;
    returns (build seconds, sampling seconds)"""
    nm, d = len(steps), steps[0].shape[0]
    t = np.ascontiguousarray(np.stack([U.reshape(-1, order="F") for U in steps]).reshape(-1))
    ct = np.zeros_like(t)
    hs = np.zeros(nm * d * d, dtype=np.int32)
    t0 = time.perf_counter()
    O.lib.or_mc_hash_build(d, nm, t, ct, hs)
    tb = time.perf_counter() - t0
    mk = np.arange(nm, dtype=np.int32)
    ys = np.zeros(chunk * nm * nscen, dtype=np.int16)
    ts = 0.0
    for b0 in range(0, nblocks, chunk):
        nb = min(chunk, nblocks - b0)
        t0 = time.perf_counter()
        O.lib.or_mc_hash_run(d, ct, hs, nb, nscen, nm, y0, mk, 5489 + b0, ys)
        ts += time.perf_counter() - t0
    return tb, ts


def cpu_clv(O, nth):
    """clv (benchmarks/clv/Program.cs:46-133) restated: local vol on 128
    nodes, 10 generators valid to 30, 60, ..., 300 days (r = 0.05,
    sigma = 0.25 x, CLVModel.cs:86-136 in FP64), the plan of 40 weekly steps
    in FP64 (first steps, DGEMM squarings, chain products), then 4096*25
    paths x 120 batches x 40 steps through the hash sampler.  The
    reference's piecewise-uniform pivot grid is replaced by a uniform grid
    on [0, 400] of the same size (the same work)."""
    x, x0 = O.xgrid(128, 0.0, 400.0, 100.0)
    gens = [O.lv_generator(x, 0.05 * x, 0.25 * x) for _ in range(10)]
    pieces, chains = plan_days([7 * (k + 1) for k in range(40)], [30 * (i + 1) for i in range(10)])
    tp = float("inf")
    for _ in range(3):  # best of three: the plan's products are small, thread start-up shows
        t0 = time.perf_counter()
        steps, flop = cpu_plan_exec(O, gens, pieces, chains)
        tp = min(tp, time.perf_counter() - t0)
    nscen, nb = 4096 * 25, 120
    tb, ts = cpu_mc_hash(O, steps, x0, nb, nscen)
    O.set_threads(1)
    _, ts1 = cpu_mc_hash(O, steps, x0, 2, nscen, chunk=2)
    O.set_threads(nth)
    out = {"gflops": flop / tp / 1e9, "plan_seconds": tp, "kernels": len(pieces),
           "mc_M_eval_per_s": 40.0 * nb * nscen / ts / 1e6, "threads": nth,
           "mc_M_eval_per_s_one_thread": 40.0 * 2 * nscen / ts1 / 1e6, "table_build_s": tb,
           "sample": "LV d=128, 10 monthly generators, 40 weekly steps (FP64 DGEMM squarings and "
                     "chains), 4096*25 paths x 120 batches x 40 steps, hash-bracket sampler; one "
                     "thread: 2 batches"}
    log(f"[bench] CPU clv: {out['gflops']:.1f} GF/s, {out['mc_M_eval_per_s']:.0f} M eval/s "
        f"({out['mc_M_eval_per_s_one_thread']:.0f} on one thread)")
    return out


def cpu_csv(O, nth):
    """csv (benchmarks/csv/Program.cs:48-147) restated: stochastic vol on 64
    price nodes x 8 vol regimes (0.55 .. 2.0, spot regime 4), 10 generators
    from the term structure (vol 0.25, beta 0.5, volvol 0.5, mean reversion
    0.5 to 3, jump sizes -3 / 0: CSVModel.mkgen, CSVModel.cs:103-172; the
    generator CSVModel.host_d_gen, CSVModel.cs:187-379, in FP64), the plan of
    40 weekly steps in FP64, then 4096*25 paths x 100 batches x 40 steps
    through the hash sampler.  Uniform price grid on [0, 400] of the same
    size in place of the pivot grid."""
    nx, nr = 64, 8
    x, x0 = O.xgrid(nx, 0.0, 400.0, 100.0)
    rval = np.array([0.55, 0.75, 0.90, 1.0, 1.1, 1.3, 1.6, 2.0])
    y0 = x0 + nx * 4
    tau = np.array([[v, v, v, 0.0] for v in (0.25, 0.5, 0.5, 0.5, 0.5, 3.0, -3.0, 0.0)])
    gen_days = [30 * (i + 1) for i in range(10)]
    svol, volvol, voldrift, jm, jp = O.sv_params(x, rval, y0, [g / 365.25 for g in gen_days], tau)
    xs, rs = np.tile(x, nr), np.repeat(rval, nx)
    gens = [O.sv_generator(nx, nr, xs, rs, 0.05 * xs, svol[i], volvol[i], voldrift[i], jm[i], jp[i])
            for i in range(10)]
    pieces, chains = plan_days([7 * (k + 1) for k in range(40)], gen_days)
    tp = float("inf")
    for _ in range(3):  # best of three: the plan's products are small, thread start-up shows
        t0 = time.perf_counter()
        steps, flop = cpu_plan_exec(O, gens, pieces, chains)
        tp = min(tp, time.perf_counter() - t0)
    nscen, nb = 4096 * 25, 100
    tb, ts = cpu_mc_hash(O, steps, y0, nb, nscen)
    out = {"gflops": flop / tp / 1e9, "plan_seconds": tp, "kernels": len(pieces),
           "mc_M_eval_per_s": 40.0 * nb * nscen / ts / 1e6, "threads": nth, "table_build_s": tb,
           "sample": "SV 64 x 8 = 512 states, 10 monthly generators, 40 weekly steps (FP64 DGEMM "
                     "squarings and chains), 4096*25 paths x 100 batches x 40 steps, hash-bracket "
                     "sampler"}
    log(f"[bench] CPU csv: {out['gflops']:.1f} GF/s, {out['mc_M_eval_per_s']:.0f} M eval/s")
    return out


def cpu_components(O, nth, U_daily):
    """the reference's CPU benchmark programs restated, on the host cores:
    cmt (MT uniforms), csgemm (SGEMM d = 576 x 100, the host BLAS) and the
    clv-style chain sampler (bracket tables + bisection, rangen.cpp:96-153)
    on the daily 2048-state kernel"""
    import ctypes as C
    import threading
    comp = {}
    # cmt: 2^20 uniforms per thread, one MT19937 stream each
    n = 1 << 20
    bufs = [np.zeros(n, dtype=np.float32) for _ in range(nth)]
    ths = [threading.Thread(target=O.lib.or_mt_uniforms, args=(5489 + i, n, bufs[i])) for i in range(nth)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    t = time.perf_counter() - t0
    comp["cmt"] = {"value": nth * n / t / 1e6, "unit": "M uniforms/s", "threads": nth,
                   "sample": f"{nth} x 2^20 MT19937 uniforms"}
    # csgemm: d = 576 SGEMM x 100 on the host BLAS (numpy)
    a = np.random.default_rng(0).standard_normal((576, 576)).astype(np.float32)
    b = a.T.copy()
    a @ b
    t0 = time.perf_counter()
    for _ in range(100):
        c = a @ b
    t = time.perf_counter() - t0
    comp["csgemm"] = {"value": 100 * 2.0 * 576 ** 3 / t / 1e9, "unit": "GFLOP/s",
                      "sample": "SGEMM d=576 x 100 through numpy's BLAS"}
    # clv-style sampler on the daily kernel: 64 blocks x 2048 paths x 252 steps
    if U_daily is not None:
        d = U_daily.shape[0]
        t0 = time.perf_counter()
        y, ct, hs = O.mc_hash(U_daily[None].astype(np.float64), 1, 64, 64 + 256, [0], 1)
        tb = time.perf_counter() - t0
        mk = np.zeros(252, dtype=np.int32)
        nb, ns = 64, 2048
        ys = np.zeros(nb * 252 * ns, dtype=np.int16)
        t0 = time.perf_counter()
        O.lib.or_mc_hash_run(d, ct, hs, nb, ns, 252, 64 + 256, mk, 5489, ys)
        t = time.perf_counter() - t0
        comp["clv_mc"] = {"value": nb * ns * 252 / t / 1e6, "unit": "M transitions/s", "threads": nth,
                          "table_build_s": tb,
                          "sample": f"{nb} x {ns} paths x 252 steps on the daily 2048-state kernel, "
                                    "bracket tables + bisection, MT19937 per path block"}
    log(f"[bench] CPU components: {comp}")
    return comp


if __name__ == "__main__":
    main()
