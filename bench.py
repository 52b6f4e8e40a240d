#!/usr/bin/env python3
"""Throughput benchmark: DQMC sweeps/s (all chains) of the batched Z2-gauge engine.

One step = one sweep (left + right pass, 4*N*Ltrot link-flip proposals per
chain) of every chain of the workload, inputs resident in HBM.  Default workload
(the configuration the metric is quoted on): 64 chains, L=16 (N=256), beta=20,
Ltrot=200, Nwrap=10, h in {0.4,0.6,0.8,1.0} x 16, J=1, t=1, mu=0.

Prints one JSON line.  `roofline` times the dominant kernel live with HIP events
on the engine's stream; `cpu_baseline` times the CPU oracle (plain-C FP64
restatement of the reference path) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DQMC sweeps/s (all chains) at L=16,β=20: mixed-prec vs FP64; max|ΔG|"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

CONFIGS = {
    # name: (L, Ltrot, beta, Nwrap, n_chains, h values, mu)
    "c4": (16, 200, 20.0, 10, 64, (0.4, 0.6, 0.8, 1.0), 0.0),
    "c2": (12, 120, 12.0, 10, 1, (0.8,), 0.0),
    "c1": (4, 40, 4.0, 10, 1, (0.5,), 0.0),
}


def chain_specs(name):
    import z2_gauge_zfy_amd as z
    L, T, beta, w, nch, hs, mu = CONFIGS[name]
    per = max(1, nch // len(hs))
    return [z.ChainSpec(RT=1.0, J=1.0, h=hs[min(k // per, len(hs) - 1)], mu=mu, Beta=beta,
                        seed=12345 + 2 * 7919 * k, absolute=1) for k in range(nch)]


def roofline(kind, N, kms, klaunch, work, precision, traffic=None):
    """achieved rate of one kernel class: its algorithmic work (flops, or bytes for
    the memory-bound classes; the flip kernel: 4 N^2 flops per accepted update)
    over its summed HIP-event time"""
    bound = "hbm" if kind in ("kin_rows", "kin_cols", "obs") else "mfma"
    avg_s = kms / 1e3 / max(1, klaunch)
    per_launch = work / max(1, klaunch)
    if bound == "hbm":
        achieved, peak, unit = per_launch / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        peak = FP32_PEAK_TFLOPS if (precision == "mixed" and kind == "flip") else FP64_PEAK_TFLOPS
        achieved, unit = per_launch / avg_s / 1e12, "TFLOP/s"
    return {"kernel": kind, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic, "avg_launch_us": avg_s * 1e6,
            "launches": klaunch, "work_per_launch": per_launch}


def step_roofline(N, T, w, nch, precision, ms_per_step):
    """whole-step bound: the checkerboard wraps at HBM speed (G two-sided and
    U one-sided, read + write every slice step: 8 T N^2 s bytes per chain-sweep)
    plus the stabilization at the FP64 peak (about 12 N^3 flops per recompute,
    2 Ltrot/Nwrap recomputes per sweep), against the measured time per step"""
    s = 4 if precision == "mixed" else 8
    wrap_bytes = 8.0 * T * N * N * s * nch
    stab_flops = 12.0 * N ** 3 * 2 * (T // w) * nch
    t_roof = wrap_bytes / (HBM_PEAK_GBS * 1e9) + stab_flops / (FP64_PEAK_TFLOPS * 1e12)
    return {"wrap_bytes": wrap_bytes, "stab_flops": stab_flops, "t_roofline_ms": t_roof * 1e3,
            "ms_per_step": ms_per_step, "frac": t_roof * 1e3 / ms_per_step}


def measured_traffic(kind, cfg, precision):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any"""
    path = os.path.join(ROOT, "profiles", f"traffic_{kind}_{cfg}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def _cpu_worker(args):
    cfg, k, sigma, seed = args
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    L, T, beta, w, nch, hs, mu = CONFIGS[cfg]
    spec = chain_specs(cfg)[k]
    p = O.make_params(RT=spec.RT, J=spec.J, h=spec.h, mu=spec.mu, Nlx=L, Nly=L, Ltrot=T, Beta=beta,
                      Nwrap=w)
    m = O.Model(p)
    c = O.Chain(m, seed, sigma_therm=sigma)
    c.seed = seed
    c.pre()
    t0 = time.perf_counter()
    c.sweep()
    return time.perf_counter() - t0


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baseline(cfg, snapshot):
    """The oracle (plain-C FP64 restatement of the reference path, one chain per
    single-threaded process as the reference runs one chain per MPI rank) on the
    GPU's state at the start of the timed region: each process takes one chain's
    field and seed from the FP64 GPU run after its warm-up sweeps, rebuilds the
    UDV checkpoints and G (pre), and times one full sweep."""
    model, ncpu, avail = host_cpu()
    P = max(1, min(avail, 16, len(snapshot)))
    with mp.get_context("spawn").Pool(P) as pool:
        times = pool.map(_cpu_worker, [(cfg, k, snapshot[k][0], snapshot[k][1]) for k in range(P)])
    wall = max(times)
    return {"value": P / wall, "unit": "sweeps/s", "cores": P, "kind": "port",
            "host": {"cpu_model": model, "nproc": ncpu, "available": avail},
            "sample": f"oracle (plain C FP64), {P} single-threaded processes (one chain each, chains "
                      f"0..{P - 1} of the workload), each starting from its chain's field and seed as the "
                      f"FP64 GPU run held them after its warm-up sweeps (pre() rebuilds the checkpoints "
                      f"and G), one full sweep timed; slowest process {wall:.2f} s"}


PROGRESS = False


def progress(msg):
    if PROGRESS:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_mode(cfg, precision, steps, warmup, roofline_kernel, delay, extra=(), snapshot=0, specs=None):
    """Time `steps` sweeps of all chains after `warmup` sweeps (the kernel class
    `roofline_kernel` bracketed by HIP events inside the timed region); then one
    more sweep per class in `extra` for their own event timing.  snapshot > 0
    keeps the field and seed of the first `snapshot` chains as they are when the
    timed region starts (the CPU baseline starts from that state)."""
    import z2_gauge_zfy_amd as z
    L, T, beta, w, nch, hs, mu = CONFIGS[cfg]
    N = L * L
    specs = specs if specs is not None else chain_specs(cfg)
    nch = len(specs)
    eng = z.Engine(L, L, T, w, specs, precision=precision)
    if delay:
        eng.set_delay(delay)
    eng.init_fields()
    eng.pre()
    for i in range(warmup):
        eng.sweep(True)
        eng.sync()
        progress(f"{precision} ({nch} chains) warm-up sweep {i + 1}/{warmup} done")
    eng.sync()
    snap = [(eng.sigma(k).copy(), eng.seed(k)) for k in range(min(snapshot, nch))]
    acc0 = sum(eng.counters(k)[0] for k in range(nch))
    prop0 = sum(eng.counters(k)[1] for k in range(nch))
    if roofline_kernel:
        eng.profile_kernel(roofline_kernel)
    t0 = time.perf_counter()
    for i in range(steps):
        eng.sweep(True)
        if PROGRESS:
            eng.sync()
            progress(f"{precision} ({nch} chains) timed sweep {i + 1}/{steps} done")
    eng.sync()
    dt = time.perf_counter() - t0
    kms, klaunch = eng.profile_read()
    kwork = eng.profile_work()
    eng.profile_kernel("")
    acc1 = sum(eng.counters(k)[0] for k in range(nch))
    prop1 = sum(eng.counters(k)[1] for k in range(nch))
    if roofline_kernel == "flip":
        kwork = 4.0 * N * N * (acc1 - acc0)
    others = {}
    for kind in extra:
        a0 = sum(eng.counters(k)[0] for k in range(nch))
        eng.profile_kernel(kind)
        t1 = time.perf_counter()
        eng.sweep(True)
        eng.sync()
        t_sw = time.perf_counter() - t1
        ms, nl = eng.profile_read()
        wk = eng.profile_work()
        eng.profile_kernel("")
        a1 = sum(eng.counters(k)[0] for k in range(nch))
        if kind == "flip":
            wk = 4.0 * N * N * (a1 - a0)
        others[kind] = (ms, nl, wk, ms / 1e3 / t_sw)
        progress(f"{precision} extra sweep timing {kind} done")
    eng.close()
    return dict(dt=dt, sweeps_per_s=nch * steps / dt, flips_per_s=(prop1 - prop0) / dt,
                acc_per_s=(acc1 - acc0) / dt, acceptance=(acc1 - acc0) / max(1, prop1 - prop0),
                kms=kms, klaunch=klaunch, kwork=kwork, accepted=acc1 - acc0, others=others, snapshot=snap)


def shadow_max_dG(cfg, sweeps, warmup, delay):
    """max |G_mixed - G_fp64| over chains and over the propagated G just before
    every recompute, FP32 run replaying the FP64 run's decisions."""
    import numpy as np
    import z2_gauge_zfy_amd as z
    L, T, beta, w, nch, hs, mu = CONFIGS[cfg]
    ref = z.Engine(L, L, T, w, chain_specs(cfg), precision=z.Z2_PREC_FP64)
    mix = z.Engine(L, L, T, w, chain_specs(cfg), precision=z.Z2_PREC_MIXED)
    for e in (ref, mix):
        if delay:
            e.set_delay(delay)
        e.init_fields()
        e.pre()
    d = np.zeros(nch)
    for _ in range(warmup + sweeps):
        d = np.maximum(d, mix.sweep_shadow(ref, True))
    obs = {}
    for name in ("density", "pair_sc_q0", "cdw_pi", "flux_avg", "chi_flux"):
        obs[name] = max(abs(mix.obs_equal(k)[name] - ref.obs_equal(k)[name]) for k in range(nch))
    ref.close()
    mix.close()
    return float(d.max()), obs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="both", choices=("both", "mixed", "fp64"))
    ap.add_argument("--roofline-kernel", default="flip")
    ap.add_argument("--extra-kernels", default="qrp,orgqr,qt_apply,kin_rows,kin_cols,obs",
                    help="kernel classes timed in one extra sweep each (roofline_other)")
    ap.add_argument("--delay", type=int, default=0, help="delay depth k (0: engine default)")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="timed sweeps of the plain FP64 form (one chain, no batching)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-shadow", action="store_true")
    ap.add_argument("--progress", action="store_true", help="one stderr line per sweep")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first run this many launches of a plain 64 MiB copy kernel (k_copy), "
                         "the reference for scaling rocprofv3 byte counters")
    a = ap.parse_args()
    global PROGRESS
    PROGRESS = a.progress
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE={world}); "
              "chains are independent, so N GPUs would each run their own chain subset.",
              file=sys.stderr)
        sys.exit(2)

    import z2_gauge_zfy_amd as z
    L, T, beta, w, nch, hs, mu = CONFIGS[a.config]
    N = L * L
    if a.copy_calibration:
        z._lib.check(z.lib().z2_test_copy(8 << 20, a.copy_calibration))
    extra = [k for k in a.extra_kernels.split(",") if k and k != "none"]
    _, _, avail = host_cpu()
    res = {}
    if a.precision in ("both", "fp64"):
        res["fp64"] = run_mode(a.config, z.Z2_PREC_FP64, a.steps, a.warmup, a.roofline_kernel, a.delay, extra,
                               snapshot=0 if a.no_cpu else min(avail, 16, nch))
    if a.precision in ("both", "mixed"):
        res["mixed"] = run_mode(a.config, z.Z2_PREC_MIXED, a.steps, a.warmup, a.roofline_kernel, a.delay, extra)
    head = res["mixed"] if "mixed" in res else res["fp64"]
    head_name = "mixed" if "mixed" in res else "fp64"
    out = {
        "metric": METRIC,
        "value": head["sweeps_per_s"],
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["dt"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_name == "mixed" else "f64",
        "data": "synthetic (seeded random initial links, absolute=1)",
        "config": {"workload": f"{a.config}: {nch} chains, L={L} (N={N}), beta={beta}, Ltrot={T}, "
                               f"Nwrap={w}, h in {list(hs)}, J=1, t=1, mu={mu}",
                   "n_chains": nch, "L": L, "Ltrot": T, "beta": beta, "Nwrap": w,
                   "precision": head_name},
        "flip_updates_per_s": head["flips_per_s"],
        "accepted_updates_per_s": head["acc_per_s"],
        "acceptance": head["acceptance"],
    }
    for k, r in res.items():
        out[f"{k}_sweeps_per_s"] = r["sweeps_per_s"]
        out[f"{k}_ms_per_step"] = r["dt"] / a.steps * 1e3
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = res["mixed"]["sweeps_per_s"] / res["fp64"]["sweeps_per_s"]
    if a.plain_steps > 0:
        # the plain FP64 GPU form: one chain (the reference's one chain per rank) at the
        # same lattice, no batching -- the denominator of the batched-mixed speed-up
        plain = run_mode(a.config, z.Z2_PREC_FP64, a.plain_steps, 1, "", a.delay, (),
                         specs=chain_specs(a.config)[:1])
        out["plain_fp64_1chain_sweeps_per_s"] = plain["sweeps_per_s"]
        out["batched_mixed_over_plain_fp64"] = (res["mixed"]["sweeps_per_s"] / plain["sweeps_per_s"]
                                                if "mixed" in res else None)
        out["batched_fp64_over_plain_fp64"] = (res["fp64"]["sweeps_per_s"] / plain["sweeps_per_s"]
                                               if "fp64" in res else None)
    rf = roofline(a.roofline_kernel, N, head["kms"], head["klaunch"], head["kwork"], head_name,
                  measured_traffic(a.roofline_kernel, a.config, head_name))
    rf["kernel_share_of_step"] = (head["kms"] / 1e3) / head["dt"]
    out["roofline"] = rf
    out["roofline_step"] = {k: step_roofline(N, T, w, nch, k, r["dt"] / a.steps * 1e3) for k, r in res.items()}
    for k, r in res.items():
        other = []
        for kind, (ms, nl, wk, share) in r["others"].items():
            ro = roofline(kind, N, ms, nl, wk, k, measured_traffic(kind, a.config, k))
            ro["kernel_share_of_sweep"] = share
            other.append(ro)
        out["roofline_other" if k == head_name else f"roofline_other_{k}"] = other
    if a.precision == "both" and not a.no_shadow:
        dg, dobs = shadow_max_dG(a.config, 1, a.warmup, a.delay)
        out["max_dG"] = dg
        out["max_dobs"] = dobs
        progress("shadow sweep done")
    if not a.no_cpu and "fp64" in res:
        out["cpu_baseline"] = cpu_baseline(a.config, res["fp64"]["snapshot"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
