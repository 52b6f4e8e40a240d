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

# the workloads: (chains as (L, Ltrot, Nwrap, h, beta, mu) per chain, precision, groups,
# autotune drift tolerance (0: no tuning), time-displaced cadence (0: none))
def _grid(L, T, w, beta, mu, hs, nch):
    per = max(1, nch // len(hs))
    return [(L, T, w, hs[min(k // per, len(hs) - 1)], beta, mu) for k in range(nch)]


def _c3():
    hs = [0.1 + 1.4 * i / 31 for i in range(32)]
    out, k = [], 0
    for L, n in ((8, 86), (10, 85), (12, 85)):
        for _ in range(n):
            out.append((L, 100, 10, hs[k % 32], 10.0, 0.0))
            k += 1
    return out


CONFIGS = {
    "c4": dict(chains=_grid(16, 200, 10, 20.0, 0.0, (0.4, 0.6, 0.8, 1.0), 64), precision="both", groups=3,
               extra_groups=1,
               desc="64 chains, L=16 (N=256), beta=20, Ltrot=200, Nwrap=10, h in [0.4, 0.6, 0.8, 1.0] x 16, "
                    "J=1, t=1, mu=0"),
    "c3": dict(chains=_c3(), precision="fp64", groups=1, tune=5e-3,
               desc="256 chains, L in {8, 10, 12} (86/85/85), beta=10, Ltrot=100, h on 32 values over "
                    "[0.1, 1.5], J=1, t=1, mu=0; delay depth, tiles, Nwrap in {5, 10, 20} and execution "
                    "mode autotuned"),
    "c5": dict(chains=_grid(20, 400, 8, 40.0, -0.3, (0.6,), 32), precision="mixed", groups=1, tau_every=4,
               desc="32 chains, L=20 (N=400), beta=40, Ltrot=400, Nwrap=8, h=0.6, J=1, t=1, mu=-0.3; "
                    "time-displaced measurement every 4 sweeps"),
    "c2": dict(chains=_grid(12, 120, 10, 12.0, 0.0, (0.8,), 1), precision="fp64", groups=1,
               desc="1 chain, L=12 (N=144), beta=12, Ltrot=120, Nwrap=10, h=0.8, J=1, t=1, mu=0"),
    "c1": dict(chains=_grid(4, 40, 10, 4.0, 0.0, (0.5,), 1), precision="fp64", groups=1,
               desc="1 chain, L=4 (N=16), beta=4, Ltrot=40, Nwrap=10, h=0.5, J=1, t=1, mu=0"),
}


def batch_chains(cfg, first=None):
    """the workload's chains: seed 12345 + 2*7919*k for chain k, random initial links (absolute=1)"""
    import z2_gauge_zfy_amd as z
    ch = CONFIGS[cfg]["chains"]
    ks = range(len(ch)) if first is None else range(first)
    return [z.BatchChain(ch[k][0], ch[k][1], ch[k][2],
                         z.ChainSpec(RT=1.0, J=1.0, h=ch[k][3], mu=ch[k][5], Beta=ch[k][4],
                                     seed=12345 + 2 * 7919 * k, absolute=1)) for k in ks]


def chain_specs(cfg):
    return [c.spec for c in batch_chains(cfg)]


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


def step_roofline(shapes, precision, ms_per_step):
    """whole-step bound: the checkerboard wraps at HBM speed (G two-sided and
    U one-sided, read + write every slice step: 8 T N^2 s bytes per chain-sweep)
    plus the stabilization at the FP64 peak (about 12 N^3 flops per recompute,
    2 Ltrot/Nwrap recomputes per sweep), summed over the chains ((N, Ltrot, Nwrap)
    each), against the measured time per step"""
    s = 4 if precision == "mixed" else 8
    wrap_bytes = sum(8.0 * T * N * N * s for N, T, w in shapes)
    stab_flops = sum(12.0 * N ** 3 * 2 * (T // w) for N, T, w in shapes)
    t_roof = wrap_bytes / (HBM_PEAK_GBS * 1e9) + stab_flops / (FP64_PEAK_TFLOPS * 1e12)
    return {"wrap_bytes": wrap_bytes, "stab_flops": stab_flops, "t_roofline_ms": t_roof * 1e3,
            "ms_per_step": ms_per_step, "frac": t_roof * 1e3 / ms_per_step}


def measured_traffic(kind, cfg, precision, chains_per_launch):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any,
    scaled from the chains per launch it was measured at to chains_per_launch"""
    path = os.path.join(ROOT, "profiles", f"traffic_{kind}_{cfg}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        j = json.load(f)
    b, at = j.get("bytes_per_launch"), j.get("chains_per_launch")
    return b * chains_per_launch / at if (b is not None and at) else b


def _cpu_worker(args):
    cfg, k, sigma, seed = args
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    L, T, w, h, beta, mu = CONFIGS[cfg]["chains"][k]
    p = O.make_params(RT=1.0, J=1.0, h=h, mu=mu, Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=w)
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
    field and seed from the GPU run (FP64 when it ran) after its warm-up sweeps, rebuilds the
    UDV checkpoints and G (pre), and times one full sweep.  snapshot: {chain: (sigma, seed)}."""
    model, ncpu, avail = host_cpu()
    ks = sorted(snapshot)
    P = max(1, min(avail, 16, len(ks)))
    ks = ks[:P]
    with mp.get_context("spawn").Pool(P) as pool:
        times = pool.map(_cpu_worker, [(cfg, k, snapshot[k][0], snapshot[k][1]) for k in ks])
    wall = max(times)
    # aggregate rate of P processes, each one chain: chains of different size are
    # weighted by the slowest, as a job of P ranks waits for its slowest rank
    return {"value": P / wall, "unit": "sweeps/s", "cores": P, "kind": "port",
            "host": {"cpu_model": model, "nproc": ncpu, "available": avail},
            "sample": f"oracle (plain C FP64), {P} single-threaded processes, one chain each (chains "
                      f"{ks[0]}..{ks[-1]} of the workload), each starting from its chain's field and seed as "
                      f"the FP64 GPU run held them after its warm-up sweeps (pre() rebuilds the checkpoints "
                      f"and G), one full sweep timed; slowest process {wall:.2f} s"}


PROGRESS = False


def progress(msg):
    if PROGRESS:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_tuning(path):
    """group lines of an autotuner file: (L, Ltrot, n_chains, Nwrap, delay, rows_cb, cols_rb), and the exec mode"""
    groups, ex = [], None
    with open(path) as f:
        for line in f:
            w = line.split()
            if w and w[0] == "group":
                lx, ly, T, prec, nch, nw, kd, cb, rb = map(int, w[1:10])
                groups.append(dict(L=lx, Ltrot=T, n_chains=nch, Nwrap=nw, delay=kd, rows_cb=cb, cols_rb=rb))
            elif w and w[0] == "exec":
                ex = "graph" if int(w[1]) == 1 else "eager"
    return groups, ex


def run_mode(cfg, precision, steps, warmup, roofline_kernel, delay, extra=(), snapshot=(), chains=None,
             groups=1, exec_mode="eager", tune=0.0, tau_every=0):
    """Time `steps` sweeps of all chains of the workload (a batch: one context per
    shape group, split into `groups` parts) after `warmup` sweeps.  The kernel
    class `roofline_kernel` is bracketed by HIP events inside the timed region
    (eager execution; with graph execution, in one more eager sweep); then one
    more sweep per class in `extra` for their own event timing.  `snapshot`: the
    chains whose field and seed are kept as they are when the timed region starts
    (the CPU baseline starts from that state)."""
    import z2_gauge_zfy_amd as z
    prec = z.Z2_PREC_MIXED if precision == "mixed" else z.Z2_PREC_FP64
    chains = chains if chains is not None else batch_chains(cfg)
    nch = len(chains)
    max_group = -(-nch // groups) if groups > 1 else 0
    b = z.Batch(chains, precision=prec, max_group=max_group)
    b.init_fields()
    ngroups = b.info()[1]
    tuning = None
    if tune > 0:
        path = os.path.join("/tmp", f"z2_tune_{cfg}_{precision}_{os.getpid()}.txt")
        t0 = time.perf_counter()
        b.autotune(warm=1, sweeps=1, drift_tol=tune, path=path)
        tl, ex = read_tuning(path)
        tuning = {"groups": tl, "exec": ex, "tuning_s": time.perf_counter() - t0}
        progress(f"{precision} autotuned in {tuning['tuning_s']:.1f} s: {tuning}")
    else:
        b.set_exec(z.Z2_EXEC_GRAPH if exec_mode == "graph" else z.Z2_EXEC_EAGER)
        if delay:
            for g in range(ngroups):
                b.group(g).set_delay(delay)
    if tau_every:
        b.enable_tau()
    b.pre()
    graph = b.exec_mode() == z.Z2_EXEC_GRAPH
    toggle = lambda i: bool(tau_every) and i % tau_every == 0
    for i in range(warmup):
        b.sweep(True, toggle(i))
        b.sync()
        progress(f"{precision} ({nch} chains) warm-up sweep {i + 1}/{warmup} done")
    b.sync()
    snap = {k: (b.sigma(k).copy(), b.seed(k)) for k in snapshot}
    grp_of = [b.where(k) for k in range(nch)]
    Ns = [c.L * c.L for c in chains]

    def accepted():
        return [b.counters(k) for k in range(nch)]

    def profile(kind):
        for g in range(ngroups):
            b.group(g).profile_kernel(kind)

    def spans_union():
        """(union of the timed launches' intervals over the groups, their sum) in ms"""
        ivs = []
        for g in range(ngroups):
            st, en = b.group(g).profile_spans(b.group(0))
            ivs.extend(zip(st.tolist(), en.tolist()))
        ivs.sort()
        union, cur0, cur1 = 0.0, None, None
        for s0, s1 in ivs:
            if cur1 is None or s0 > cur1:
                if cur1 is not None:
                    union += cur1 - cur0
                cur0, cur1 = s0, s1
            else:
                cur1 = max(cur1, s1)
        if cur1 is not None:
            union += cur1 - cur0
        return union, sum(e - s for s, e in ivs)

    def profile_read(kind, c0, c1):
        ms, nl, wk = 0.0, 0, 0.0
        span = spans_union() if ngroups > 1 else None
        for g in range(ngroups):
            e = b.group(g)
            t, n = e.profile_read()
            ms, nl, wk = ms + t, nl + n, wk + e.profile_work()
            e.profile_kernel("")
        if kind == "flip":  # 4 N^2 flops per accepted rank-2 update
            wk = sum(4.0 * Ns[k] ** 2 * (c1[k][0] - c0[k][0]) for k in range(nch))
        return ms, nl, wk, span

    c0 = accepted()
    live = roofline_kernel and not graph
    if live:
        profile(roofline_kernel)
    t0 = time.perf_counter()
    for i in range(steps):
        b.sweep(True, toggle(warmup + i))
        if PROGRESS:
            b.sync()
            progress(f"{precision} ({nch} chains) timed sweep {i + 1}/{steps} done")
    b.sync()
    dt = time.perf_counter() - t0
    c1 = accepted()
    kms = klaunch = kwork = 0
    kspan = None
    if live:
        kms, klaunch, kwork, kspan = profile_read(roofline_kernel, c0, c1)
    others = {}
    kinds = list(extra) + ([roofline_kernel] if roofline_kernel and graph else [])
    if kinds:
        b.set_exec(z.Z2_EXEC_EAGER)  # HIP events bracket eager launches
    for kind in kinds:
        a0 = accepted()
        profile(kind)
        t1 = time.perf_counter()
        b.sweep(True)
        b.sync()
        t_sw = time.perf_counter() - t1
        ms, nl, wk, _ = profile_read(kind, a0, accepted())
        others[kind] = (ms, nl, wk, ms / 1e3 / t_sw)
        progress(f"{precision} extra sweep timing {kind} done")
    if roofline_kernel and graph:
        kms, klaunch, kwork = others.pop(roofline_kernel)[:3]
    b.close()
    acc = sum(c1[k][0] - c0[k][0] for k in range(nch))
    prop = sum(c1[k][1] - c0[k][1] for k in range(nch))
    shapes = [(c.L * c.L, c.Ltrot, c.Nwrap) for c in chains]
    if tuning:
        nw = {(g["L"], g["Ltrot"]): g["Nwrap"] for g in tuning["groups"]}
        shapes = [(c.L * c.L, c.Ltrot, nw.get((c.L, c.Ltrot), c.Nwrap)) for c in chains]
    return dict(dt=dt, sweeps_per_s=nch * steps / dt, flips_per_s=prop / dt, acc_per_s=acc / dt,
                acceptance=acc / max(1, prop), kms=kms, klaunch=klaunch, kwork=kwork, kspan=kspan, accepted=acc,
                others=others, snapshot=snap, groups=ngroups, exec="graph" if graph else "eager",
                roofline_live=bool(live), tuning=tuning, shapes=shapes)


def shadow_max_dG(cfg, sweeps, warmup, delay):
    """max |G_mixed - G_fp64| over chains and over the propagated G just before
    every recompute of `sweeps` sweeps after `warmup` ones, the mixed run
    replaying the FP64 run's decisions (one shape only)."""
    import numpy as np
    import z2_gauge_zfy_amd as z
    ch = batch_chains(cfg)
    L, T, w = ch[0].L, ch[0].Ltrot, ch[0].Nwrap
    specs = [c.spec for c in ch]
    ref = z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_FP64)
    mix = z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_MIXED)
    for e in (ref, mix):
        if delay:
            e.set_delay(delay)
        e.init_fields()
        e.pre()
    for _ in range(warmup):  # the sweeps from the random start, as the timed runs' warm-up
        mix.sweep_shadow(ref, True)
    d = np.zeros(len(specs))
    for _ in range(sweeps):
        d = np.maximum(d, mix.sweep_shadow(ref, True))
    obs = {}
    for name in ("density", "pair_sc_q0", "cdw_pi", "flux_avg", "chi_flux"):
        obs[name] = max(abs(mix.obs_equal(k)[name] - ref.obs_equal(k)[name]) for k in range(len(specs)))
    ref.close()
    mix.close()
    return float(d.max()), obs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default=None, choices=("both", "mixed", "fp64"),
                    help="default: the workload's own (c4: both, headline mixed)")
    ap.add_argument("--groups", type=int, default=None,
                    help="contexts (HIP streams) the chains of one shape are split into (default: the workload's)")
    ap.add_argument("--exec", default="eager", choices=("eager", "graph"),
                    help="execution of untuned workloads: launches issued per step, or one captured graph per sweep")
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
    W = CONFIGS[a.config]
    precision = a.precision or W["precision"]
    groups = a.groups or W["groups"]
    chains = batch_chains(a.config)
    nch = len(chains)
    if a.copy_calibration:
        z._lib.check(z.lib().z2_test_copy(8 << 20, a.copy_calibration))
    extra = [k for k in a.extra_kernels.split(",") if k and k != "none"]
    _, _, avail = host_cpu()
    P = min(avail, 16, nch)
    snap_ks = sorted({round(i * (nch - 1) / max(1, P - 1)) for i in range(P)}) if not a.no_cpu else []
    common = dict(groups=groups, exec_mode=a.exec, tune=W.get("tune", 0.0), tau_every=W.get("tau_every", 0))
    res = {}
    if precision in ("both", "fp64"):
        res["fp64"] = run_mode(a.config, "fp64", a.steps, a.warmup, a.roofline_kernel, a.delay, extra,
                               snapshot=snap_ks, **common)
    if precision in ("both", "mixed"):
        res["mixed"] = run_mode(a.config, "mixed", a.steps, a.warmup, a.roofline_kernel, a.delay, extra,
                                snapshot=snap_ks if "fp64" not in res else (), **common)
    head_name = "mixed" if "mixed" in res else "fp64"
    head = res[head_name]
    Ls = sorted({c.L for c in chains})
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
        "config": {"workload": f"{a.config}: {W['desc']}", "n_chains": nch, "L": Ls[0] if len(Ls) == 1 else Ls,
                   "precision": head_name, "groups": head["groups"], "exec": head["exec"]},
        "flip_updates_per_s": head["flips_per_s"],
        "accepted_updates_per_s": head["acc_per_s"],
        "acceptance": head["acceptance"],
    }
    for k, r in res.items():
        out[f"{k}_sweeps_per_s"] = r["sweeps_per_s"]
        out[f"{k}_ms_per_step"] = r["dt"] / a.steps * 1e3
        if r["tuning"]:
            out[f"{k}_autotune"] = r["tuning"]
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = res["mixed"]["sweeps_per_s"] / res["fp64"]["sweeps_per_s"]
    if W.get("extra_groups") and a.groups is None and "mixed" in res:
        # the same chains in another number of contexts (HIP streams), timed on their own:
        # with several, one group's kernels fill the gaps and idle CUs of another's
        eg = W["extra_groups"]
        r2 = run_mode(a.config, "mixed", a.steps, a.warmup, "", a.delay, (), groups=eg, exec_mode=a.exec)
        out[f"mixed_{eg}groups_sweeps_per_s"] = r2["sweeps_per_s"]
    if a.plain_steps > 0 and a.config == "c4":
        # the plain FP64 GPU form: one chain (the reference's one chain per rank) at the
        # same lattice, no batching -- the denominator of the batched-mixed speed-up
        plain = run_mode(a.config, "fp64", a.plain_steps, 1, "", a.delay, (), chains=batch_chains(a.config, 1))
        out["plain_fp64_1chain_sweeps_per_s"] = plain["sweeps_per_s"]
        out["batched_mixed_over_plain_fp64"] = (res["mixed"]["sweeps_per_s"] / plain["sweeps_per_s"]
                                                if "mixed" in res else None)
        out["batched_fp64_over_plain_fp64"] = (res["fp64"]["sweeps_per_s"] / plain["sweeps_per_s"]
                                               if "fp64" in res else None)
    Nrep = chains[0].L ** 2
    rf = roofline(a.roofline_kernel, Nrep, head["kms"], head["klaunch"], head["kwork"], head_name,
                  measured_traffic(a.roofline_kernel, a.config, head_name, nch / head["groups"]))
    rf["kernel_share_of_step"] = (head["kms"] / 1e3) / head["dt"] if head["roofline_live"] else None
    if head.get("kspan"):
        # several stream groups: their launches overlap; the rate over the time any
        # launch of the class runs (union of the groups' launch intervals)
        union_ms, sum_ms = head["kspan"]
        rf["launches_in_flight"] = sum_ms / union_ms
        rf["achieved_all_groups"] = head["kwork"] / (union_ms / 1e3) / (1e12 if rf["unit"] == "TFLOP/s" else 1e9)
        rf["frac_all_groups"] = rf["achieved_all_groups"] / rf["peak"]
        rf["union_share_of_step"] = (union_ms / 1e3) / head["dt"]
    rf["measured_in"] = "timed sweeps" if head["roofline_live"] else "one extra eager sweep after the timed ones"
    out["roofline"] = rf
    out["roofline_step"] = {k: step_roofline(r["shapes"], k, r["dt"] / a.steps * 1e3) for k, r in res.items()}
    for k, r in res.items():
        other = []
        for kind, (ms, nl, wk, share) in r["others"].items():
            ro = roofline(kind, Nrep, ms, nl, wk, k, measured_traffic(kind, a.config, k, nch / r["groups"]))
            ro["kernel_share_of_sweep"] = share
            other.append(ro)
        out["roofline_other" if k == head_name else f"roofline_other_{k}"] = other
    if precision == "both" and not a.no_shadow:
        dg, dobs = shadow_max_dG(a.config, 2, a.warmup, a.delay)
        out["max_dG"] = dg
        out["max_dobs"] = dobs
        progress("shadow sweep done")
    if not a.no_cpu:
        src = res["fp64"] if "fp64" in res else res["mixed"]
        out["cpu_baseline"] = cpu_baseline(a.config, src["snapshot"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
