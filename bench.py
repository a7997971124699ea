"""Throughput of the TDVMC hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
    python bench.py --drift STEPS [--walkers W]     (real-time quench, E(t) in both forms)

Workload (default): 3-D bulk, N = 343 He-4-like bosons (Lennard-Jones,
density 0.365 sigma^-3), Jastrow = 40-parameter pair spline + 256 k-space terms
(P = 296), 65536 walkers.  One step = one RK4 TDVP step = 4 sampling stages;
each stage = one sweep (N single-particle Metropolis moves) of every walker,
one measurement (ln psi, O_k, E_L) per walker, the moments S / F and the FP64
Cholesky solve -- all on the device, replayed from a captured hipGraph.

Printed: one JSON line.  value = walker-samples/s of the mixed-precision form
(FP32 pair / spline / phase arithmetic, split-bf16 MFMA moments, FP64 solve);
the same line carries the FP64 plain form's rate, TDVP steps/s, the roofline
of the dominant kernel (the sweep), and the reference CPU path timed on this
host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X published vector / matrix FP32 peak
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense bf16 MFMA
PEAK_HBM_GBPS = 8000.0

def measured_traffic(workload, W):
    """HBM bytes per launch of the sweep and the measurement, and per stage of
    the moments, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept in
    profiles/ for this workload and walker count (its `traffic-json:` line),
    or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic-json:"):
                    d = json.loads(line.split(":", 1)[1])
                    if d.get("walkers") == W:
                        d["source"] = os.path.relpath(path, ROOT)
                        return d
    except (OSError, ValueError):
        pass
    return None


def sweep_flops_per_move(N, D, K):
    """Algorithmic flops of one trial move: 2 (N-1) pair evaluations of
    f_pv = 4D + 9 flops, 2K phase / rho updates of 2D + 8 flops, + 10."""
    return 2 * (N - 1) * (4 * D + 9) + 2 * K * (2 * D + 8) + 10


def measure_flops(N, D, K):
    return N * (N - 1) / 2 * (8 * D + 29) + N * K * (2 * D + 12) + N * (4 * D + 2)


def obs_flops(N, D, V):
    """Algorithmic flops of one walker's g(r) and S(k): every pair with
    D subtractions, D minimum images (2 flops each), r^2 (2D - 1), sqrt and the
    bin (2): 5D + 2; every (particle, wave vector) with the phase (2D - 1),
    cos, sin and the two sums: 2D + 3."""
    return N * (N - 1) / 2 * (5 * D + 2) + N * V * (2 * D + 3)


def run_observables(cfg, samples=4, bins=200):
    """Pair distribution and S(k) of every walker (BosonsBulk's additional
    observables on the 50 wave-vector shells), `samples` observations per
    precision without moves in between; kernel time from HIP events on the
    context stream."""
    import tdvmc_amd as t

    cfg = dict(cfg, GR_BIN_COUNT=bins, MC_NADDITIONALSTEPS=samples, MC_NADDITIONALTHERMSTEPS=0,
               MC_NADDITIONALINITIALIZATIONSTEPS=0)
    out = {}
    with t.Context(cfg) as ctx:
        info = ctx.get_info()
        for i in range(2):
            ctx.sweep(info.N, seed=21 + i, precision=t.MIXED)
        gg, sg = ctx.observables_info()
        for name, prec in (("mixed", t.MIXED), ("fp64", t.FP64)):
            ctx.additional_properties(seed=5, n_samples=1, precision=prec)    # warm-up
            ctx.profile_enable(True)
            t0 = time.perf_counter()
            ctx.additional_properties(seed=5, precision=prec)
            el = time.perf_counter() - t0
            tot, n = ctx.profile_get("observables")
            ctx.profile_enable(False)
            out[name + "_ms"] = tot / max(n, 1)
            out[name + "_call_ms"] = el * 1e3
        N, D, W = info.N, info.DIM, info.n_walkers
    V = {1: 50, 2: 103, 3: 307}[D]
    out.update({"gr_bins": len(gg), "sk_shells": len(sg), "wave_vectors": V, "walkers": W,
                "samples_per_call": samples,
                "roofline_mixed": kernel_roofline(f"obs_kernel<float, {D}>", out["mixed_ms"], W * obs_flops(N, D, V),
                                                  PEAK_FP32_TFLOPS)})
    return out


def run_gpu(cfg, steps, warmup, precision, use_graph=True, ode=2):
    """Times `steps` RK4 TDVP steps.  Kernel durations come from HIP events
    recorded on the context stream around every sweep / measurement /
    moments / solve launch of the timed steps themselves (inside the captured
    graph they are event-record nodes, re-recorded by every replay)."""
    import tdvmc_amd as t

    ctx = t.Context(cfg)
    info = ctx.get_info()
    W = info.n_walkers
    # thermalise the lattice start a little (untimed; launches of the same
    # shape as the timed ones, N moves each), then let the autotuner pick the
    # sweep's lanes per walker on these walkers
    for i in range(2):
        ctx.sweep(info.N, seed=11 + i, precision=precision)
    lanes = ctx.autotune(precision)
    dt = info.timestep
    ctx.profile_enable(True)
    for _ in range(warmup):
        ctx.tdvp_step(dt, ode_type=ode, precision=precision, use_graph=use_graph)
    ctx.synchronize()
    ctx.profile_enable(True)          # clears the counters; the graph is kept
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.tdvp_step(dt, ode_type=ode, precision=precision, use_graph=use_graph)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = {k: ctx.profile_get(k) for k in ("sweep", "eval", "moments", "solve")}
    ctx.profile_enable(False)
    stages = 4 * steps
    samples = stages * W * info.samples_per_walker
    acc, trials = ctx.acceptance()
    uR, _, _, _ = ctx.get_params()
    ok = bool(np.all(np.isfinite(uR)))
    ctx.close()
    return dict(elapsed=elapsed, samples=samples, steps=steps, W=W, prof=prof,
                acceptance=acc / max(trials, 1), finite=ok, info=info, lanes=lanes)


def run_drift(cfg, steps, therm_sweeps=100):
    """Real-time TDVP after the quench: the same start (walkers thermalised
    with `therm_sweeps` sweeps of N moves from the lattice start, the
    configured parameters) integrated `steps` RK4 steps in the FP64 form and
    in the mixed form.  <E_L>(t) comes from a device measurement of the
    current walkers after every step (not timed)."""
    import tdvmc_amd as t

    with t.Context(cfg) as c0:
        info = c0.get_info()
        for i in range(therm_sweeps):
            c0.sweep(info.N, seed=5 + i, precision=t.MIXED)
        R0 = c0.get_walkers()
        er = c0.eval_configs(R0, precision=t.FP64)["ER"]
        print(f"drift: start after {therm_sweeps} sweeps, <E_L> = {er.mean():.6f}", file=sys.stderr, flush=True)
    W = R0.shape[0]
    sigma = float(er.std() / np.sqrt(W))
    dt = info.timestep
    out = {}
    for name, prec in (("fp64", t.FP64), ("mixed", t.MIXED)):
        with t.Context(cfg) as ctx:
            ctx.set_walkers(R0)
            E = [ctx.measure(prec)["localEnergyR"]]
            spent = 0.0
            for s in range(steps):
                ctx.synchronize()
                t0 = time.perf_counter()
                ctx.tdvp_step(dt, ode_type=2, precision=prec, use_graph=True)
                ctx.synchronize()
                spent += time.perf_counter() - t0
                E.append(ctx.measure(prec)["localEnergyR"])
                if (s + 1) % 20 == 0:
                    print(f"drift {name}: step {s + 1}/{steps}, E = {E[-1]:.6f}, {spent:.1f} s", file=sys.stderr,
                          flush=True)
            uR = ctx.get_params()[0]
            E = np.array(E)
            out[name] = dict(E=E, steps_per_s=steps / spent, walker_samples_per_s=4 * steps * W / spent,
                             finite=bool(np.all(np.isfinite(uR)) and np.all(np.isfinite(E))))
    # control: the same walkers and sweeps (4 stages of MC_NTHERMSTEPS moves per
    # step) with the parameters held at their start values.  What it shows is
    # the chains still relaxing towards |psi|^2, not the TDVP dynamics.
    with t.Context(cfg) as ctx:
        ctx.set_walkers(R0)
        E = [ctx.measure(t.MIXED)["localEnergyR"]]
        for s in range(steps):
            for k in range(4):
                ctx.sweep(info.therm_moves, seed=900 + 4 * s + k, precision=t.MIXED)
            E.append(ctx.measure(t.MIXED)["localEnergyR"])
        out["fixed_params"] = dict(E=np.array(E))
    d = out["mixed"]["E"] - out["fp64"]["E"]
    res = {
        "mode": "drift", "workload": None, "walkers": W, "N": info.N, "DIM": info.DIM, "P": info.P,
        "K": info.K, "steps": steps, "dt": dt, "ode": "RK4, real time", "therm_sweeps": therm_sweeps,
        "sigma_E": sigma,
        "max_abs_mixed_minus_fp64": float(np.max(np.abs(d))),
        "max_abs_mixed_minus_fp64_over_sigma": float(np.max(np.abs(d)) / sigma),
    }
    Ef = out["fixed_params"]["E"]
    res["fixed_params"] = {"E0": float(Ef[0]), "E_end": float(Ef[-1]), "drift": float(Ef[-1] - Ef[0]),
                           "E_t": [float(x) for x in Ef],
                           "what": "same walkers and sweeps, parameters held at their start values (mixed form)"}
    for name in ("fp64", "mixed"):
        E = out[name]["E"]
        res[name] = {"E0": float(E[0]), "E_end": float(E[-1]), "drift": float(E[-1] - E[0]),
                     "max_abs_dev_from_E0": float(np.max(np.abs(E - E[0]))),
                     "tdvp_steps_per_s": out[name]["steps_per_s"],
                     "walker_samples_per_s": out[name]["walker_samples_per_s"], "finite": out[name]["finite"],
                     "E_t": [float(x) for x in E]}
    return res


def copy_probe(mib):
    """A plain elementwise kernel of known size (reads and writes `mib` MiB
    each, 10 launches), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(10):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "launches": 10, "bytes_read": n * 4, "bytes_written": n * 4}))


def _cpu_worker(args):
    """One host core: the reference BosonsBulk plug-in (oracle/_ref/libref.so,
    compiled from the reference sources) sampling one walker for budget_s:
    walker-samples of N moves + one measurement, as the reference's loop."""
    import ctypes as C

    from oracle import oracle as o
    from tdvmc_amd.workloads import mcmillan_params

    seed, budget_s = args
    r = o.ref()
    N, D, P = 343, 3, 40
    L = (N / (0.365 / 2.556 ** 3)) ** (1 / 3)
    n = 7
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    R = np.ascontiguousarray((g + rng.uniform(-0.05, 0.05, g.shape)) * (L / n), dtype=np.float64)
    uR = np.ascontiguousarray(mcmillan_params(P, L / 2, 1.1 * 2.556, 0.6 * 2.556))
    sp = np.array([2.2, 1.0])
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    mean = C.c_double()
    t0 = time.perf_counter()
    r.ref_bosonsbulk_sample(N, D, L, P, f(sp), 2, f(R), f(uR), 0.3 * 2.556, seed, 1, N, C.byref(mean))
    t1 = time.perf_counter() - t0
    n_samp = max(2, int(budget_s / max(t1, 1e-6)))
    t0 = time.perf_counter()
    r.ref_bosonsbulk_sample(N, D, L, P, f(sp), 2, f(R), f(uR), 0.3 * 2.556, seed + 1000, n_samp, N,
                            C.byref(mean))
    return n_samp, time.perf_counter() - t0


def host_cores():
    """Cores this process may run on, at most 16 (one GPU's share of a box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=15.0):
    """Reference BosonsBulk plug-in on every host core this process may use
    (one forked worker per core, each an independent walker, before this
    process touches the GPU): N = 343 3-D, 40 pair-spline parameters,
    square-well pair potential (the reference's BosonsBulk has the square well
    built in and no k-space term, so this is the pair-only part of the
    workload).  Throughput = total walker-samples / wall time."""
    try:
        from oracle import oracle as o

        if not o.ref_available():
            return None
    except Exception:
        return None
    import multiprocessing as mp

    cores = host_cores()
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(11 + 7 * c, budget_s) for c in range(cores)])
    wall = time.perf_counter() - t0
    total = sum(n for n, _ in res)
    # each worker's own rate, summed (the pool's start-up is not sample time)
    rate = sum(n / el for n, el in res)
    return dict(value=rate, unit="walker-samples/s", cores=cores, kind="reference", cpu=cpu_model(),
                nproc=os.cpu_count(), wall_s=wall,
                sample=f"{total} walker-samples (N=343 3-D, {343} moves + 1 measurement each) of the reference "
                       f"BosonsBulk plug-in, 40 pair parameters, square-well potential, no k-space term; "
                       f"{cores} processes (one walker each) for about {budget_s:.0f} s")


def run_batch_c4(steps, warmup):
    """64 trapped 2-D drops of uneven N as one batch: FP32 sweeps and
    measurement, FP64 moments and solves, one RK4 step per iteration."""
    import tdvmc_amd as t
    from tdvmc_amd.workloads import trap2d_batch64

    cfgs = trap2d_batch64()
    with t.Batch(cfgs) as b:
        b.sweep(0, seed=3, precision=t.MIXED)         # thermalise (each system N moves)
        waves, lanes = b.autotune(t.MIXED)            # measurement and sweep shapes per system
        for _ in range(warmup):
            b.tdvp_step(1e-3, ode_type=2, precision=t.MIXED, use_graph=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            b.tdvp_step(1e-3, ode_type=2, precision=t.MIXED, use_graph=True)
        el = time.perf_counter() - t0
        W = sum(c.info.n_walkers for c in b.systems)
        Ns = [c.N for c in b.systems]
    by_n, by_l = {}, {}
    for n, wv, ln in zip(Ns, waves, lanes):
        by_n.setdefault(n, set()).add(wv)
        by_l.setdefault(n, set()).add(ln)
    return {"systems": len(cfgs), "walkers": W, "N_range": [min(Ns), max(Ns)],
            "measure_waves_per_walker": {str(n): sorted(v) for n, v in sorted(by_n.items())},
            "sweep_lanes_per_walker": {str(n): sorted(v) for n, v in sorted(by_l.items())},
            "walker_samples_per_s": 4 * steps * W / el, "tdvp_steps_per_s": steps / el,
            "ms_per_step": el / steps * 1e3,
            "step": "RK4: 4 x (one batched sweep of N moves per system + batched measurement + "
                    "FP64 moments + FP64 solve per system)"}


def kernel_roofline(name, avg_ms, flops, peak):
    """`peak` is the published FP32 figure, which counts packed FP32 (two FMAs
    per lane per cycle).  The kernels are built without packed FP32, so their
    own vector-FMA limit is half of it; `frac_of_unpacked` is relative to that."""
    achieved = flops / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    unpacked = peak / 2
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": None, "kernel": name, "flops_per_launch": flops, "avg_ms": avg_ms,
            "peak_unpacked": unpacked, "frac_of_unpacked": achieved / unpacked}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="bulk3d_lj343_k256")
    ap.add_argument("--walkers", type=int, default=0)
    ap.add_argument("--fp64-steps", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched trap workload")
    ap.add_argument("--no-obs", action="store_true", help="skip the g(r) / S(k) observables")
    ap.add_argument("--no-reuse", action="store_true", help="skip the RK4 step that reuses the first stage's samples")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise kernel of this size (counter calibration)")
    ap.add_argument("--drift", type=int, default=0, metavar="STEPS",
                    help="only run the real-time quench workload for STEPS RK4 steps in both forms "
                         "and report <E_L>(t) and its drift")
    ap.add_argument("--drift-therm", type=int, default=100, metavar="SWEEPS",
                    help="thermalisation sweeps before the drift run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU path is design-only "
                                   "(walker shards + one all-reduce of the moments per stage)"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    from tdvmc_amd.workloads import WORKLOADS

    if args.drift:
        name = args.workload if args.workload != "bulk3d_lj343_k256" else "bulk2d_gauss400_k512"
        kw = {"W": args.walkers} if args.walkers else {}
        res = run_drift(WORKLOADS[name](**kw), args.drift, args.drift_therm)
        res["workload"] = name
        print(json.dumps(res))
        return

    # the host baseline forks its workers before this process touches the GPU
    cpu = None if args.no_cpu else cpu_baseline()
    make = WORKLOADS[args.workload]
    kw = {"W": args.walkers} if args.walkers else {}
    cfg = make(**kw)
    cfg["PRECISION"] = "MIXED"
    mixed = run_gpu(cfg, args.steps, args.warmup, precision=1, use_graph=not args.no_graph)
    fp64_steps = args.fp64_steps if args.fp64_steps >= 0 else max(1, args.steps // 2)
    fp64 = run_gpu(cfg, fp64_steps, 1, precision=0, use_graph=not args.no_graph) if fp64_steps > 0 else None

    info = mixed["info"]
    N, D, K = info.N, info.DIM, info.K
    W = mixed["W"]
    value = mixed["samples"] / mixed["elapsed"]
    ms_step = mixed["elapsed"] / mixed["steps"] * 1e3
    avg = {k: (v[0] / max(v[1], 1)) for k, v in mixed["prof"].items()}
    launches = {k: v[1] for k, v in mixed["prof"].items()}
    lpw = mixed["lanes"]
    kpl = next(m for m in (0, 2, 4, 8) if m * lpw >= info.K_terms)   # k terms per lane of the sweep template
    # algorithmic flops per launch: one sweep = N trial moves of every walker,
    # one measurement = every pair and every (particle, k) term of every walker
    rl = {
        "sweep": kernel_roofline(f"sweep_kernel<float, {D}, {lpw}, {kpl}, true>", avg["sweep"],
                                 W * info.therm_moves * sweep_flops_per_move(N, D, K), PEAK_FP32_TFLOPS),
        "eval": kernel_roofline(f"eval_kernel<float, {D}, false, true, float>", avg["eval"],
                                W * measure_flops(N, D, K), PEAK_FP32_TFLOPS),
    }
    dom = max(rl, key=lambda k: rl[k]["avg_ms"])
    # the moments S, F of one stage (column sums and means, centring pack,
    # split-bf16 GEMM, tile reduction) against the dense bf16 MFMA peak: six
    # bf16 products per entry of the lower triangle over the computed columns
    # (one per k / -k pair, plus E_R and E_I)
    n_col = info.one_body + info.P_pair + info.K_terms + 2
    mom_flops = W * n_col * (n_col + 1) / 2 * 2 * 6
    mom_ms = avg["moments"]
    rl["moments"] = {"bound": "mfma", "achieved": mom_flops / (mom_ms * 1e-3) / 1e12 if mom_ms > 0 else 0.0,
                     "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "traffic": None,
                     "kernel": "moments segment (colsum, colmean, pack_bf16, gemm_bf16x6, reduce_tiles)",
                     "flops_per_launch": mom_flops, "avg_ms": mom_ms, "columns": n_col}
    rl["moments"]["frac"] = rl["moments"]["achieved"] / PEAK_BF16_TFLOPS
    tr = measured_traffic(args.workload, W)
    if tr:
        for k in rl:
            if k in tr:
                rl[k]["traffic"] = tr[k]
                rl[k]["traffic_source"] = tr["source"]
        if rl["moments"]["traffic"] and mom_ms > 0:
            gbps = rl["moments"]["traffic"] / (mom_ms * 1e-3) / 1e9
            rl["moments"]["hbm_GBps"] = gbps
            rl["moments"]["hbm_frac"] = gbps / PEAK_HBM_GBPS
    line = {
        "metric": "walker-samples/s + TDVP steps/s, N=343 3D, mixed vs FP64 @ S,F rel err 1e-6",
        "value": value,
        "unit": "walker-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair/spline/phase + split-bf16 (bf16x6) MFMA moments + fp64 solve",
        "data": "synthetic (McMillan-initialised spline, seeded k-space amplitudes, jittered-lattice walkers)",
        "config": {"workload": args.workload, "N": N, "DIM": D, "P": info.P, "K": K,
                   "walkers": W, "step": "RK4 TDVP step = 4 x (1 sweep of N moves + measure + "
                   "moments + FP64 solve)", "graph": not args.no_graph,
                   "sweep_lanes_per_walker": {"mixed": mixed["lanes"], "fp64": fp64["lanes"] if fp64 else None}},
        "tdvp_steps_per_s": mixed["steps"] / mixed["elapsed"],
        "acceptance": mixed["acceptance"],
        "kernel_ms": avg,
        "kernel_launches": launches,
        "roofline": rl[dom],
        "roofline_other": {k: v for k, v in rl.items() if k != dom},
    }
    if fp64 is not None:
        v64 = fp64["samples"] / fp64["elapsed"]
        line["fp64_value"] = v64
        line["mixed_over_fp64"] = value / v64
        line["fp64_tdvp_steps_per_s"] = fp64["steps"] / fp64["elapsed"]
        line["fp64_kernel_ms"] = {k: (v[0] / max(v[1], 1)) for k, v in fp64["prof"].items()}
        # the plain FP64 form's two big kernels against the FP64 vector peak
        # (same algorithmic flops as the mixed form's)
        f64ms = line["fp64_kernel_ms"]
        lpw64 = fp64["lanes"]
        kpl64 = next(m for m in (0, 2, 4, 8) if m * lpw64 >= info.K_terms)
        line["roofline_fp64"] = {}
        for k, name, fl in (("sweep", f"sweep_kernel<double, {D}, {lpw64}, {kpl64}, true>",
                             W * info.therm_moves * sweep_flops_per_move(N, D, K)),
                            ("eval", f"eval_kernel<double, {D}, false, true, double>", W * measure_flops(N, D, K))):
            ach = fl / (f64ms[k] * 1e-3) / 1e12 if f64ms.get(k, 0) > 0 else 0.0
            line["roofline_fp64"][k] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                                        "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "kernel": name,
                                        "flops_per_launch": fl, "avg_ms": f64ms[k]}
    if not args.no_batch:
        line["batch_trap2d_64"] = run_batch_c4(max(1, args.steps // 2), 1)
    if not args.no_obs:
        line["observables_gr_sk"] = run_observables(cfg)
    if not args.no_reuse:
        # ODE type 3: one sampling stage and three measurements of the same samples per step
        r3 = run_gpu(cfg, max(1, args.steps // 2), 1, precision=1, use_graph=not args.no_graph, ode=3)
        line["rk4_reuse_samples"] = {"tdvp_steps_per_s": r3["steps"] / r3["elapsed"],
                                     "ms_per_step": r3["elapsed"] / r3["steps"] * 1e3,
                                     "kernel_ms": {k: (v[0] / max(v[1], 1)) for k, v in r3["prof"].items()},
                                     "kernel_launches": {k: v[1] for k, v in r3["prof"].items()}}
    line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
