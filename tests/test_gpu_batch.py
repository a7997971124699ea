"""Batches of independent systems of uneven particle count (2-D harmonic
traps with a one-body factor): one sweep launch and one measurement launch
for all systems must reproduce the single-system path exactly (same random
stream per walker), and a batched TDVP step must match the oracle."""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import make_cfg, oracle_system, rel

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")

NS = (3, 10, 20, 64)


def _cfgs(W=300, therm=0, ns=NS, **extra):
    out = []
    for i, N in enumerate(ns):
        out.append(make_cfg(N=N, D=2, L=0.0, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 2.0), one_body=1,
                            trap_omega=1.0, rc=6.0, hbar2_2m=0.5, W=W, step=0.4, therm=therm, SEED=11 + i,
                            **extra))
    return out


def _params(cfg, i):
    s = oracle_system(cfg)
    rng = np.random.default_rng(100 + i)
    uR = rng.normal(0, 0.05, s.P)
    uR[0] = 0.5 / rng.uniform(0.8, 1.2) ** 2      # one-body width parameter 1/(2 w^2)
    uI = rng.normal(0, 0.01, s.P)
    return s, uR, uI


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_sweep_matches_single_contexts(prec):
    cfgs = _cfgs()
    seed = 77
    print("batch sweep", prec, flush=True)
    with tdvmc.Batch(cfgs) as b:
        starts = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            starts.append((c.get_walkers(), uR, uI))
        b.set_sweep_lanes(32)
        b.sweep(40, seed=seed, precision=prec)
        got = [c.get_walkers() for c in b.systems]
        acc = [c.acceptance() for c in b.systems]
    for i, cfg in enumerate(cfgs):
        R0, uR, uI = starts[i]
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R0)
            ctx.set_sweep_lanes(32)
            ctx.sweep(40, seed=tdvmc.batch_system_seed(seed, i), precision=prec)
            assert np.array_equal(ctx.get_walkers(), got[i]), f"system {i} (N={NS[i]})"
            assert ctx.acceptance() == acc[i]


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_sweep_per_system_lanes_match_single_contexts(prec):
    """Without a shared lane setting each system sweeps with the smallest of
    8, 16, 32 lanes that covers N (32 above).  A walker's random stream does
    not depend on the lane count, and lanes past N add exact zeros to the
    group sum, so every system matches a single context with 16 or 32 lanes."""
    cfgs = _cfgs()
    seed = 91
    print("batch sweep per-system lanes", prec, flush=True)
    with tdvmc.Batch(cfgs) as b:
        starts = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            starts.append((c.get_walkers(), uR, uI))
        b.sweep(40, seed=seed, precision=prec)
        got = [c.get_walkers() for c in b.systems]
    for i, cfg in enumerate(cfgs):
        R0, uR, uI = starts[i]
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R0)
            ctx.set_sweep_lanes(16 if NS[i] <= 16 else 32)
            ctx.sweep(40, seed=tdvmc.batch_system_seed(seed, i), precision=prec)
            assert np.array_equal(ctx.get_walkers(), got[i]), f"system {i} (N={NS[i]})"


def test_batch_measure_matches_single_and_oracle():
    cfgs = _cfgs()
    print("batch measure", flush=True)
    with tdvmc.Batch(cfgs) as b:
        walkers = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            walkers.append((s, c.get_walkers(), uR, uI))
        res = b.measure(precision=tdvmc.FP64)
    for i, (s, R, uR, uI) in enumerate(walkers):
        out = o.eval_configs(s, R, uR, uI)
        m = o.moments(out["O"], out["ER"], out["EI"])
        r = res[i]
        assert rel(r["mO"], m["mO"]) < 1e-10
        assert rel(r["mE"], m["mE"]) < 1e-10
        assert rel(r["S"], m["S"]) < 1e-10
        assert rel(r["FR"], m["FR"]) < 1e-9
        assert rel(r["FI"], m["FI"]) < 1e-9


def test_batch_measure_boundary_families():
    """Each system of the batch with its own boundary-condition families."""
    pairs = [(1, 4), (3, 3), (4, 1), (2, 3)]
    cfgs = [dict(c, SPLINE_BC_ORIGIN=a, SPLINE_BC_CUTOFF=b) for c, (a, b) in zip(_cfgs(), pairs)]
    print("batch measure, boundary families", flush=True)
    with tdvmc.Batch(cfgs) as b:
        walkers = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            walkers.append((s, c.get_walkers(), uR, uI))
        res = b.measure(precision=tdvmc.FP64)
    for i, (s, R, uR, uI) in enumerate(walkers):
        assert (s.bc_origin, s.bc_cutoff) == pairs[i]
        out = o.eval_configs(s, R, uR, uI)
        m = o.moments(out["O"], out["ER"], out["EI"])
        r = res[i]
        assert rel(r["mO"], m["mO"]) < 1e-10
        assert rel(r["mE"], m["mE"]) < 1e-10
        assert rel(r["S"], m["S"]) < 1e-10
        assert rel(r["FR"], m["FR"]) < 1e-9


@pytest.mark.parametrize("solver", [0, 1])
def test_batch_euler_step_vs_oracle(solver):
    cfgs = _cfgs(therm=0, IMAGINARY_TIME=1, LINEAR_EQUATION_SOLVER_TYPE=solver, REG_EPS=1e-3)
    dt = 0.01
    print("batch euler", solver, flush=True)
    with tdvmc.Batch(cfgs) as b:
        ref = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            R = c.get_walkers()
            out = o.eval_configs(s, R, uR, uI)
            m = o.moments(out["O"], out["ER"], out["EI"])
            x = o.solve_param_dot(m, imag_time=1, precondition=0, eps=1e-3, solver_type=solver)
            ref.append((uR + dt * x["xR"], uI + dt * x["xI"], dt * np.max(np.abs(x["xR"])),
                        dt * np.max(np.abs(x["xI"]))))
        b.tdvp_step(dt, ode_type=0, precision=tdvmc.FP64)
        for i, c in enumerate(b.systems):
            gR, gI, _, _ = c.get_params()
            xR, xI, sR, sI = ref[i]
            assert np.max(np.abs(gR - xR)) < 1e-9 * sR + 1e-13, f"system {i}"
            assert np.max(np.abs(gI - xI)) < 1e-9 * sI + 1e-13, f"system {i}"


def test_batch_graph_replays_match_eager():
    # (these small, weakly constrained systems stay stable in imaginary time
    # at dt = 1e-3 with 300 walkers; larger N or dt diverge in single
    # contexts just the same)
    cfgs = _cfgs(therm=20, ns=(3, 10, 20), IMAGINARY_TIME=1)
    res = []
    for graph in (False, True):
        print("batch graph", graph, flush=True)
        with tdvmc.Batch(cfgs) as b:
            for i, c in enumerate(b.systems):
                s, uR, uI = _params(cfgs[i], i)
                c.set_params(uR, uI)
            for _ in range(3):
                b.tdvp_step(1e-3, ode_type=2, precision=tdvmc.FP64, use_graph=graph)
                b.measure(precision=tdvmc.FP64)             # eager work between replays
            res.append([np.concatenate(c.get_params()[:2]) for c in b.systems]
                       + [c.get_walkers() for c in b.systems])
    for a, g in zip(res[0], res[1]):
        assert np.all(np.isfinite(a))
        assert np.allclose(a, g, rtol=1e-11, atol=1e-13)


def test_batch_autotune_keeps_results():
    """The measurement's walker shape (1, 2 or 4 wavefronts per walker, per
    system) changes the work split only: FP64 results before and after the
    autotuner's choice agree, and match the oracle."""
    ns = (3, 10, 33, 64, 100)                    # default shapes 1, 1, 2, 2, 4
    cfgs = _cfgs(ns=ns)
    print("batch autotune", flush=True)
    with tdvmc.Batch(cfgs) as b:
        walkers = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            walkers.append((s, c.get_walkers(), uR, uI))
        R0 = [c.get_walkers() for c in b.systems]
        before = b.measure(precision=tdvmc.FP64)
        waves, lanes = b.autotune(tdvmc.FP64)
        assert len(waves) == len(ns) and all(w in (1, 2, 4) for w in waves)
        assert all(l in (8, 16, 32, 64) for l in lanes)
        # the sweep timings ran on scratch copies: walkers and counters untouched
        for c, r in zip(b.systems, R0):
            assert np.array_equal(c.get_walkers(), r) and c.acceptance() == (0, 0)
        after = b.measure(precision=tdvmc.FP64)
        w2, l2 = b.autotune(tdvmc.MIXED)
        assert all(w in (1, 2, 4) for w in w2) and all(l in (8, 16, 32, 64) for l in l2)
        mixed = b.measure(precision=tdvmc.MIXED)
    print("waves per walker", waves, "lanes per walker", lanes)
    for i, (s, R, uR, uI) in enumerate(walkers):
        out = o.eval_configs(s, R, uR, uI)
        m = o.moments(out["O"], out["ER"], out["EI"])
        for k in ("mO", "mE", "S", "FR", "FI"):
            assert np.allclose(after[i][k], before[i][k], rtol=1e-11, atol=1e-14), (i, k)
        assert rel(after[i]["mO"], m["mO"]) < 1e-10
        assert rel(after[i]["S"], m["S"]) < 1e-10
        assert rel(after[i]["FR"], m["FR"]) < 1e-9
        assert rel(mixed[i]["mO"], m["mO"]) < 2e-6


def _f(s, R, x, eps=1e-3):
    P = s.P
    out = o.eval_configs(s, R, x[:P], x[P:2 * P])
    m = o.moments(out["O"], out["ER"], out["EI"])
    r = o.solve_param_dot(m, imag_time=1, precondition=0, eps=eps)
    return np.concatenate([r["xR"], r["xI"], [r["phiDotR"], r["phiDotI"]]])


def _oracle_step(s, R, x0, ode, dt):
    """Predictor-corrector (one or six correctors) or RK4 on the samples R:
    oracle.ode_step (src/TDVMC.cpp:1855-1967, :2037-2103) over the oracle's
    evaluation, moments and solve."""
    return o.ode_step(lambda x: _f(s, R, x), x0, ode, dt)


@pytest.mark.parametrize("ode", [1, 3, 5])
def test_batch_pc_and_reuse_steps_vs_oracle(ode):
    # type 1 samples afresh for its corrector, so it runs without moves; types
    # 3 and 5 run with moves: their later stages measure the samples of k1,
    # which are the walkers each system's step leaves behind
    therm = 0 if ode == 1 else 20
    cfgs = _cfgs(therm=therm, ns=(3, 10, 20), IMAGINARY_TIME=1, REG_EPS=1e-3)
    dt = 0.01
    print("batch ode", ode, flush=True)
    with tdvmc.Batch(cfgs) as b:
        start = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI, 0.1, -0.2)
            start.append((s, np.concatenate([uR, uI, [0.1, -0.2]]), c.get_walkers()))
        b.tdvp_step(dt, ode_type=ode, precision=tdvmc.FP64)
        for i, c in enumerate(b.systems):
            s, x0, R0 = start[i]
            R = c.get_walkers()
            assert (ode == 1) == np.array_equal(R, R0), f"system {i}"
            x1 = _oracle_step(s, R, x0, ode, dt)
            gR, gI, pR, pI = c.get_params()
            got = np.concatenate([gR, gI, [pR, pI]])
            assert np.max(np.abs(got - x1)) < 1e-9 * np.max(np.abs(x1 - x0)) + 1e-13, f"system {i}"


def test_batch_quench_crosses_switch_time_like_single_contexts():
    """Gaussian pair potentials quenched at t_q = SYSTEM_PARAMS[2]
    (src/PhysicalSystems/BosonsBulk.cpp:236-243): the batched Euler steps
    before and after t_q use the potential in force, as single contexts do.
    The batch's device descriptors copy each system's potential, so they must
    follow the switch; the captured step replays with the new values."""
    sp = (0.5, 2.0, 0.015, 0.5, 3.5)
    cfgs = _cfgs(therm=0, ns=(3, 10, 20), IMAGINARY_TIME=0, REG_EPS=1e-3, W=256)
    for c in cfgs:
        c["SYSTEM_PARAMS"] = list(sp)
    dt = 0.01
    print("batch quench", flush=True)
    for graph in (False, True):
        with tdvmc.Batch(cfgs) as b:
            start = []
            for i, c in enumerate(b.systems):
                s, uR, uI = _params(cfgs[i], i)
                c.set_params(uR, uI)
                start.append((uR, uI, c.get_walkers()))
            energies = []
            for _ in range(3):            # t = 0, 0.01 before t_q; t = 0.02 after
                m = b.measure(precision=tdvmc.FP64)
                energies.append([r["mE"][0] for r in m])
                b.tdvp_step(dt, ode_type=0, precision=tdvmc.FP64, use_graph=graph)
            got = [np.concatenate(c.get_params()[:2]) for c in b.systems]
        for i, cfg in enumerate(cfgs):
            uR, uI, R = start[i]
            with tdvmc.Context(cfg) as ctx:
                ctx.set_params(uR, uI)
                ctx.set_walkers(R)
                for k in range(3):
                    e = ctx.measure(precision=tdvmc.FP64)["localEnergyR"]
                    assert abs(e - energies[k][i]) < 1e-10 * max(1.0, abs(e)), (graph, i, k)
                    ctx.tdvp_step(dt, ode_type=0, precision=tdvmc.FP64)
                ref = np.concatenate(ctx.get_params()[:2])
            # the batched FP64 moments sum in another order than a single
            # context's; the solve passes those last-bit differences on
            change = np.max(np.abs(ref - np.concatenate([uR, uI])))
            diff = np.max(np.abs(got[i] - ref))
            print("system", i, "graph", graph, "change", change, "batch vs single", diff, flush=True)
            assert diff < 1e-8 * change + 1e-13, (graph, i)
    # the switch changes the energy: the third measurement sees the new potential
    assert abs(energies[2][0] - energies[1][0]) > 1e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_member_walker_count_change(prec):
    """A member context given a new walker count (new device buffers) is swept,
    measured and stepped by the batch with its new walkers: batched sweep equal
    to a single context, batched measurement equal to the oracle, a captured
    step rebuilt for the new launch shapes."""
    cfgs = _cfgs(W=200)
    seed = 123
    print("batch walker count", prec, flush=True)
    with tdvmc.Batch(cfgs) as b:
        params = []
        for i, c in enumerate(b.systems):
            s, uR, uI = _params(cfgs[i], i)
            c.set_params(uR, uI)
            params.append((s, uR, uI))
        b.set_sweep_lanes(32)
        b.sweep(10, seed=seed, precision=prec)          # descriptors built for W = 200
        b.tdvp_step(1e-4, ode_type=0, precision=prec, use_graph=True)
        sysd = b.systems[1]
        rng = np.random.default_rng(5)
        R_new = rng.normal(0.0, 1.0, (640, cfgs[1]["N"], 2))
        sysd.set_walkers(R_new)                         # W 200 -> 640: buffers reallocated
        for i, c in enumerate(b.systems):
            c.set_params(params[i][1], params[i][2])
        starts = [c.get_walkers() for c in b.systems]
        b.sweep(30, seed=seed + 1, precision=prec)
        got = [c.get_walkers() for c in b.systems]
        res = b.measure(precision=tdvmc.FP64)
        b.tdvp_step(1e-4, ode_type=0, precision=prec, use_graph=True)
        after = [np.concatenate(c.get_params()[:2]) for c in b.systems]
    assert got[1].shape == (640, cfgs[1]["N"], 2)
    for i, cfg in enumerate(cfgs):
        s, uR, uI = params[i]
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(starts[i])
            ctx.set_sweep_lanes(32)
            # the batch's stream epoch before this sweep: its first sweep and the
            # step's sweep launch advanced it once each
            ctx.set_rng_state(ctx.get_rng_state()[0], 2)
            ctx.sweep(30, seed=tdvmc.batch_system_seed(seed + 1, i), precision=prec)
            assert np.array_equal(ctx.get_walkers(), got[i]), f"system {i}"
        out = o.eval_configs(s, got[i], uR, uI)
        m = o.moments(out["O"], out["ER"], out["EI"])
        assert rel(res[i]["mO"], m["mO"]) < 1e-10, f"system {i}"
        assert rel(res[i]["S"], m["S"]) < 1e-10, f"system {i}"
        assert np.all(np.isfinite(after[i]))
