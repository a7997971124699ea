"""GPU TDVP step (Euler, predictor-corrector, RK4, and the variants that reuse
the samples of the first stage) against the oracle, and hipGraph replay.

With no Metropolis moves between stages (MC_NTHERMSTEPS = 0) every RK stage
measures the same fed walkers, so the step is deterministic: the parameters
after one step must match the oracle's restatement of
CalculateNextParametersRK4 / ...Euler / ...PC / ...RK4ReuseSamples /
...PCReuseSamples (src/TDVMC.cpp:1834-2103) built from the oracle's evaluation,
moments and solve, within 1e-9.  The reuse variants are also checked after
real moves: their later stages evaluate the samples of the first stage, which
are the walkers the step leaves behind.
"""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import make_cfg, oracle_system

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")


def _setup(imag, K=4):
    N, D, L, P = 9, 3, 3.2, 6
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.5, 1.2), K=K, hbar2_2m=0.5, W=512,
                   therm=0, init=0, IMAGINARY_TIME=imag, REG_EPS=1e-3)
    s = oracle_system(cfg)
    rng = np.random.default_rng(7 + imag)
    R = rng.uniform(-L / 2, L / 2, (512, N, D))
    uR = rng.normal(0, 0.2, s.P)
    uI = rng.normal(0, 0.05, s.P)
    uR[P:] *= 0.05
    uI[P:] *= 0.05
    return cfg, s, R, uR, uI


def _f(s, R, x, imag, eps=1e-3):
    P = s.P
    uR, uI = x[:P], x[P:2 * P]
    out = o.eval_configs(s, R, uR, uI)
    m = o.moments(out["O"], out["ER"], out["EI"])
    r = o.solve_param_dot(m, imag_time=imag, precondition=0, eps=eps)
    return np.concatenate([r["xR"], r["xI"], [r["phiDotR"], r["phiDotI"]]])


def _oracle_step(s, R, x0, imag, ode, dt):
    """One step of ODE type `ode` on fixed samples R: oracle.ode_step (the
    integrators of src/TDVMC.cpp:1834-2103, pinned on a linear ODE) over the
    oracle's evaluation, moments and solve."""
    return o.ode_step(lambda x: _f(s, R, x, imag), x0, ode, dt)


@pytest.mark.parametrize("imag", [0, 1])
@pytest.mark.parametrize("ode", [0, 1, 2, 3, 5])
def test_step_vs_oracle(imag, ode):
    cfg, s, R, uR, uI = _setup(imag)
    dt = 0.01
    x0 = np.concatenate([uR, uI, [0.1, -0.2]])
    x1 = _oracle_step(s, R, x0, imag, ode, dt)
    print("tdvp", imag, ode, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, -0.2)
        ctx.set_walkers(R)
        ctx.tdvp_step(dt, ode_type=ode, precision=tdvmc.FP64)
        gR, gI, pR, pI = ctx.get_params()
        assert np.array_equal(ctx.get_walkers(), R)
    got = np.concatenate([gR, gI, [pR, pI]])
    step = x1 - x0
    assert np.max(np.abs(got - x1)) < 1e-9 * np.max(np.abs(step)) + 1e-13


def test_mixed_step_close_to_fp64():
    cfg, s, R, uR, uI = _setup(1)
    res = {}
    for prec in (0, 1):
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R)
            ctx.tdvp_step(0.01, ode_type=2, precision=prec)
            res[prec] = np.concatenate(ctx.get_params()[:2])
    step = res[0] - np.concatenate([uR, uI])
    err = np.max(np.abs(res[1] - res[0])) / np.max(np.abs(step))
    print(f"mixed vs fp64 RK4 step relative difference {err:.2e}")
    # measured 5e-7 (FP32 pair arithmetic, split-bf16 moments, FP64 solve)
    assert err < 1e-5


@pytest.mark.parametrize("therm", [0, 9])
def test_graph_replays_match_eager(therm):
    cfg, s, R, uR, uI = _setup(1)
    cfg["MC_NTHERMSTEPS"] = therm
    cfg["MC_STEP"] = 0.3
    res = []
    for graph in (False, True):
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R)
            traj = []
            for _ in range(3):
                ctx.tdvp_step(0.01, ode_type=2, precision=tdvmc.FP64, use_graph=graph)
                traj.append(np.concatenate(ctx.get_params()[:2]))
                ctx.measure(precision=tdvmc.FP64)      # eager work between replays
            res.append((np.array(traj), ctx.get_walkers()))
    # the evaluation sums pair terms with LDS atomics, whose order (and so the
    # last bit of FP64 sums) varies between launches
    assert np.allclose(res[0][0], res[1][0], rtol=1e-11, atol=1e-14)
    assert np.allclose(res[0][1], res[1][1], rtol=1e-12, atol=1e-12)


def test_profiled_graph_replays():
    """Event-record nodes inside the captured step time every replay: one
    sweep / measurement / moments / solve per RK stage, positive durations,
    and the same parameters as the unprofiled graph."""
    cfg, s, R, uR, uI = _setup(1)
    cfg["MC_NTHERMSTEPS"] = 9
    cfg["MC_STEP"] = 0.3
    res = []
    for prof in (False, True):
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R)
            ctx.profile_enable(prof)
            for _ in range(3):
                ctx.tdvp_step(0.01, ode_type=2, precision=tdvmc.FP64, use_graph=True)
            if prof:
                for name in ("sweep", "eval", "moments", "solve"):
                    ms, n = ctx.profile_get(name)
                    print(name, ms, n, flush=True)
                    assert n == 12 and ms > 0.0
            res.append(np.concatenate(ctx.get_params()[:2]))
    assert np.allclose(res[0], res[1], rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("solver", [0, 1])
def test_euler_step_normalize_adjust_vs_oracle(solver):
    """USE_NORMALIZE_WF (phi R := -exponent of walker 0 at the last sample)
    and USE_ADJUST_PARAMETERS after an Euler step, with either linear solver
    (src/TDVMC.cpp:3757-3775).  AdjustParameters (:2531-2567) subtracts the
    last pair-spline parameter from the N_PARAM pair-spline parameters only;
    the k-space amplitudes after them must be left as the step made them (a
    shift there would change the state, not only its normalisation)."""
    imag = 1
    cfg, s, R, uR, uI = _setup(imag)
    cfg.update(USE_NORMALIZE_WF=1, USE_ADJUST_PARAMETERS=1, LINEAR_EQUATION_SOLVER_TYPE=solver)
    dt = 0.01
    P = s.P
    out = o.eval_configs(s, R, uR, uI)
    mo = o.moments(out["O"], out["ER"], out["EI"])
    r = o.solve_param_dot(mo, imag_time=imag, precondition=0, eps=1e-3, solver_type=solver)
    xR = uR + dt * r["xR"]
    xI = uI + dt * r["xI"]
    Ps = cfg["N_PARAM"]
    assert P > Ps  # the k-space block follows the pair block
    xR[:Ps] -= xR[Ps - 1]
    xI[:Ps] -= xI[Ps - 1]
    phiR = -out["expo"][0]
    phiI = -0.2 + dt * r["phiDotI"]
    print("normalize/adjust", solver, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, -0.2)
        ctx.set_walkers(R)
        ctx.tdvp_step(dt, ode_type=0, precision=tdvmc.FP64)
        assert ctx.step_status() == 0
        gR, gI, pR, pI = ctx.get_params()
    scale = dt * max(np.max(np.abs(r["xR"])), np.max(np.abs(r["xI"])))
    assert np.max(np.abs(gR - xR)) < 1e-9 * scale + 1e-13
    assert np.max(np.abs(gI - xI)) < 1e-9 * scale + 1e-13
    assert gR[Ps - 1] == 0.0 and gI[Ps - 1] == 0.0
    assert np.all(gR[Ps:] != 0.0)
    assert abs(pR - phiR) < 1e-10 * max(1.0, abs(phiR))
    assert abs(pI - phiI) < 1e-10


def test_parameter_acceptance_check_rolls_back():
    """PARAMETER_ACCEPTANCE_CHECK_TYPE 2 (src/TDVMC.cpp:2610-2630): from the
    second step on, a relative parameter change above 0.1 rejects the step;
    the parameters and the time stay as they were."""
    cfg, s, R, uR, uI = _setup(1)
    cfg.update(USE_PARAMETER_ACCEPTANCE_CHECK=1, PARAMETER_ACCEPTANCE_CHECK_TYPE=2)
    print("acceptance check", flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI)
        ctx.set_walkers(R)
        ctx.tdvp_step(1e-4, ode_type=0, precision=tdvmc.FP64)     # first step: accepted
        before = np.concatenate(ctx.get_params()[:2])
        t0 = ctx.get_info().time
        with pytest.raises(tdvmc.TdvmcError) as e:
            ctx.tdvp_step(1e3, ode_type=0, precision=tdvmc.FP64)  # huge step: >10 % change
        assert e.value.code == tdvmc.ERR_REJECTED
        after = np.concatenate(ctx.get_params()[:2])
        assert np.array_equal(before, after)
        assert ctx.get_info().time == t0
        ctx.tdvp_step(1e-6, ode_type=0, precision=tdvmc.FP64)     # a small step is accepted again


@pytest.mark.parametrize("ode", [3, 5])
@pytest.mark.parametrize("graph", [False, True])
def test_reuse_after_moves_vs_oracle(ode, graph):
    # the first stage samples after N moves per walker; the later stages
    # evaluate those samples, which are the walkers the step leaves behind
    cfg, s, R, uR, uI = _setup(0)
    cfg.update(MC_NTHERMSTEPS=9, MC_STEP=0.3)
    dt = 0.01
    x0 = np.concatenate([uR, uI, [0.1, -0.2]])
    print("tdvp reuse after moves", ode, graph, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, -0.2)
        ctx.set_walkers(R)
        ctx.tdvp_step(dt, ode_type=ode, precision=tdvmc.FP64, use_graph=graph)
        gR, gI, pR, pI = ctx.get_params()
        R1 = ctx.get_walkers()
    assert not np.array_equal(R1, R)
    x1 = _oracle_step(s, R1, x0, 0, ode, dt)
    got = np.concatenate([gR, gI, [pR, pI]])
    assert np.max(np.abs(got - x1)) < 1e-9 * np.max(np.abs(x1 - x0)) + 1e-13


@pytest.mark.parametrize("ode", [3, 5])
def test_reuse_two_samples_per_walker(ode):
    # two samples per walker without moves: the kept copies hold the same
    # walkers twice, and the moments of the doubled set equal those of one
    cfg, s, R, uR, uI = _setup(1)
    cfg.update(SAMPLES_PER_WALKER=2)
    dt = 0.01
    x0 = np.concatenate([uR, uI, [0.1, -0.2]])
    x1 = _oracle_step(s, R, x0, 1, ode, dt)
    print("tdvp reuse spw=2", ode, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, -0.2)
        ctx.set_walkers(R)
        ctx.tdvp_step(dt, ode_type=ode, precision=tdvmc.FP64)
        gR, gI, pR, pI = ctx.get_params()
    got = np.concatenate([gR, gI, [pR, pI]])
    assert np.max(np.abs(got - x1)) < 1e-9 * np.max(np.abs(x1 - x0)) + 1e-13


def test_unsupported_ode_type_refused():
    cfg, s, R, uR, uI = _setup(0)
    with tdvmc.Context(cfg) as ctx:
        with pytest.raises(tdvmc.TdvmcError) as ei:
            ctx.tdvp_step(0.01, ode_type=4, precision=tdvmc.FP64)
    assert ei.value.code == tdvmc.ERR_BAD_ARG


@pytest.mark.parametrize("ode", [0, 1, 2, 3])
def test_trial_counts_graph_and_eager(ode):
    """tdvmc_get_acceptance counts every trial move once: steps replayed from a
    captured graph count what eager steps count, and both equal the fresh
    sampling stages of the ODE type (Euler 1, predictor-corrector 2, RK4 4,
    RK4 reusing samples 1) x (MC_NINITIALIZATIONSTEPS + MC_NTHERMSTEPS) x W."""
    N, D, L, Ps, W = 8, 3, 3.0, 6, 128
    therm, init = 16, 8
    cfg = make_cfg(N=N, D=D, L=L, P_s=Ps, pot="GAUSSIAN", sysparams=(0.5, 1.0), hbar2_2m=0.5, W=W,
                   therm=therm, init=init, IMAGINARY_TIME=1, REG_EPS=1e-3)
    stages = {0: 1, 1: 2, 2: 4, 3: 1}[ode]
    counts = []
    for graph in (False, True):
        print("trial counts", ode, graph, flush=True)
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(np.full(Ps, 0.01))
            for _ in range(3):
                ctx.tdvp_step(1e-4, ode_type=ode, precision=tdvmc.FP64, use_graph=graph)
            acc, trials = ctx.acceptance()
            counts.append(trials)
            assert 0 < acc <= trials
    assert counts[0] == counts[1] == 3 * stages * (init + therm) * W


def test_reference_config_samples_mc_nsteps(golden_dir):
    """The reference's own config/BosonsBulk1D.config (1-D, N = 20, 100
    converged parameters, MC_NSTEPS = 5000, MC_NTHERMSTEPS = 100,
    MC_NINITIALIZATIONSTEPS = 200): one sampling stage draws MC_NSTEPS
    samples per walker, as the reference does per rank (src/TDVMC.cpp:1073-1080),
    each after MC_NTHERMSTEPS moves."""
    import os

    path = os.path.join(golden_dir, "config_BosonsBulk1D.json")
    print("reference config sampling", flush=True)
    with tdvmc.Context(path) as ctx:
        info = ctx.get_info()
        assert info.samples_per_walker == 5000 and info.n_walkers == 1
        d = ctx.sample(seed=4, precision=tdvmc.FP64)
        acc, trials = ctx.acceptance()
    assert d["n_samples"] == 5000
    assert trials == 200 + 5000 * 100
    assert 0 < acc < trials
    assert np.isfinite(d["localEnergyR"]) and np.all(np.isfinite(d["S"]))
