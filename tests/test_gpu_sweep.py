"""Statistical parity of the Metropolis sweep kernel (FP64 and mixed).

Sampled trajectories cannot match the reference move for move (it draws from
libstdc++'s normal_distribution on mt19937_64; the kernel uses Philox), so the
sampled distribution is checked instead:
  * two particles on a ring: the relative distance has density exp(2 u(r))
    (independent walkers -> chi-square test);
  * Gaussian one-body factor in open space: <sum r^2> = N D / (4 alpha);
  * a 3-D system with the k-space term: <E_L> and <O_k> agree with the
    oracle's own Metropolis chain within combined error bars.
"""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import make_cfg, oracle_system

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")


def _u(s, uR, r):
    R = np.zeros((len(r), 2, 1))
    R[:, 1, 0] = r
    return o.eval_configs(s, R, uR)["expo"]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("order", ["RANDOM", "SEQUENTIAL"])
def test_two_particles_on_ring(precision, order):
    L, P = 6.0, 6
    W = 65536
    cfg = make_cfg(N=2, D=1, L=L, P_s=P, pot="NONE", sysparams=(), W=W, step=1.2, MOVE_ORDER=order)
    s = oracle_system(cfg)
    uR = np.array([-1.2, -0.8, -0.2, 0.1, 0.05, 0.0])
    print("ring", precision, order, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        ctx.sweep(400, seed=1, precision=precision)
        R = ctx.get_walkers()
        acc, trials = ctx.acceptance()
    d = R[:, 0, 0] - R[:, 1, 0]
    r = np.abs(d - L * np.round(d / L))
    edges = np.linspace(0, L / 2, 25)
    hist, _ = np.histogram(r, bins=edges)
    grid = np.linspace(0, L / 2, 4801)
    dens = np.exp(2 * _u(s, uR, grid))
    cdf = np.concatenate([[0], np.cumsum((dens[1:] + dens[:-1]) / 2 * np.diff(grid))])
    mass = np.diff(np.interp(edges, grid, cdf))
    expect = mass / mass.sum() * W
    chi2 = float(np.sum((hist - expect) ** 2 / expect))
    print(f"chi2 = {chi2:.1f} (24 bins), acceptance {acc / trials:.3f}")
    assert chi2 < 60
    assert 0.1 < acc / trials < 0.95


@pytest.mark.parametrize("precision", [0, 1])
def test_gaussian_one_body_open_space(precision):
    N, D, alpha = 5, 2, 0.35
    W = 32768
    cfg = make_cfg(N=N, D=D, L=0.0, P_s=6, pot="NONE", sysparams=(), one_body=1, rc=2.0, W=W, step=0.8,
                   hbar2_2m=0.5, therm=0)
    uR = np.zeros(7)
    uR[0] = alpha
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        ctx.sweep(60 * N, seed=3, precision=precision)
        R = ctx.get_walkers()
        step = ctx.measure(precision=precision)
    r2 = (R ** 2).sum(axis=(1, 2))
    expect = N * D / (4 * alpha)
    err = r2.std() / np.sqrt(W)
    assert abs(r2.mean() - expect) < 5 * err, (r2.mean(), expect, err)
    assert abs(-step["localOperators"][0] - expect) < 5 * err


@pytest.mark.parametrize("precision", [0, 1])
def test_energy_vs_oracle_chain(precision):
    N, D, L, P, K = 8, 3, 3.0, 6, 6
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.6, 1.5), K=K, hbar2_2m=0.5, W=16384,
                   step=0.45, therm=0)
    s = oracle_system(cfg)
    rng = np.random.default_rng(2)
    uR = np.concatenate([rng.normal(0, 0.15, P), rng.normal(0, 0.02, K)])
    uR[P - 1] = 0.0
    print("chain", precision, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        ctx.sweep(100 * N, seed=5, precision=precision)
        g = ctx.measure(precision=precision)
        R = ctx.get_walkers()
    out = o.eval_configs(s, R, uR)
    E_gpu, E_err = out["ER"].mean(), out["ER"].std() / np.sqrt(len(out["ER"]))
    assert abs(g["localEnergyR"] - E_gpu) < 1e-4 * max(1.0, abs(E_gpu))
    # oracle chain: 1 walker, long run, blocking error bar
    R0 = R[0]
    _, rec, acc = o.metropolis(s, R0, uR, step=0.45, seed=11, n_samples=30000, n_therm=N)
    ref = o.eval_configs(s, rec[::3], uR)
    e = ref["ER"]
    blocks = e[: len(e) // 50 * 50].reshape(50, -1).mean(axis=1)
    E_ref, E_ref_err = e.mean(), blocks.std() / np.sqrt(50)
    tol = 4.5 * np.hypot(E_err, E_ref_err)
    print(f"<E_L> gpu {E_gpu:.5f} +- {E_err:.5f}, oracle {E_ref:.5f} +- {E_ref_err:.5f}")
    assert abs(E_gpu - E_ref) < tol
    o_gpu, o_ref = out["O"].mean(0), ref["O"].mean(0)
    o_err = np.hypot(out["O"].std(0) / np.sqrt(len(out["O"])),
                     ref["O"][: len(e) // 50 * 50].reshape(50, -1, s.P).mean(1).std(0) / np.sqrt(50))
    assert np.all(np.abs(o_gpu - o_ref) < 5 * o_err + 1e-12)


def test_sweep_is_deterministic_per_seed():
    cfg = make_cfg(N=10, D=2, L=4.0, P_s=6, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=4, W=256, step=0.4)
    outs = []
    for _ in range(2):
        with tdvmc.Context(cfg) as ctx:
            ctx.sweep(50, seed=42, precision=0)
            outs.append(ctx.get_walkers())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("precision", [0, 1])
def test_sampling_stage_rho_handoff(precision):
    """tdvmc_sample takes rho_k from the sweep instead of recomputing it; the
    stage's moments must equal those of an evaluation of the final walkers."""
    N, D, L, P, K = 27, 3, 4.0, 8, 26
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=K, hbar2_2m=0.5, W=2048,
                   step=0.3, therm=N)
    rng = np.random.default_rng(4)
    uR = np.concatenate([rng.normal(0, 0.15, P), rng.normal(0, 0.01, K)])
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        st = ctx.sample(seed=3, precision=precision)
        R = ctx.get_walkers()
        ev = ctx.eval_configs(R, precision=precision)
    tol = 1e-12 if precision == 0 else 2e-5
    assert np.allclose(st["localOperators"], ev["O"].mean(0), rtol=tol, atol=tol)
    assert abs(st["localEnergyR"] - ev["ER"].mean()) < tol * max(1.0, abs(ev["ER"].mean())) * 10


def test_autotune_keeps_walkers_and_stream():
    """The autotuner times sweeps of a scratch copy: the context's walkers,
    acceptance counts and random stream are the same as without tuning, and
    the chosen lane count is one of 16 / 32 / 64."""
    cfg = make_cfg(N=40, D=3, L=5.0, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=10, W=700, step=0.3)
    res = []
    for tune in (False, True):
        with tdvmc.Context(cfg) as ctx:
            ctx.init_walkers(5)
            R0 = ctx.get_walkers()
            if tune:
                print("autotune", flush=True)
                lanes = ctx.autotune(tdvmc.FP64)
                assert lanes in (16, 32, 64)
                assert np.array_equal(ctx.get_walkers(), R0)
                assert ctx.acceptance() == (0, 0)
                ctx.set_sweep_lanes(32)           # same lane count as the default: identical chain
            ctx.sweep(80, seed=3, precision=tdvmc.FP64)
            res.append(ctx.get_walkers())
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("precision", [0, 1])
def test_sampled_energy_vs_reference_chains(golden_dir, precision):
    """<E_R> sampled by the GPU sweep against 256 chains of the compiled
    reference (BosonsBulk's own move ratio in the DoMetropolisStep loop,
    tests/golden/sample_bulk3d_sw16.json): a 3-D square-well system of 16
    bosons.  The two chains use different generators, so they agree within
    combined error bars, not move for move."""
    import json
    import os

    g = json.load(open(os.path.join(golden_dir, "sample_bulk3d_sw16.json")))
    N, D, L = g["N"], g["D"], g["L"]
    W, n_meas = 8192, 60
    cfg = make_cfg(N=N, D=D, L=L, P_s=g["P"], pot="SQUARE_WELL", sysparams=tuple(g["sysparams"]), hbar2_2m=1.0,
                   W=W, step=g["step"])
    print("energy vs reference chains", precision, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(np.array(g["uR"]))
        ctx.sweep(g["burn"] * N, seed=3, precision=precision)
        E = np.zeros((n_meas, W))
        for k in range(n_meas):
            ctx.sweep(N, seed=100 + k, precision=precision)
            E[k] = ctx.eval_configs(ctx.get_walkers(), precision=tdvmc.FP64)["ER"]
    per_walker = E.mean(0)                     # walkers are independent chains
    mean, sigma = per_walker.mean(), per_walker.std(ddof=1) / np.sqrt(W)
    diff = abs(mean - g["mean"])
    comb = np.hypot(sigma, g["sigma"])
    print(f"GPU <E_R> {mean:.4f} +- {sigma:.4f}, reference {g['mean']:.4f} +- {g['sigma']:.4f}: "
          f"{diff / comb:.2f} sigma", flush=True)
    assert diff < 4.5 * comb


def test_rng_state_resume(tmp_path):
    """A chain stopped and resumed from saved walkers and a saved (seed, epoch)
    (tdvmc_get_rng_state, tdvmc_files_write_rng_state / read, tdvmc_set_rng_state)
    continues bit for bit as the uninterrupted chain."""
    import ctypes as C

    cfg = make_cfg(N=16, D=3, L=3.0, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=6, hbar2_2m=0.5, W=256,
                   step=0.3, SEED=17)
    uR = np.random.default_rng(2).normal(0, 0.1, 14)
    print("rng resume", flush=True)
    with tdvmc.Context(cfg) as a:
        a.set_params(uR)
        a.sweep(160, seed=5)
        R = a.get_walkers()
        seed, epoch = a.get_rng_state()
        L = tdvmc.lib()
        assert L.tdvmc_files_write_rng_state((str(tmp_path) + "/").encode(), seed, epoch, 0) == 0
        a.sweep(96, seed=9)
        a.sweep(32, seed=11)
        R_cont = a.get_walkers()
    s, e = C.c_uint64(), C.c_uint32()
    assert L.tdvmc_files_read_rng_state(str(tmp_path / "random" / "state_philox_0.dat").encode(), C.byref(s),
                                        C.byref(e)) == 0
    assert (s.value, e.value) == (17, epoch) and epoch == 1
    with tdvmc.Context(cfg) as b:
        b.set_params(uR)
        b.set_walkers(R)
        b.set_rng_state(s.value, e.value)
        b.sweep(96, seed=9)
        b.sweep(32, seed=11)
        assert np.array_equal(b.get_walkers(), R_cont)
