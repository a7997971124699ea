"""Non-uniform pair-spline knots (USE_NURBS / NURBS_GRID) on the GPU.

The reference's own NURBS evaluations (tests/golden/eval_nurbs_*.npz) are
checked by test_gpu_eval.py::test_fp64_vs_reference_golden.  Here:
  * FP64 and mixed evaluation against the oracle on non-uniform knots, with
    the Lennard-Jones and Gaussian potentials, k-space terms and pairs
    placed exactly on knots and at rc;
  * the Metropolis sweep samples exp(2 u(r)) of the non-uniform spline (two
    particles on a ring, chi-square), FP64 and mixed;
  * <E_L> and <O> of a 3-D sweep against the oracle's own chain;
  * one TDVP step runs in both forms and agrees between them.
"""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import energy_scale, make_cfg, oracle_system, rel

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")


def _grid(P, rc, a, seed):
    rng = np.random.default_rng(seed)
    g = rc * (np.arange(P) / (P - 1)) ** a
    gaps = np.diff(g)
    g[1:-1] += rng.uniform(-0.15, 0.15, P - 2) * np.minimum(gaps[:-1], gaps[1:])
    return g


CASES = {
    "lj3d_k26": (dict(N=27, D=3, L=4.2, P_s=12, pot="LJ", sysparams=(1.0, 1.5), K=26, hbar2_2m=0.7), 1.5, 6),
    "gauss2d_k24": (dict(N=30, D=2, L=6.0, P_s=10, pot="GAUSSIAN", sysparams=(0.5, 2.0), K=24, hbar2_2m=0.5), 0.8, 6),
    "gauss1d_n20": (dict(N=20, D=1, L=20.0, P_s=14, pot="GAUSSIAN", sysparams=(1.0, 1.0), hbar2_2m=0.5), 1.2, 8),
    "lj3d_n100": (dict(N=100, D=3, L=7.0, P_s=18, pot="LJ", sysparams=(1.0, 1.0)), 1.3, 3),
}


def _case(name, seed=0):
    kw, a, W = CASES[name]
    cfg = make_cfg(**kw)
    g = _grid(kw["P_s"], kw["L"] / 2.0, a, seed + 7)
    cfg.update(USE_NURBS=1, NURBS_GRID=g.tolist())
    s = oracle_system(cfg)
    rng = np.random.default_rng(seed + 100)
    N, D = s.N, s.D
    n = int(np.ceil(N ** (1.0 / D) - 1e-9))
    lat = np.stack(np.meshgrid(*[np.arange(n)] * D, indexing="ij"), -1).reshape(-1, D)[:N]
    R = (lat[None] + rng.uniform(-0.25, 0.25, (W, N, D))) * (s.L / n)
    R += rng.integers(-1, 2, (W, N, D)) * s.L
    # one pair per walker on a knot (or at rc): the interval tie rule
    for w in range(W):
        d = np.zeros(D)
        d[0] = g[1 + w % (len(g) - 2)]
        R[w, 1] = R[w, 0] + d
    uR = rng.normal(0, 0.2, s.P)
    uI = rng.normal(0, 0.1, s.P)
    if s.K:
        uR[s.P_s:] *= 0.01
        uI[s.P_s:] *= 0.01
    return cfg, s, R, uR, uI


@pytest.mark.parametrize("name", list(CASES))
def test_nurbs_eval_vs_oracle(name):
    cfg, s, R, uR, uI = _case(name)
    print("nurbs eval", name, flush=True)
    with tdvmc.Context(cfg) as ctx:
        assert tdvmc.parse_config(cfg).use_nurbs == 1
        ctx.set_params(uR, uI, 0.1, 0.0)
        f64 = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=True)
        nopp = ctx.eval_configs(R, precision=tdvmc.FP64)
        mix = ctx.eval_configs(R, precision=tdvmc.MIXED, per_particle=True)
    ref = o.eval_configs(s, R, uR, uI, phiR=0.1)
    sc = energy_scale(ref)
    assert rel(f64["expo"], ref["expo"]) < 1e-10
    assert rel(f64["O"], ref["O"]) < 1e-10
    assert np.max(np.abs(f64["ER"] - ref["ER"]) / sc) < 1e-10
    assert np.max(np.abs(f64["EI"] - ref["EI"]) / sc) < 1e-10
    assert rel(f64["grad"], ref["grad"]) < 1e-10
    assert rel(f64["lap"], ref["lap"]) < 1e-10
    assert np.max(np.abs(nopp["ER"] - ref["ER"]) / sc) < 1e-10
    e_O = rel(mix["O"], ref["O"])
    e_E = np.max(np.abs(mix["ER"] - ref["ER"]) / sc)
    e_g = rel(mix["grad"], ref["grad"])
    print(f"mixed errors {name}: O {e_O:.2e} E_L {e_E:.2e} grad {e_g:.2e}")
    # the bounds of the uniform-knot mixed form (test_gpu_eval.py::test_mixed_vs_oracle)
    assert e_O < 2e-6
    assert e_E < 1e-5
    assert e_g < 1e-5


def _u(s, uR, r):
    R = np.zeros((len(r), 2, 1))
    R[:, 1, 0] = r
    return o.eval_configs(s, R, uR)["expo"]


@pytest.mark.parametrize("precision", [0, 1])
def test_nurbs_two_particles_on_ring(precision):
    L, P = 6.0, 6
    W = 65536
    g = [0.0, 0.25, 0.7, 1.4, 2.1, 3.0]
    cfg = make_cfg(N=2, D=1, L=L, P_s=P, pot="NONE", sysparams=(), W=W, step=1.2)
    cfg.update(USE_NURBS=1, NURBS_GRID=g)
    s = oracle_system(cfg)
    uR = np.array([-1.2, -0.8, -0.2, 0.1, 0.05, 0.0])
    print("nurbs ring", precision, flush=True)
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
def test_nurbs_energy_vs_oracle_chain(precision):
    N, D, L, P, K = 8, 3, 3.0, 6, 6
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.6, 1.5), K=K, hbar2_2m=0.5, W=16384,
                   step=0.45, therm=0)
    cfg.update(USE_NURBS=1, NURBS_GRID=[0.0, 0.15, 0.45, 0.8, 1.15, 1.5])
    s = oracle_system(cfg)
    rng = np.random.default_rng(2)
    uR = np.concatenate([rng.normal(0, 0.15, P), rng.normal(0, 0.02, K)])
    uR[P - 1] = 0.0
    print("nurbs chain", precision, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        ctx.sweep(100 * N, seed=5, precision=precision)
        g = ctx.measure(precision=precision)
        R = ctx.get_walkers()
    out = o.eval_configs(s, R, uR)
    E_gpu, E_err = out["ER"].mean(), out["ER"].std() / np.sqrt(len(out["ER"]))
    assert abs(g["localEnergyR"] - E_gpu) < 1e-4 * max(1.0, abs(E_gpu))
    _, rec, acc = o.metropolis(s, R[0], uR, step=0.45, seed=11, n_samples=30000, n_therm=N)
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


def test_nurbs_tdvp_step_mixed_vs_fp64():
    """One RK4 step (measurement-only stages on fixed walkers, as in
    test_gpu_tdvp.py::test_mixed_step_close_to_fp64) on non-uniform knots."""
    N, D, L, P, K, W = 9, 3, 3.2, 6, 4, 512
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.5, 1.2), K=K, hbar2_2m=0.5, W=W,
                   therm=0, init=0, IMAGINARY_TIME=1, REG_EPS=1e-3, USE_NURBS=1,
                   NURBS_GRID=[0.0, 0.12, 0.4, 0.75, 1.2, 1.6])
    rng = np.random.default_rng(8)
    R = rng.uniform(-L / 2, L / 2, (W, N, D))
    uR = rng.normal(0, 0.2, P + K)
    uI = rng.normal(0, 0.05, P + K)
    uR[P:] *= 0.05
    uI[P:] *= 0.05
    res = {}
    for prec in (tdvmc.FP64, tdvmc.MIXED):
        print("nurbs tdvp step", prec, flush=True)
        with tdvmc.Context(cfg) as ctx:
            ctx.set_params(uR, uI)
            ctx.set_walkers(R)
            ctx.tdvp_step(0.01, ode_type=2, precision=prec)
            res[prec] = np.concatenate(ctx.get_params()[:2])
    step = res[tdvmc.FP64] - np.concatenate([uR, uI])
    assert np.all(np.isfinite(res[tdvmc.FP64])) and np.all(np.isfinite(res[tdvmc.MIXED]))
    err = np.max(np.abs(res[tdvmc.MIXED] - res[tdvmc.FP64])) / np.max(np.abs(step))
    print(f"nurbs RK4 step: mixed vs FP64 relative difference {err:.2e}")
    # the bound of the uniform-knot step (test_gpu_tdvp.py::test_mixed_step_close_to_fp64)
    assert err < 1e-5


@pytest.mark.parametrize("knots", ["uniform", "nurbs"])
def test_basis_sums_vs_oracle(knots):
    """tdvmc_basis_sums (WFParts::SplinedFunction::splineSums of the C++ view):
    against the oracle's B_s, the partition of unity (sum_s B_s = pairs within
    rc), and the boundary-condition map onto the evaluation's O."""
    import ctypes as C

    cfg, s, R, uR, uI = _case("gauss2d_k24", seed=3)
    if knots == "uniform":
        cfg.pop("USE_NURBS")
        cfg.pop("NURBS_GRID")
        s = oracle_system(cfg)
    W, P = R.shape[0], s.P_s
    B_ref, n_in = o.basis_sums(s, R)
    assert np.allclose(B_ref.sum(1), n_in, rtol=1e-13)
    dp = C.POINTER(C.c_double)
    print("basis sums", knots, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI)
        out = {}
        for prec in (tdvmc.FP64, tdvmc.MIXED):
            B = np.zeros((W, P + 2))
            Rc = np.ascontiguousarray(R)
            assert tdvmc.lib().tdvmc_basis_sums(ctx._h, Rc.ctypes.data_as(dp), W, prec,
                                                B.ctypes.data_as(dp)) == tdvmc.OK
            out[prec] = B
        O = ctx.eval_configs(R, precision=tdvmc.FP64)["O"][:, :P]
    assert rel(out[tdvmc.FP64], B_ref) < 1e-12
    assert rel(out[tdvmc.MIXED], B_ref) < 2e-6
    B = out[tdvmc.FP64]
    T_B = np.concatenate([B[:, [1]], B[:, [0]] + B[:, [2]], B[:, 3:P - 1], B[:, [P - 1]] + B[:, [P + 1]], B[:, [P]]], 1)
    assert rel(T_B, O) < 1e-12
