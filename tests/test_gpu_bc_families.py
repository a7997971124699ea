"""Boundary-condition families of the pair spline (SPLINE_BC_ORIGIN /
SPLINE_BC_CUTOFF) on the GPU, through the C ABI, against the oracle whose map
tests/test_bc_families_cpu.py pins to the reference's factors and to the
conditions each family imposes.

* Evaluation on fed configurations for every accepted pair of families, with
  complex parameters and k-space terms: FP64 within 1e-10 of the oracle, with
  and without per-particle outputs (the two paths sum u'' differently), and
  the mixed form within O and ln psi 2e-6, E_L 1e-5 of the magnitude of its
  cancelling parts, gradients 1e-5 (measured at most 2.9e-7, 5.7e-7 and
  6.9e-7).  One case on NURBS knots.
* tdvmc_spline_bc returns the oracle's rows.
* One RK4 TDVP step on fed walkers within 1e-9 of the oracle's step.
* The sampled <E_L> of a sweep against an oracle Metropolis chain.
"""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import energy_scale, make_cfg, oracle_system, rel

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")

ROWS = {1: 3, 2: 2, 3: 2, 4: 1}
VALID = [(a, b) for a in ROWS for b in ROWS if ROWS[a] + ROWS[b] == 4]


def _case(pair, W=6, nurbs=False):
    N, D, L, P, K = 20, 3, 4.0, 9, 6
    extra = dict(SPLINE_BC_ORIGIN=pair[0], SPLINE_BC_CUTOFF=pair[1])
    if nurbs:
        extra.update(USE_NURBS=1, NURBS_GRID=[0.0, 0.2, 0.5, 0.75, 1.0, 1.3, 1.55, 1.75, 1.95])
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=K, hbar2_2m=0.5, W=W, **extra)
    s = oracle_system(cfg)
    rng = np.random.default_rng(10 * pair[0] + pair[1])
    R = rng.uniform(-L / 2, L / 2, (W, N, D))
    uR = rng.normal(0, 0.3, s.P)
    uI = rng.normal(0, 0.1, s.P)
    uR[P:] *= 0.05
    uI[P:] *= 0.05
    return cfg, s, R, uR, uI


def _check(out, ref, tol_o, tol_e, tol_g, pp, tag=""):
    if tag:
        sc = energy_scale(ref)
        print(f"bc errors {tag}: O {rel(out['O'], ref['O']):.2e} E_L "
              f"{np.max(np.abs(out['ER'] - ref['ER']) / sc):.2e}"
              + (f" grad {rel(out['grad'], ref['grad']):.2e}" if pp else ""))
    assert rel(out["expo"], ref["expo"]) < tol_o
    assert rel(out["O"], ref["O"]) < tol_o
    sc = energy_scale(ref)
    assert np.max(np.abs(out["ER"] - ref["ER"]) / sc) < tol_e
    assert np.max(np.abs(out["EI"] - ref["EI"]) / sc) < tol_e
    if pp:
        assert rel(out["grad"], ref["grad"]) < tol_g
        assert rel(out["lap"], ref["lap"]) < tol_g


@pytest.mark.parametrize("pair", VALID)
def test_eval_vs_oracle(pair):
    cfg, s, R, uR, uI = _case(pair)
    ref = o.eval_configs(s, R, uR, uI)
    print("bc", pair, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI)
        f_pp = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=True)
        f = ctx.eval_configs(R, precision=tdvmc.FP64)
        m = ctx.eval_configs(R, precision=tdvmc.MIXED, per_particle=True)
        a, b = ctx.spline_bc()
    assert np.array_equal(a, o.bc_factors(pair[0], False)) and np.array_equal(b, o.bc_factors(pair[1], True))
    _check(f_pp, ref, 1e-10, 1e-10, 1e-10, True)
    _check(f, ref, 1e-10, 1e-10, 1e-10, False)
    _check(m, ref, 2e-6, 1e-5, 1e-5, True, tag=f"mixed {pair}")


def test_eval_vs_oracle_nurbs():
    cfg, s, R, uR, uI = _case((3, 3), nurbs=True)
    ref = o.eval_configs(s, R, uR, uI)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI)
        f = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=True)
        m = ctx.eval_configs(R, precision=tdvmc.MIXED, per_particle=True)
    _check(f, ref, 1e-10, 1e-10, 1e-10, True)
    _check(m, ref, 2e-6, 1e-5, 1e-5, True, tag="mixed nurbs")


def _f(s, R, x, eps=1e-3):
    P = s.P
    out = o.eval_configs(s, R, x[:P], x[P:2 * P])
    mo = o.moments(out["O"], out["ER"], out["EI"])
    r = o.solve_param_dot(mo, imag_time=0, precondition=0, eps=eps)
    return np.concatenate([r["xR"], r["xI"], [r["phiDotR"], r["phiDotI"]]])


@pytest.mark.parametrize("pair", [(1, 4), (3, 3)])
def test_rk4_step_vs_oracle(pair):
    cfg, s, R, uR, uI = _case(pair, W=256)
    cfg.update(MC_NTHERMSTEPS=0, MC_NINITIALIZATIONSTEPS=0, IMAGINARY_TIME=0, REG_EPS=1e-3)
    dt = 0.005
    x0 = np.concatenate([uR, uI, [0.1, -0.2]])
    x1 = o.ode_step(lambda x: _f(s, R, x), x0, 2, dt)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, -0.2)
        ctx.set_walkers(R)
        ctx.tdvp_step(dt, ode_type=2, precision=tdvmc.FP64)
        gR, gI, pR, pI = ctx.get_params()
    got = np.concatenate([gR, gI, [pR, pI]])
    assert np.max(np.abs(got - x1)) < 1e-9 * np.max(np.abs(x1 - x0)) + 1e-13


def test_sampled_energy_vs_oracle_chain():
    N, D, L, P, K = 8, 3, 3.0, 6, 4
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="GAUSSIAN", sysparams=(0.6, 1.5), K=K, hbar2_2m=0.5, W=16384,
                   step=0.45, therm=0, SPLINE_BC_ORIGIN=4, SPLINE_BC_CUTOFF=1)
    s = oracle_system(cfg)
    rng = np.random.default_rng(12)
    uR = np.concatenate([rng.normal(0, 0.15, P), rng.normal(0, 0.02, K)])
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        ctx.sweep(100 * N, seed=9, precision=tdvmc.FP64)
        R = ctx.get_walkers()
    out = o.eval_configs(s, R, uR)
    E_gpu, E_err = out["ER"].mean(), out["ER"].std() / np.sqrt(len(out["ER"]))
    _, rec, _ = o.metropolis(s, R[0], uR, step=0.45, seed=13, n_samples=30000, n_therm=N)
    e = o.eval_configs(s, rec[::3], uR)["ER"]
    blocks = e[: len(e) // 50 * 50].reshape(50, -1).mean(axis=1)
    E_ref, E_ref_err = e.mean(), blocks.std() / np.sqrt(50)
    print(f"<E_L> gpu {E_gpu:.5f} +- {E_err:.5f}, oracle {E_ref:.5f} +- {E_ref_err:.5f}")
    assert abs(E_gpu - E_ref) < 4.5 * np.hypot(E_err, E_ref_err)
