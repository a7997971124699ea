"""GPU parity of the evaluation kernel (ln psi, O_k, E_L, gradient, Laplacian)
on fed configurations, through the C ABI.

FP64 plain form: within 1e-10 of the oracle (relative to the largest entry of
each quantity; E_L relative to the magnitude of its cancelling parts), and
against the reference's own golden vectors with the tolerance those carry
(see test_oracle_golden.py).
Mixed form (FP32 pair / spline / phase arithmetic, FP64 accumulation of the
per-walker sums), with and without per-particle outputs: O_k within 2e-6,
ln psi within 5e-7, E_L within 1e-5 of the magnitude of its cancelling parts,
gradients within 2e-5 (measured at most 4.3e-7, 5.1e-8, 2.1e-6 and 3.9e-6
over the cases); on the configuration-file cases up to N = 400 (measured
1.4e-6, 9e-8, 4.3e-6, 5.5e-6 at N = 343, K = 256) O_k within 5e-6 and E_L
within 2e-5.  The measured errors are printed.
"""
import glob
import os
import zlib

import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import energy_scale, make_cfg, oracle_system, ref_golden_tol, rel

pytestmark = pytest.mark.gpu

tdvmc = pytest.importorskip("tdvmc_amd")

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "eval_*.npz")))


def _ctx(cfg):
    return tdvmc.Context(cfg)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_fp64_vs_reference_golden(path):
    z = np.load(path, allow_pickle=False)
    N, D, L, P = int(z["N"]), int(z["D"]), float(z["L"]), int(z["P"])
    cfg = make_cfg(N, D, L, P, pot="SQUARE_WELL", sysparams=[float(x) for x in z["sysparams"]])
    if "grid" in z.files:
        # eval_nurbs_*: non-uniform knots
        cfg.update(USE_NURBS=1, NURBS_GRID=[float(x) for x in z["grid"]])
    print("case", os.path.basename(path), flush=True)
    with _ctx(cfg) as ctx:
        ctx.set_params(z["uR"], z["uI"], float(z["phiR"]), float(z["phiI"]))
        ctx.set_time(float(z["time"]))
        out = ctx.eval_configs(z["R"], precision=tdvmc.FP64, per_particle=True)
    # against the oracle (same formulation, pinned to the reference and to exact arithmetic)
    s = oracle_system(cfg)
    ref = o.eval_configs(s, z["R"], z["uR"], z["uI"], phiR=float(z["phiR"]), time=float(z["time"]))
    assert rel(out["expo"], ref["expo"]) < 1e-10
    assert rel(out["O"], ref["O"]) < 1e-10
    sc = energy_scale(ref)
    assert np.max(np.abs(out["ER"] - ref["ER"]) / sc) < 1e-10
    assert np.max(np.abs(out["EI"] - ref["EI"]) / sc) < 1e-10
    assert rel(out["grad"], ref["grad"]) < 1e-10
    assert rel(out["lap"], ref["lap"]) < 1e-10
    for k in range(9):
        if k == 2:
            continue
        assert rel(out["other"][:, k], ref["other"][:, k]) < 1e-10, k
    # against the reference's golden vectors directly
    tol = ref_golden_tol(P)
    assert rel(out["expo"], z["expo"]) < tol
    assert rel(out["O"], z["O"]) < tol
    assert rel(out["ER"], z["ER"]) < tol
    assert rel(out["EI"], z["EI"]) < tol


CASES = {
    # name: (cfg kwargs, walkers)
    "lj3d_k26": (dict(N=27, D=3, L=4.2, P_s=12, pot="LJ", sysparams=(1.0, 1.5), K=26, hbar2_2m=0.7), 6),
    "gauss2d_k24_quench": (dict(N=30, D=2, L=6.0, P_s=10, pot="GAUSSIAN", sysparams=(0.5, 2.0, 0.0, 0.4, 3.0),
                                K=24, hbar2_2m=0.5), 6),
    "gauss1d_n20": (dict(N=20, D=1, L=20.0, P_s=16, pot="GAUSSIAN", sysparams=(1.0, 1.0), hbar2_2m=0.5), 8),
    "trap2d_onebody": (dict(N=10, D=2, L=0.0, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 2.0), one_body=1,
                            hbar2_2m=0.5, trap_omega=1.0, rc=3.0), 8),
    "lj3d_n64_even": (dict(N=64, D=3, L=5.6, P_s=20, pot="LJ", sysparams=(1.0, 1.0), K=8), 4),
    "bulk3d_n343_k256": (dict(N=343, D=3, L=9.79, P_s=40, pot="LJ", sysparams=(1.0, 1.0), K=256), 2),
    # 90 = 64 + 26: the last block is met twice per ring (folded); 100 = 64 + 36: not folded
    "sqwell2d_n90": (dict(N=90, D=2, L=10.0, P_s=14, pot="SQUARE_WELL", sysparams=(0.6, 1.5), K=12), 3),
    "lj3d_n100": (dict(N=100, D=3, L=7.0, P_s=18, pot="LJ", sysparams=(1.0, 1.0)), 3),
}


def _case_data(name, seed=0):
    kw, W = CASES[name]
    cfg = make_cfg(**kw)
    s = oracle_system(cfg)
    rng = np.random.default_rng(zlib.crc32(name.encode()) + seed)
    N, D = s.N, s.D
    if s.L > 0:
        # jittered lattice so that LJ cores do not overlap
        n = int(np.ceil(N ** (1.0 / D) - 1e-9))
        grid = np.stack(np.meshgrid(*[np.arange(n)] * D, indexing="ij"), -1).reshape(-1, D)[:N]
        R = (grid[None] + rng.uniform(-0.25, 0.25, (W, N, D))) * (s.L / n)
        R += rng.integers(-1, 2, (W, N, D)) * s.L
    else:
        R = rng.normal(0, 1.0, (W, N, D))
    uR = rng.normal(0, 0.2, s.P)
    uI = rng.normal(0, 0.1, s.P)
    if s.one_body:
        uR[0], uI[0] = 0.45, 0.03
    if s.K:
        uR[s.P_s + s.one_body:] *= 0.01
        uI[s.P_s + s.one_body:] *= 0.01
    return cfg, s, R, uR, uI


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_vs_oracle(name):
    cfg, s, R, uR, uI = _case_data(name)
    print("case", name, flush=True)
    with _ctx(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, 0.0)
        ctx.set_time(1.0)
        out = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=True)
        nopp = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=False)
    ref = o.eval_configs(s, R, uR, uI, phiR=0.1, time=1.0)
    sc = energy_scale(ref)
    assert rel(out["expo"], ref["expo"]) < 1e-10
    assert rel(out["O"], ref["O"]) < 1e-10
    assert np.max(np.abs(out["ER"] - ref["ER"]) / sc) < 1e-10
    assert np.max(np.abs(out["EI"] - ref["EI"]) / sc) < 1e-10
    assert rel(out["grad"], ref["grad"]) < 1e-10
    assert rel(out["lap"], ref["lap"]) < 1e-10
    # the variant without per-particle outputs sums the Laplacian per pair
    assert np.max(np.abs(nopp["ER"] - ref["ER"]) / sc) < 1e-10
    assert rel(nopp["O"], ref["O"]) < 1e-10


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_vs_oracle(name):
    cfg, s, R, uR, uI = _case_data(name)
    print("case", name, flush=True)
    with _ctx(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, 0.0)
        ctx.set_time(1.0)
        out = ctx.eval_configs(R, precision=tdvmc.MIXED, per_particle=True)
        # without per-particle outputs (the sampling stages' path: sum of u''
        # from the interval moments)
        agg = ctx.eval_configs(R, precision=tdvmc.MIXED)
    ref = o.eval_configs(s, R, uR, uI, phiR=0.1, time=1.0)
    sc = energy_scale(ref)
    e_O = rel(out["O"], ref["O"])
    e_x = np.max(np.abs(out["expo"] - ref["expo"]) / (np.abs(ref["O"]) @ np.abs(uR) + 1e-300))
    e_E = np.max(np.abs(out["ER"] - ref["ER"]) / sc)
    e_g = rel(out["grad"], ref["grad"])
    e_Ea = max(np.max(np.abs(agg["ER"] - ref["ER"]) / sc), np.max(np.abs(agg["EI"] - ref["EI"]) / sc))
    e_Oa = rel(agg["O"], ref["O"])
    print(f"mixed errors {name}: O {e_O:.2e} lnpsi {e_x:.2e} E_L {e_E:.2e} grad {e_g:.2e} "
          f"(aggregate path: O {e_Oa:.2e} E_L {e_Ea:.2e})")
    assert e_O < 2e-6
    assert e_x < 5e-7
    assert e_E < 1e-5
    assert e_g < 2e-5
    assert e_Oa < 2e-6
    assert e_Ea < 1e-5


@pytest.mark.parametrize("name", list(CASES))
def test_real_parameters_vs_oracle(name):
    """All imaginary parameters zero (imaginary-time runs from a real start):
    the kernels skip the imaginary parts, which are exact zeros."""
    cfg, s, R, uR, uI = _case_data(name, seed=1)
    uI = np.zeros_like(uI)
    print("case real", name, flush=True)
    ref = o.eval_configs(s, R, uR, uI, phiR=0.1, time=1.0)
    sc = energy_scale(ref)
    with _ctx(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, 0.0)
        ctx.set_time(1.0)
        f64 = ctx.eval_configs(R, precision=tdvmc.FP64)
        mix = ctx.eval_configs(R, precision=tdvmc.MIXED)
    assert np.all(f64["EI"] == 0.0) and np.all(mix["EI"] == 0.0)
    assert rel(f64["O"], ref["O"]) < 1e-10
    assert rel(f64["expo"], ref["expo"]) < 1e-10
    assert np.max(np.abs(f64["ER"] - ref["ER"]) / sc) < 1e-10
    for k in (0, 1, 3, 4, 5, 6, 7, 8):
        assert rel(f64["other"][:, k], ref["other"][:, k]) < 1e-10 * max(1.0, np.max(np.abs(ref["other"][:, 0])) /
                                                                        max(np.max(np.abs(ref["other"][:, k])), 1e-300)), k
    e_O, e_E = rel(mix["O"], ref["O"]), np.max(np.abs(mix["ER"] - ref["ER"]) / sc)
    print(f"mixed errors real {name}: O {e_O:.2e} E_L {e_E:.2e}")
    assert e_O < 2e-6
    assert e_E < 1e-5


# The reference's own particle configurations (config/particleconfiguration_*.csv,
# kept as data under tests/golden/refconf_*.csv: one line of N*D comma-separated
# coordinates) fed to the benchmark workloads.  particleconfiguration_400 lies
# in the 20 x 20 box of the 2-D N = 400 workload as it is; the 64- and
# 343-particle files come from boxes of other sizes (their coordinates span
# about 4 and 7 x 2.15443), so they are scaled to the workload's box.
REFCONF = {
    "bulk2d_gauss400_k512": ("refconf_particleconfiguration_400.csv", 2, None),
    "bulk3d_lj64": ("refconf_particleconfiguration_64.csv", 3, 4.0),
    "bulk3d_lj343_k256": ("refconf_particleconfiguration_343_.csv", 3, 7 * 2.15443),
}


def _refconf_data(name, golden_dir):
    from tdvmc_amd import workloads

    fname, D, file_box = REFCONF[name]
    cfg = workloads.WORKLOADS[name](W=4)
    s = oracle_system(cfg)
    vals = np.array([float(x) for x in open(os.path.join(golden_dir, fname)).read().replace("\n", ",").split(",")
                     if x.strip()])
    R0 = vals.reshape(s.N, D)
    if file_box:
        R0 = R0 * (s.L / file_box)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    jitter = 0.02 * s.L / round(s.N ** (1.0 / D))
    R = np.stack([R0] + [R0 + rng.uniform(-jitter, jitter, R0.shape) for _ in range(3)])
    uR = np.array(cfg["PARAMS_REAL"], dtype=np.float64) + rng.normal(0, 0.05, s.P)
    uI = rng.normal(0, 0.05, s.P)
    uR[s.P_s:] = 1e-3 * rng.normal(0, 1, s.P - s.P_s)
    uI[s.P_s:] *= 0.01
    return cfg, s, R, uR, uI


@pytest.mark.parametrize("name", list(REFCONF))
def test_workloads_on_reference_configurations(name, golden_dir):
    """The 2-D N = 400 quench workload (Gaussian pair potential, V0 = 5 -> 7
    at t = 0, 48 spline parameters, 512 k-space terms), the 3-D N = 64 and
    N = 343 Lennard-Jones workloads, each on the reference's own particle
    configuration (and three jittered copies): FP64 within 1e-10 of the
    oracle, the mixed form within its stated bounds."""
    cfg, s, R, uR, uI = _refconf_data(name, golden_dir)
    t = 0.0                                   # at t_q = 0 the quenched V0 = 7 is in force
    print("reference configuration", name, flush=True)
    with _ctx(cfg) as ctx:
        ctx.set_params(uR, uI, 0.1, 0.0)
        ctx.set_time(t)
        f64 = ctx.eval_configs(R, precision=tdvmc.FP64, per_particle=True)
        mix = ctx.eval_configs(R, precision=tdvmc.MIXED, per_particle=True)
    ref = o.eval_configs(s, R, uR, uI, phiR=0.1, time=t)
    sc = energy_scale(ref)
    assert rel(f64["expo"], ref["expo"]) < 1e-10
    assert rel(f64["O"], ref["O"]) < 1e-10
    assert np.max(np.abs(f64["ER"] - ref["ER"]) / sc) < 1e-10
    assert np.max(np.abs(f64["EI"] - ref["EI"]) / sc) < 1e-10
    assert rel(f64["grad"], ref["grad"]) < 1e-10
    assert rel(f64["lap"], ref["lap"]) < 1e-10
    e_O = rel(mix["O"], ref["O"])
    e_x = np.max(np.abs(mix["expo"] - ref["expo"]) / (np.abs(ref["O"]) @ np.abs(uR) + 1e-300))
    e_E = np.max(np.abs(mix["ER"] - ref["ER"]) / sc)
    e_g = rel(mix["grad"], ref["grad"])
    print(f"mixed errors {name}: O {e_O:.2e} lnpsi {e_x:.2e} E_L {e_E:.2e} grad {e_g:.2e}", flush=True)
    assert e_O < 5e-6
    assert e_x < 5e-7
    assert e_E < 2e-5
    assert e_g < 2e-5
