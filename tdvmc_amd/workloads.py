"""Synthetic workloads (config dictionaries in the reference's config schema).

    bulk1d_gauss   1-D bulk bosons, N = 20, L = 20, Gaussian pair potential
                   exp(-r^2/2), hbar^2/2m = 1/2, 16 spline parameters ~ N(0, 0.1),
                   1024 walkers, Gaussian step 0.5
    bulk3d_lj64    3-D He-4-like bulk, N = 64 at density 0.365 sigma^-3 (Angstrom,
                   Kelvin), Lennard-Jones (sigma = 2.556, eps = 10.22), 32 spline
                   parameters from the McMillan form u = -(b/r)^5, b = 1.1 sigma,
                   16384 walkers on a jittered simple-cubic lattice, step 0.3 sigma
    bulk3d_lj343_k256
                   N = 343 at the same density, 40 spline parameters (McMillan) +
                   256 k-space terms a_k ~ N(0, 1e-3): P = 296, 65536 walkers
    bulk2d_gauss400_k512
                   2-D bulk, N = 400, L = 20, Gaussian pair potential V0 = 5,
                   width 0.2, quenched to V0 = 7 at t = 0; 48 spline parameters +
                   512 k-space terms: P = 560, 131072 walkers, RK4, dt = 1e-3,
                   real time, Tikhonov eps = 1e-4
    trap2d_drop    2-D harmonic trap (omega = 1), Gaussian pair repulsion
                   V0 = 2, width 0.5, one-body width w ~ U(0.8, 1.2) and 24 spline
                   parameters; N from {3, 6, 10, 20, 64, 100}
    trap2d_batch64 64 such drops with N and w drawn from those ranges, 4096
                   walkers each, as one batch
"""
from __future__ import annotations

import numpy as np

SIGMA_HE = 2.556          # Angstrom (src/Potentials/LJ_He_He.cpp:13-14)
EPS_HE = 10.22            # Kelvin
HBAR2_2M_HE4 = 6.0596     # K Angstrom^2 for He-4


def _base(**kw):
    cfg = {
        "CONFIG_VERSION": "0.24",
        "SYSTEM_TYPE": "BosonsBulk",
        "OUTPUT_DIRECTORY": "",
        "IMAGINARY_TIME": 1,
        "ODE_SOLVER_TYPE": 2,
        "LINEAR_EQUATION_SOLVER_TYPE": 0,
        "MC_NINITIALIZATIONSTEPS": 0,
        "SAMPLES_PER_WALKER": 1,
    }
    cfg.update(kw)
    return cfg


def mcmillan_params(P_s, rc, b, r_min):
    """Spline parameters alpha_k = u(k h) of u(r) = -(b / r)^5 sampled at the
    knots k h (k = 0 .. P_s-1, h = rc / (P_s - 1)), with r clipped below at
    r_min (the r = 0 knot diverges) and shifted so that u(rc) = 0."""
    h = rc / (P_s - 1)
    r = np.maximum(np.arange(P_s) * h, r_min)
    u = -(b / r) ** 5
    return u - (-(b / rc) ** 5)


def bulk1d_gauss(W=1024, seed=1):
    rng = np.random.default_rng(seed)
    P = 16
    return _base(N=20, DIM=1, LBOX=20.0, N_PARAM=P, HBAR2_2M=0.5, PAIR_POTENTIAL="GAUSSIAN",
                 SYSTEM_PARAMS=[1.0, 1.0], PARAMS_REAL=rng.normal(0.0, 0.1, P).tolist(),
                 MC_STEP=0.5, MC_NTHERMSTEPS=20, N_WALKERS=W, TIMESTEP=1e-3, ODE_SOLVER_TYPE=0,
                 PRECISION="FP64")


def _lj_bulk(N, P_s, K, W, seed, step_sigma=0.3, precision="MIXED"):
    rho = 0.365 / SIGMA_HE ** 3
    L = (N / rho) ** (1.0 / 3.0)
    rc = L / 2
    pair = mcmillan_params(P_s, rc, 1.1 * SIGMA_HE, 0.6 * SIGMA_HE)
    rng = np.random.default_rng(seed)
    ak = rng.normal(0.0, 1e-3, K) if K else np.zeros(0)
    params = np.concatenate([pair, ak])
    return _base(N=N, DIM=3, LBOX=L, N_PARAM=P_s, K_VECTORS=K, HBAR2_2M=HBAR2_2M_HE4, PAIR_POTENTIAL="LJ",
                 SYSTEM_PARAMS=[SIGMA_HE, EPS_HE], PARAMS_REAL=params.tolist(), MC_STEP=step_sigma * SIGMA_HE,
                 MC_NTHERMSTEPS=N, N_WALKERS=W, TIMESTEP=1e-4, PRECISION=precision)


def bulk3d_lj64(W=16384, seed=3):
    cfg = _lj_bulk(64, 32, 0, W, seed, precision="FP64")
    cfg["ODE_SOLVER_TYPE"] = 0
    return cfg


def bulk3d_lj343_k256(W=65536, seed=4, precision="MIXED"):
    return _lj_bulk(343, 40, 256, W, seed, precision=precision)


def bulk2d_gauss400_k512(W=131072, seed=7, precision="MIXED"):
    rng = np.random.default_rng(seed)
    P_s, K = 48, 512
    params = np.concatenate([np.zeros(P_s), rng.normal(0.0, 1e-4, K)])
    return _base(N=400, DIM=2, LBOX=20.0, N_PARAM=P_s, K_VECTORS=K, HBAR2_2M=0.5, PAIR_POTENTIAL="GAUSSIAN",
                 SYSTEM_PARAMS=[0.2, 5.0, 0.0, 0.2, 7.0], PARAMS_REAL=params.tolist(), MC_STEP=0.3,
                 MC_NTHERMSTEPS=400, N_WALKERS=W, IMAGINARY_TIME=0, TIMESTEP=1e-3, TOTALTIME=0.2,
                 REG_EPS=1e-4, PRECISION=precision)


def trap2d_drop(N, width, W=4096, seed=5):
    rng = np.random.default_rng(seed)
    P_s = 24
    params = np.concatenate([[1.0 / (2.0 * width ** 2)], rng.normal(0.0, 0.05, P_s)])
    params[-1] = 0.0
    return _base(SYSTEM_TYPE="BosonsTrap", N=N, DIM=2, LBOX=0.0, N_PARAM=P_s, ONE_BODY=1, TRAP_OMEGA=1.0,
                 SPLINE_CUTOFF=6.0, HBAR2_2M=0.5, PAIR_POTENTIAL="GAUSSIAN", SYSTEM_PARAMS=[0.5, 2.0],
                 PARAMS_REAL=params.tolist(), MC_STEP=0.5, MC_NTHERMSTEPS=N, N_WALKERS=W, PRECISION="MIXED")


def trap2d_batch64(W=4096, n_systems=64, seed_n=5, seed_w=6):
    """64 independent 2-D trapped drops (one batch): N drawn uniformly from
    {3, 6, 10, 20, 64, 100} (seed 5), one-body width w ~ U(0.8, 1.2) (seed 6),
    24 pair-spline parameters, W walkers each."""
    rn = np.random.default_rng(seed_n)
    rw = np.random.default_rng(seed_w)
    Ns = rn.choice([3, 6, 10, 20, 64, 100], size=n_systems)
    widths = rw.uniform(0.8, 1.2, size=n_systems)
    cfgs = []
    for i, (N, w) in enumerate(zip(Ns, widths)):
        cfg = trap2d_drop(int(N), float(w), W=W, seed=100 + i)
        cfg["SEED"] = 1000 + i
        cfgs.append(cfg)
    return cfgs


WORKLOADS = {
    "bulk1d_gauss": bulk1d_gauss,
    "bulk3d_lj64": bulk3d_lj64,
    "bulk3d_lj343_k256": bulk3d_lj343_k256,
    "bulk2d_gauss400_k512": bulk2d_gauss400_k512,
}
