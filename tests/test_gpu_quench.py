"""Real-time TDVP after a quench (2-D bulk, N = 400, V0 = 5 -> 7 at t = 0,
48 spline + 512 k-space parameters, RK4, Tikhonov eps = 1e-4), reduced to
8192 walkers and a few steps: the mixed form and the FP64 form start from
the same walkers and parameters with the same random numbers, so their
chains part only where an accept decision differs by rounding.  Their
<E_L>(t) differ by at most 1.1e-3 of the statistical error of the two runs
over 4 steps (measured); the bound is 0.05 of it, and two independent runs
would differ by about one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")


def _run(prec, steps, W):
    from tdvmc_amd.workloads import bulk2d_gauss400_k512

    cfg = bulk2d_gauss400_k512(W=W)
    energies = []
    with tdvmc.Context(cfg) as ctx:
        info = ctx.get_info()
        assert (info.N, info.DIM, info.P, info.K) == (400, 2, 560, 512)
        ctx.sweep(2 * info.N, seed=5, precision=prec)      # same seed, same start in both forms
        for _ in range(steps):
            ctx.tdvp_step(info.timestep, ode_type=2, precision=prec, use_graph=True)
            er = ctx.eval_configs(ctx.get_walkers(), precision=prec)["ER"]
            energies.append((er.mean(), er.std() / np.sqrt(er.size)))
        uR = ctx.get_params()[0]
    return np.array(energies), uR


def test_quench_mixed_tracks_fp64():
    steps, W = 4, 8192
    print("quench fp64", flush=True)
    e64, u64 = _run(tdvmc.FP64, steps, W)
    print("quench mixed", flush=True)
    emx, umx = _run(tdvmc.MIXED, steps, W)
    assert np.all(np.isfinite(u64)) and np.all(np.isfinite(umx))
    diff = emx[:, 0] - e64[:, 0]
    sigma = np.sqrt(emx[:, 1] ** 2 + e64[:, 1] ** 2)
    print("E(t) fp64", e64[:, 0], "mixed", emx[:, 0], "diff / sigma", diff / sigma)
    assert np.all(np.abs(diff) < 0.05 * sigma)
