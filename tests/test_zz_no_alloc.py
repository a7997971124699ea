"""tdvmc_adjoint_noalloc (the allocation-free, device-pointer adjoint) against
tdvmc_adjoint, called through the C ABI on torch device buffers: two calls in
a row, the second with fewer walkers (the fixed scratch keeps the first call's
contents), in both precisions; and the size limit of its scratch."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import make_cfg, oracle_system

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")
torch = pytest.importorskip("torch")


def _cfg():
    return make_cfg(N=27, D=3, L=3.6, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=10, hbar2_2m=0.5)


def _noalloc(ctx, R, gbar, prec):
    """One call on device buffers; outputs pre-filled with NaN so that an entry
    the call leaves unwritten fails the comparison."""
    W = R.shape[0]
    Rt = torch.tensor(R, dtype=torch.float64, device="cuda")
    gt = torch.tensor(gbar, dtype=torch.float64, device="cuda")
    dRt = torch.full(R.shape, float("nan"), dtype=torch.float64, device="cuda")
    dat = torch.full((ctx.P,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = tdvmc.lib().tdvmc_adjoint_noalloc(ctx._h, C.c_void_p(Rt.data_ptr()), W, C.c_void_p(gt.data_ptr()),
                                           int(prec), C.c_void_p(dRt.data_ptr()), C.c_void_p(dat.data_ptr()), None)
    assert rc == tdvmc.OK, tdvmc.lib().tdvmc_last_error(ctx._h)
    ctx.synchronize()
    return dRt.cpu().numpy(), dat.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_adjoint_matches_adjoint(prec):
    cfg = _cfg()
    s = oracle_system(cfg)
    rng = np.random.default_rng(21)
    uR = rng.normal(0, 0.3, s.P)
    uR[8:] *= 0.05
    uI = rng.normal(0, 0.05, s.P)
    print("noalloc adjoint", prec, flush=True)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR, uI)
        for W in (96, 37):                      # the second call is smaller
            R = rng.uniform(-1.8, 1.8, (W, 27, 3))
            gbar = rng.normal(0.0, 1.0, W)
            dR_ref, da_ref = ctx.adjoint(R, gbar, precision=prec)
            dR, da = _noalloc(ctx, R, gbar, prec)
            assert np.all(np.isfinite(dR)) and np.all(np.isfinite(da)), W
            # same kernels on the same inputs: only the order of the LDS atomic
            # FP64 sums may differ between launches
            assert np.allclose(dR, dR_ref, rtol=1e-11, atol=1e-12 * np.max(np.abs(dR_ref))), W
            assert np.allclose(da, da_ref, rtol=1e-11, atol=1e-12 * np.max(np.abs(da_ref))), W


def test_noalloc_adjoint_rejects_oversize():
    cfg = _cfg()
    with tdvmc.Context(cfg) as ctx:
        P, N, D = ctx.P, 27, 3
        per = (P + 2 * N * D + 2 * N + 4) * 8
        wmax = (64 << 20) // per
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        p = C.c_void_p(t.data_ptr())
        rc = tdvmc.lib().tdvmc_adjoint_noalloc(ctx._h, p, wmax + 1, p, 0, p, p, None)
        assert rc == tdvmc.ERR_BAD_ARG
        assert b"scratch" in tdvmc.lib().tdvmc_last_error(ctx._h)
