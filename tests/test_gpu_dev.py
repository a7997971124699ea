"""Device-pointer entry points (tdvmc_*_dev) on torch tensors: walker exchange,
measurement, sampling stage and fed-configuration evaluation with the
results left in device memory, ordered on a torch stream; each against the
host-pointer entry point of a twin context with the same inputs and seed."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import make_cfg

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")
torch = pytest.importorskip("torch")


def _cfg():
    return make_cfg(N=20, D=3, L=3.4, P_s=8, pot="GAUSSIAN", sysparams=(0.5, 1.5), K=12, hbar2_2m=0.5, W=300,
                    step=0.3, therm=20, init=10)


def _params(P):
    rng = np.random.default_rng(8)
    uR = rng.normal(0, 0.2, P)
    uI = rng.normal(0, 0.05, P)
    uR[8:] *= 0.02
    uI[8:] *= 0.02
    return uR, uI


def _p(t):
    return C.c_void_p(t.data_ptr())


def _close(a, b, tol=1e-11):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-300)


def test_dev_entry_points_match_host_entry_points():
    cfg = _cfg()
    L = tdvmc.lib()
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    print("dev entry points", flush=True)
    with tdvmc.Context(cfg) as a, tdvmc.Context(cfg) as b:
        P, N, D, W = a.P, 20, 3, 300
        uR, uI = _params(P)
        a.set_params(uR, uI, 0.1, 0.0)
        b.set_params(uR, uI, 0.1, 0.0)
        rng = np.random.default_rng(3)
        R = rng.uniform(-1.7, 1.7, (W, N, D))
        b.set_walkers(R)
        Rt = torch.tensor(R, dtype=torch.float64, device="cuda")
        mean = torch.full((P + 2,), float("nan"), dtype=torch.float64, device="cuda")
        S = torch.full((P, P), float("nan"), dtype=torch.float64, device="cuda")
        F = torch.full((2, P), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            assert L.tdvmc_set_walkers_dev(a._h, _p(Rt), W, sh) == tdvmc.OK
            assert L.tdvmc_measure_dev(a._h, tdvmc.FP64, _p(mean), _p(S), _p(F), sh) == tdvmc.OK
            # work queued on the torch stream after the call sees the results
            S2 = S * 2.0
        stream.synchronize()
        ref = b.measure(precision=tdvmc.FP64)
        assert _close(mean[:P].cpu().numpy(), ref["localOperators"])
        assert abs(mean[P].item() - ref["localEnergyR"]) <= 1e-11 * abs(ref["localEnergyR"])
        assert _close(S.cpu().numpy(), ref["S"]) and _close(S2.cpu().numpy(), 2.0 * ref["S"])
        assert _close(F[0].cpu().numpy(), ref["FR"]) and _close(F[1].cpu().numpy(), ref["FI"])

        # one sampling stage (moves + measurement + moments) with the same seed
        for prec in (tdvmc.FP64, tdvmc.MIXED):
            with torch.cuda.stream(stream):
                assert L.tdvmc_sample_dev(a._h, 77 + prec, prec, _p(mean), _p(S), _p(F), sh) == tdvmc.OK
                Rout = torch.full_like(Rt, float("nan"))
                assert L.tdvmc_get_walkers_dev(a._h, _p(Rout), W, sh) == tdvmc.OK
            stream.synchronize()
            ref = b.sample(seed=77 + prec, precision=prec)
            assert np.array_equal(Rout.cpu().numpy(), b.get_walkers()), prec
            assert _close(S.cpu().numpy(), ref["S"], 1e-10), prec
            assert _close(F[0].cpu().numpy(), ref["FR"], 1e-10), prec
            assert abs(mean[P].item() - ref["localEnergyR"]) <= 1e-10 * abs(ref["localEnergyR"]), prec
        assert a.acceptance() == b.acceptance()

        # fed configurations, results in device memory
        Rf = rng.uniform(-1.7, 1.7, (64, N, D))
        Rft = torch.tensor(Rf, dtype=torch.float64, device="cuda")
        ex = torch.zeros(64, dtype=torch.float64, device="cuda")
        er, ei = torch.zeros_like(ex), torch.zeros_like(ex)
        O = torch.zeros((64, P), dtype=torch.float64, device="cuda")
        g = torch.zeros((64, N, D, 2), dtype=torch.float64, device="cuda")
        lp = torch.zeros((64, N, 2), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert L.tdvmc_eval_configs_dev(a._h, _p(Rft), 64, tdvmc.FP64, _p(ex), _p(er), _p(ei), _p(O), _p(g), _p(lp),
                                        sh) == tdvmc.OK
        stream.synchronize()
        a.synchronize()
        host = b.eval_configs(Rf, precision=tdvmc.FP64, per_particle=True)
        assert _close(ex.cpu().numpy(), host["expo"]) and _close(O.cpu().numpy(), host["O"])
        assert _close(er.cpu().numpy(), host["ER"], 1e-10)
        assert _close(g.cpu().numpy(), host["grad"], 1e-10) and _close(lp.cpu().numpy(), host["lap"], 1e-10)


def test_dev_walkers_wrong_count_is_an_error():
    with tdvmc.Context(_cfg()) as a:
        t = torch.zeros((10, 20, 3), dtype=torch.float64, device="cuda")
        assert tdvmc.lib().tdvmc_set_walkers_dev(a._h, _p(t), 10, None) == tdvmc.ERR_BAD_ARG
