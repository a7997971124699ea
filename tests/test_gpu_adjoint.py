"""GPU adjoint of ln psi (tdvmc_adjoint) and its torch.autograd.Function
wrapper: against the oracle's O_k / gradients, and against finite
differences of the library's own forward (torch.autograd.gradcheck, FP64)."""
import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import make_cfg, oracle_system, rel

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")
torch = pytest.importorskip("torch")


def _system(D=3, N=12, L=3.2, P_s=8, K=6):
    cfg = make_cfg(N=N, D=D, L=L, P_s=P_s, pot="GAUSSIAN", sysparams=(0.5, 1.0), K=K, hbar2_2m=0.5)
    s = oracle_system(cfg)
    rng = np.random.default_rng(3)
    R = rng.uniform(-L / 2, L / 2, (3, N, D))
    uR = rng.normal(0, 0.3, s.P)
    uR[P_s:] *= 0.05
    return cfg, s, R, uR


def test_adjoint_vs_oracle():
    cfg, s, R, uR = _system()
    gbar = np.array([0.7, -1.3, 0.4])
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        dR, da = ctx.adjoint(R, gbar, precision=tdvmc.FP64)
    ref = o.eval_configs(s, R, uR)
    # d ln psi / d alpha = O, plus the outer-pair count on the last pair parameter
    expect = gbar @ ref["O"]
    outer = np.zeros(len(gbar))
    for w in range(R.shape[0]):
        e = np.zeros(s.P)
        e[s.P_s - 1] = 1.0
        outer[w] = o.eval_configs(s, R[w:w + 1], uR + e)["expo"][0] - ref["expo"][w] - ref["O"][w, s.P_s - 1]
    expect[s.P_s - 1] += gbar @ outer
    assert rel(da, expect) < 1e-10
    assert rel(dR, gbar[:, None, None] * ref["grad"][..., 0]) < 1e-10


@pytest.mark.parametrize("D", [1, 3])
def test_autograd_gradcheck(D):
    from tdvmc_amd.autograd import LogPsi

    cfg, s, R, uR = _system(D=D, N=7 if D == 3 else 9, L=3.0 if D == 3 else 6.0, P_s=6, K=4)
    with tdvmc.Context(cfg) as ctx:
        Rt = torch.tensor(R[:2], dtype=torch.float64, requires_grad=True)
        ut = torch.tensor(uR, dtype=torch.float64, requires_grad=True)
        f = lambda Rx, ux: LogPsi.apply(ctx, Rx, ux, 0)  # noqa: E731
        # LDS-atomic summation order makes repeated backward calls differ in the
        # last bits; nondet_tol allows that
        assert torch.autograd.gradcheck(f, (Rt, ut), eps=1e-6, atol=1e-6, rtol=1e-5, nondet_tol=1e-12)


def test_mixed_adjoint_vs_fp64_and_oracle():
    """The mixed-form adjoint (FP32 pair arithmetic, FP64 accumulation): dR and
    dalpha against the FP64 adjoint and the oracle, at the mixed form's
    evaluation bounds (gradient 5e-5, O 2e-6 relative)."""
    cfg, s, R, uR = _system(N=27, L=4.2, P_s=10, K=12)
    rng = np.random.default_rng(9)
    R = rng.uniform(-2.1, 2.1, (8, 27, 3))
    gbar = rng.normal(0, 1, 8)
    with tdvmc.Context(cfg) as ctx:
        ctx.set_params(uR)
        dR64, da64 = ctx.adjoint(R, gbar, precision=tdvmc.FP64)
        dR32, da32 = ctx.adjoint(R, gbar, precision=tdvmc.MIXED)
    ref = o.eval_configs(s, R, uR)
    e_R = rel(dR32, gbar[:, None, None] * ref["grad"][..., 0])
    e_a = rel(da32, da64)
    print(f"mixed adjoint: dR {e_R:.2e} dalpha {e_a:.2e} (relative to FP64 / oracle)")
    assert rel(dR64, gbar[:, None, None] * ref["grad"][..., 0]) < 1e-10
    assert e_R < 5e-5
    assert e_a < 2e-6
