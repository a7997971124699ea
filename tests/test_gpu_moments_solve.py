"""GPU parity of the TDVP moments and of the FP64 solve, through the C ABI.

Moments: FP64 mode within 1e-12 (relative Frobenius) of the oracle's two-pass
FP64 moments; split-bf16 MFMA mode within 1e-6 (relative Frobenius of S and of
F) at 65536 samples with the statistics of a real measurement.
Solve: within 1e-10 of the oracle's restatement of SolveForParametersDot
(real / imaginary time, with and without diagonal scaling), and the count of
non-positive pivots on a singular system.
"""
import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu
tdvmc = pytest.importorskip("tdvmc_amd")

from tests.helpers import make_cfg  # noqa: E402


def _ctx(P_s=8, K=0):
    return tdvmc.Context(make_cfg(N=4, D=3, L=4.0, P_s=P_s, K=K))


def _samples(W, P, seed, scale=1.0, mean=0.0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(P, P)) / np.sqrt(P)
    O = mean + scale * (rng.normal(size=(W, P)) @ A + 0.3 * rng.normal(size=(W, 1)))
    ER = O @ rng.normal(size=P) * 0.1 + rng.normal(size=W)
    EI = rng.normal(size=W) * 0.2 + 0.05 * O[:, 0]
    return O, ER, EI


def fro(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("W,P", [(1000, 10), (4097, 37), (3000, 296)])
def test_moments_fp64(W, P):
    O, ER, EI = _samples(W, P, seed=P, scale=3.0, mean=50.0)
    print("moments fp64", W, P, flush=True)
    with _ctx() as ctx:
        m = ctx.moments(O, ER, EI, precision=tdvmc.FP64)
    ref = o.moments(O, ER, EI)
    assert fro(m["mO"], ref["mO"]) < 1e-13
    assert abs(m["mE"][0] - ref["mE"][0]) < 1e-12 * max(1, abs(ref["mE"][0]))
    assert fro(m["S"], ref["S"]) < 1e-12
    assert fro(m["FR"], ref["FR"]) < 1e-12
    assert fro(m["FI"], ref["FI"]) < 1e-12
    assert np.array_equal(m["S"], m["S"].T)


def test_moments_split_bf16_realistic():
    # O with the statistics of a 343-particle measurement: large means, small
    # relative fluctuations (this is where un-centred bf16 products would fail)
    W, P = 65536, 296
    O, ER, EI = _samples(W, P, seed=11, scale=30.0, mean=1000.0)
    print("moments split-bf16", W, P, flush=True)
    with _ctx() as ctx:
        m = ctx.moments(O, ER, EI, precision=tdvmc.MIXED)
        m64 = ctx.moments(O, ER, EI, precision=tdvmc.FP64)
    ref = o.moments(O, ER, EI)
    eS, eR, eI = fro(m["S"], ref["S"]), fro(m["FR"], ref["FR"]), fro(m["FI"], ref["FI"])
    print(f"split-bf16 relative Frobenius errors: S {eS:.2e}  F_R {eR:.2e}  F_I {eI:.2e}")
    # measured 3.9e-9 / 1.7e-9 / 1.5e-9 (three-term split, leading product to FP64 every 16 walkers)
    assert eS < 5e-8 and eR < 5e-8 and eI < 5e-8
    assert fro(m64["S"], ref["S"]) < 1e-12
    assert fro(m["mO"], ref["mO"]) < 1e-13


def test_moments_split_bf16_small_sample():
    W, P = 1000, 20
    O, ER, EI = _samples(W, P, seed=5, scale=2.0, mean=10.0)
    with _ctx() as ctx:
        m = ctx.moments(O, ER, EI, precision=tdvmc.MIXED)
    ref = o.moments(O, ER, EI)
    e = fro(m["S"], ref["S"])
    print(f"split-bf16 W=1000 relative Frobenius error of S: {e:.2e}")
    assert e < 1e-7   # measured 8.4e-9


@pytest.mark.parametrize("imag,pre", [(0, 0), (1, 0), (0, 1), (1, 1)])
# 16-column blocks up to P = 340 and past P = 581, 32-column blocks between;
# last block of 12, 1, 8, 8, 12 (NB = 16) and 1, 8, 16 (NB = 32) columns
@pytest.mark.parametrize("P", [12, 33, 296, 600, 1100, 353, 392, 560])
def test_solve_vs_oracle(imag, pre, P):
    rng = np.random.default_rng(P + 10 * imag + pre)
    A = rng.normal(size=(P, P))
    S = A @ A.T / P + 0.01 * np.eye(P)
    m = dict(mO=rng.normal(size=P), mE=np.array([1.3, -0.4]), S=S, FR=rng.normal(size=P), FI=rng.normal(size=P))
    print("solve", P, imag, pre, flush=True)
    with _ctx() as ctx:
        r = ctx.solve(m["mO"], m["mE"], S, m["FR"], m["FI"], imag_time=imag, precondition=pre, eps=1e-3)
    ref = o.solve_param_dot(m, imag_time=imag, precondition=pre, eps=1e-3)
    assert r["bad"] == 0 and ref["bad"] == 0
    assert fro(r["xR"], ref["xR"]) < 1e-10
    assert fro(r["xI"], ref["xI"]) < 1e-10
    assert abs(r["phiDotR"] - ref["phiDotR"]) < 1e-10 * max(1, abs(ref["phiDotR"]))
    assert abs(r["phiDotI"] - ref["phiDotI"]) < 1e-10 * max(1, abs(ref["phiDotI"]))


def test_solve_too_large_is_an_error():
    """Past the largest system the LDS panel holds (P = 1129), the solve
    reports an error instead of launching."""
    P = 1200
    rng = np.random.default_rng(3)
    S = np.eye(P)
    m = dict(mO=rng.normal(size=P), mE=np.array([0.3, 0.0]), S=S, FR=rng.normal(size=P), FI=np.zeros(P))
    with _ctx() as ctx:
        with pytest.raises(tdvmc.TdvmcError):
            ctx.solve(m["mO"], m["mE"], S, m["FR"], m["FI"], imag_time=1, eps=1e-3)


@pytest.mark.parametrize("P,bad", [(70, (40, 66)), (390, (40, 388))])
def test_solve_nonpositive_pivot_in_later_panel(P, bad):
    """A non-positive pivot in a later panel (the third of 16 columns, the
    second of 32) and in the partial last one is counted and handled as in
    the reference."""
    rng = np.random.default_rng(7)
    A = rng.normal(size=(P, P))
    S = A @ A.T / P + 0.5 * np.eye(P)
    for i in bad:
        S[i, :] = 0.0
        S[:, i] = 0.0
        S[i, i] = -0.5
    m = dict(mO=rng.normal(size=P), mE=np.array([0.3, 0.0]), S=S, FR=rng.normal(size=P), FI=np.zeros(P))
    with _ctx() as ctx:
        r = ctx.solve(m["mO"], m["mE"], S, m["FR"], m["FI"], imag_time=1, eps=1e-3, check=False)
    ref = o.solve_param_dot(m, imag_time=1, eps=1e-3)
    assert r["bad"] == ref["bad"] == 2
    assert np.allclose(r["xR"], ref["xR"], rtol=1e-10, atol=1e-12)


def test_solve_nonpositive_pivot_flagged():
    P = 6
    S = np.zeros((P, P))
    S[0, 0] = 1.0
    S[1, 1] = -0.5          # negative after the +eps shift
    for i in range(2, P):
        S[i, i] = 2.0
    m = dict(mO=np.zeros(P), mE=np.array([0.5, 0.0]), S=S, FR=np.ones(P), FI=np.zeros(P))
    with _ctx() as ctx:
        r = ctx.solve(m["mO"], m["mE"], S, m["FR"], m["FI"], imag_time=1, eps=1e-3, check=False)
    ref = o.solve_param_dot(m, imag_time=1, eps=1e-3)
    assert r["bad"] == ref["bad"] == 1
    assert np.allclose(r["xR"], ref["xR"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("imag,pre", [(0, 0), (1, 1)])
@pytest.mark.parametrize("P", [12, 296])
def test_solve_type1_vs_oracle(imag, pre, P):
    """Solver type 1: the GPU's Cholesky solve of the same SPD system against
    the oracle's full-pivot QR, then the mean subtraction."""
    rng = np.random.default_rng(3 * P + 10 * imag + pre)
    A = rng.normal(size=(P, P))
    S = A @ A.T / P + 0.05 * np.eye(P)
    m = dict(mO=rng.normal(size=P), mE=np.array([0.9, 0.3]), S=S, FR=rng.normal(size=P), FI=rng.normal(size=P))
    print("solve type 1", P, imag, pre, flush=True)
    with _ctx() as ctx:
        r = ctx.solve(m["mO"], m["mE"], S, m["FR"], m["FI"], imag_time=imag, precondition=pre, solver_type=1)
    ref = o.solve_param_dot(m, imag_time=imag, precondition=pre, solver_type=1)
    assert r["bad"] == 0 and ref["bad"] == 0
    assert fro(r["xR"], ref["xR"]) < 1e-10
    assert fro(r["xI"], ref["xI"]) < 1e-10
    assert abs(r["phiDotR"] - ref["phiDotR"]) < 1e-10 * max(1, abs(ref["phiDotR"]))
    assert abs(r["phiDotI"] - ref["phiDotI"]) < 1e-10 * max(1, abs(ref["phiDotI"]))


@pytest.mark.parametrize("P,ndup", [(24, 4), (296, 128)])
def test_solve_type1_singular_pm_k(P, ndup):
    """Solver type 1 on an S made singular by identical O columns (k and -k
    share |rho_k|^2): the GPU's full-pivot QR and the oracle's pick the same
    solution (rank P - ndup)."""
    rng = np.random.default_rng(P)
    W = 3000
    base = rng.normal(size=(W, P - ndup)) * rng.uniform(0.5, 3.0, P - ndup)
    cols = list(range(P - ndup))
    twins = rng.choice(P - ndup, ndup, replace=False)
    O = np.empty((W, P))
    order = rng.permutation(P)          # scatter the twins among the columns
    src = cols + list(twins)
    for dst, c in zip(order, src):
        O[:, dst] = base[:, c]
    E = O @ rng.normal(size=P) * 0.1 + rng.normal(size=W)
    m = o.moments(O, E, 0.3 * E)
    print("solve type 1 singular", P, ndup, flush=True)
    with _ctx() as ctx:
        r = ctx.solve(m["mO"], m["mE"], m["S"], m["FR"], m["FI"], imag_time=1, precondition=0, solver_type=1)
    ref = o.solve_param_dot(m, imag_time=1, precondition=0, solver_type=1)
    assert ref["bad"] == ndup
    assert fro(r["xR"], ref["xR"]) < 1e-8
    assert fro(r["xI"], ref["xI"]) < 1e-8
    assert abs(r["phiDotR"] - ref["phiDotR"]) < 1e-8 * max(1, abs(ref["phiDotR"]))


@pytest.mark.parametrize("solver,pre", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("imag", [0, 1])
@pytest.mark.parametrize("name", ["p12", "p296_twins"])
def test_solve_vs_reference_driver(golden_dir, name, imag, solver, pre):
    """The GPU solve against the reference driver's own SolveForParametersDot
    + CalculatePhiDot (compiled reference, tests/golden/tdvp_cases.npz), on
    the driver's moment sums: a generic P = 12 set, and a P = 296 set of a
    bulk system whose +-k twins make S exactly singular (solver type 1 then
    runs at rank 168 with solutions up to 1.7e10 without the scaling)."""
    import os

    z = np.load(os.path.join(golden_dir, "tdvp_cases.npz"), allow_pickle=False)
    g = {k.split("__")[1]: z[k] for k in z.files if k.startswith(name + "__")}
    mO, mE = g["mO"], g["mE"]
    S = g["OO"] - np.outer(mO, mO)
    FR, FI = g["OER"] - mE[0] * mO, g["OEI"] - mE[1] * mO
    key = f"i{imag}_s{solver}_p{pre}"
    print("solve vs reference driver", name, key, flush=True)
    with _ctx() as ctx:
        r = ctx.solve(mO, mE, S, FR, FI, imag_time=imag, precondition=pre, eps=1e-3, solver_type=solver,
                      check=False)
    xR, xI, pd = g["xR_" + key], g["xI_" + key], g["phidot_" + key]
    scale = max(np.max(np.abs(xR)), np.max(np.abs(xI)))
    tol = 1e-7 if (name == "p296_twins" and solver == 1 and pre == 0) else 1e-10
    assert np.max(np.abs(r["xR"] - xR)) <= tol * scale
    assert np.max(np.abs(r["xI"] - xI)) <= tol * scale
    sR = np.abs(mO) @ np.abs(xR) + abs(mE[0])
    sI = np.abs(mO) @ np.abs(xI) + abs(mE[0])
    assert abs(r["phiDotR"] - pd[0]) <= tol * sR
    assert abs(r["phiDotI"] - pd[1]) <= tol * sI
