"""The C++ layer (include/tdvmc.hpp) driven by a C++ program the way the
reference's driver drives IPhysicalSystem: per-configuration
CalculateExpectationValues, a CalculateWFQuotient / AcceptMove pair and one
batched measurement into SimulationStepData, checked against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as o
from tests.helpers import make_cfg, oracle_system, rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "adapter_check")


def _run(tmp_path, cfg, R, uR, uI, phi, move):
    W, N, D = R.shape
    P = uR.size
    cpath, ipath = tmp_path / "sys.config", tmp_path / "in.txt"
    cpath.write_text(json.dumps(cfg))
    with open(ipath, "w") as f:
        f.write(f"{W} {N} {D} {P}\n")
        f.write(" ".join(repr(float(v)) for v in uR) + "\n")
        f.write(" ".join(repr(float(v)) for v in uI) + "\n")
        f.write(f"{phi[0]!r} {phi[1]!r}\n")
        f.write(" ".join(repr(float(v)) for v in R.ravel()) + "\n")
        f.write(f"{move[0]} " + " ".join(repr(float(v)) for v in move[1]) + "\n")
    print("adapter_check", W, N, D, P, flush=True)
    r = subprocess.run([BIN, str(cpath), str(ipath)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, i, v = line.split()
        out.setdefault(k, {})[int(i)] = float(v)
    return {k: np.array([d[i] for i in sorted(d)]) for k, d in out.items()}


def test_cpp_adapter_vs_oracle(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/build/adapter_check is not built"
    N, D, L, Ps, K = 12, 3, 3.6, 8, 6
    cfg = make_cfg(N=N, D=D, L=L, P_s=Ps, pot="GAUSSIAN", sysparams=(0.6, 1.5), K=K, hbar2_2m=0.5, W=5)
    s = oracle_system(cfg)
    rng = np.random.default_rng(21)
    W = 5
    R = rng.uniform(-L / 2, L / 2, (W, N, D))
    uR = rng.normal(0, 0.2, s.P)
    uI = rng.normal(0, 0.05, s.P)
    uR[Ps:] *= 0.05
    move = (4, R[0, 4] + np.array([0.11, -0.07, 0.05]))
    res = _run(tmp_path, cfg, R, uR, uI, (0.3, -0.1), move)
    ref = o.eval_configs(s, R, uR, uI)
    assert rel(res["expo"], ref["expo"]) < 1e-10
    assert rel(res["ER"], ref["ER"]) < 1e-10
    assert rel(res["EI"], ref["EI"]) < 1e-9
    assert rel(res["O"], ref["O"].ravel()) < 1e-10
    assert rel(res["OO"], ref["O"][:, -1] * ref["O"][:, 0]) < 1e-10
    R2 = R[:1].copy()
    R2[0, move[0]] = move[1]
    ref2 = o.eval_configs(s, R2, uR, uI)
    q = np.exp(2 * (ref2["expo"][0] - ref["expo"][0]))
    assert abs(res["quotient"][0] - q) < 1e-10 * q
    assert abs(res["exponent_after_accept"][0] - ref2["expo"][0]) < 1e-10 * max(1, abs(ref2["expo"][0]))
    assert abs(res["mean_ER"][0] - ref["ER"].mean()) < 1e-10 * max(1, abs(ref["ER"].mean()))
    assert rel(res["mean_O"], ref["O"].mean(0)) < 1e-10
    assert rel(res["S_diag"], ref["O"].var(0)) < 1e-9
    assert int(res["n_samples"][0]) == W
    assert int(res["error_caught"][0]) == 1       # TDVMC_ERR_BAD_ARG surfaced as tdvmc::Error


IPHYS = os.path.join(ROOT, "tests", "cpp", "build", "iphys_check")


def test_iphysicalsystem_adapter_vs_reference_bosonsbulk(tmp_path):
    """PhysicalSystems::GpuBulkBosons (include/tdvmc_iphysicalsystem.hpp)
    against the reference's own BosonsBulk, both driven through the
    reference's IPhysicalSystem* by tests/cpp/iphys_check.cpp (built with the
    reference's headers and its compiled plug-in, oracle/_ref/libref.so):
    wave function and exponent, a Metropolis chain of trial quotients and
    accepted moves, expectation values and their products, the nine 'other'
    values, g(r) and S(k) in the plug-ins' observables, and the
    correlated-sampling overloads."""
    assert os.path.exists(IPHYS), "tests/cpp/build/iphys_check is not built (make -C oracle iphys)"
    rng = np.random.default_rng(4)
    N, D, L, P = 27, 3, 4.5, 10
    cfg = make_cfg(N=N, D=D, L=L, P_s=P, pot="SQUARE_WELL", sysparams=(0.7, 2.0), W=1, GR_BIN_COUNT=40,
                   PARAMS_REAL=rng.normal(0, 0.3, P).tolist(), PARAMS_IMAGINARY=rng.normal(0, 0.2, P).tolist())
    path = tmp_path / "sys.config"
    path.write_text(json.dumps(cfg))
    print("iphys_check", flush=True)
    r = subprocess.run([IPHYS, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        out[k] = [float(x) for x in v]
    print(r.stdout)
    assert 0 < out["accepted"][0] < out["moves"][0]
    # the reference's incremental sums (clamped, src/PhysicalSystems/BosonsBulk.cpp:617-621)
    # drift from a full evaluation by round-off only
    for k in ("err_expo", "err_wf", "err_quotient", "err_O", "err_OO", "err_OE", "err_ER", "err_EI", "err_other"):
        assert out[k][0] < 1e-9, k
    assert out["gr_bins"][0] == out["gr_bins"][1] > 0
    assert out["gr_grid_err"][0] < 1e-14
    assert out["err_gr"][0] < 1e-12 and out["err_sk"][0] < 1e-10
    assert out["err_cs"][0] < 1e-12 and out["com"][0] == 0.0
    assert out["nic"][0] == 1 and out["normalization"][0] == 1 and out["n_other"][0] == 9
