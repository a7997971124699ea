"""Pins the CPU oracle against the reference's own outputs.

* minimum image: the 48 known answers of src/test/Tests.h:76-156 (tolerance
  1e-9, as there) and 200 random displacements evaluated by the reference;
* BosonsBulk evaluation: fed configurations evaluated by the compiled
  reference plug-in (tests/golden/eval_*.npz, written by make_golden.py):
  exponent, O_k, E_R, E_I, the nine 'other' values, and the per-particle
  gradient / Laplacian obtained by contracting the reference's sD / sD2
  arrays with the spline coefficients;
* single-particle move ratio (CalculateWFQuotient) against a full recompute.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as o


def test_nic_known_answers(golden_dir):
    data = json.load(open(os.path.join(golden_dir, "nic_cases.json")))
    assert len(data["known"]) == 48
    for c in data["known"]:
        d, v = o.displacement(c["ri"], c["rj"], c["L"])
        assert abs(d - c["dist"]) < 1e-9, c
        assert np.all(np.abs(v - np.array(c["disp"])) < 1e-9), c


def test_nic_random_vs_reference(golden_dir):
    data = json.load(open(os.path.join(golden_dir, "nic_cases.json")))
    for c in data["random"]:
        d, v = o.displacement(c["ri"], c["rj"], c["L"])
        assert abs(d - c["dist"]) <= 1e-12 * max(1.0, c["dist"])
        assert np.allclose(v, c["disp"], rtol=0, atol=1e-12)


def _case(path):
    z = np.load(path, allow_pickle=False)
    s = o.System(N=int(z["N"]), D=int(z["D"]), L=float(z["L"]), P_s=int(z["P"]), pot_type=o.POT_SQUARE,
                 pot=tuple(float(x) for x in z["sysparams"]), hbar2_2m=1.0)
    return z, s


CASES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "eval_*.npz")))


def _tol(P):
    # The reference evaluates each cubic piece in the monomial basis of the
    # absolute distance (coefficients ~ (knot index)^3), which carries a
    # relative roundoff of order 1e-16 * P^3 * (number of pairs)^0.5; below
    # P = 16 that is under 1e-11.  Larger cases are also checked against a
    # 40-digit evaluation in test_oracle_math.py.
    return 1e-10 if P <= 16 else 2e-7


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p) for p in CASES])
def test_eval_vs_reference(path):
    z, s = _case(path)
    out = o.eval_configs(s, z["R"], z["uR"], z["uI"], phiR=float(z["phiR"]), time=float(z["time"]))
    tol = _tol(s.P_s)

    def rel(a, b):
        return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)

    assert rel(out["expo"], z["expo"]) < tol
    assert rel(out["O"], z["O"]) < tol
    assert rel(out["ER"], z["ER"]) < tol
    assert rel(out["EI"], z["EI"]) < tol
    for k in (0, 1, 3, 4, 5, 6, 7, 8):
        assert rel(out["other"][:, k], z["other"][:, k]) < tol, k
    # other[2] = wf = exp(exponent + phiR): relative error is the absolute
    # error of the exponent
    assert np.max(np.abs(np.log(out["other"][:, 2]) - np.log(z["other"][:, 2]))) < tol * np.max(np.abs(z["expo"])) + 1e-12

    # gradient / Laplacian: contract the reference's sD / sD2 with c = T^T u
    P, S = s.P_s, s.P_s + 2
    cR, cI = np.zeros(S), np.zeros(S)
    for c, u in ((cR, z["uR"]), (cI, z["uI"])):
        c[1] += u[0]
        c[0] += u[1]
        c[2] += u[1]
        for k in range(2, P - 2):
            c[k + 1] += u[k]
        c[P - 1] += u[P - 2]
        c[P + 1] += u[P - 2]
        c[P] += u[P - 1]
    gR = np.einsum("s,wsna->wna", cR, z["sD"])
    gI = np.einsum("s,wsna->wna", cI, z["sD"])
    lR = np.einsum("s,wsn->wn", cR, z["sD2"])
    lI = np.einsum("s,wsn->wn", cI, z["sD2"])
    assert rel(out["grad"][..., 0], gR) < tol
    assert rel(out["grad"][..., 1], gI) < tol
    assert rel(out["lap"][..., 0], lR) < tol
    assert rel(out["lap"][..., 1], lI) < tol


def test_move_ratio_vs_reference(golden_dir):
    rows = json.load(open(os.path.join(golden_dir, "move_cases.json")))
    for c in rows:
        N, D, L, P = c["N"], c["D"], c["L"], c["P"]
        s = o.System(N=N, D=D, L=L, P_s=P)
        R = np.array(c["R"]).reshape(N, D)
        old = o.eval_configs(s, R, c["uR"])["expo"][0]
        R[c["p"]] = c["newpos"]
        new = o.eval_configs(s, R, c["uR"])["expo"][0]
        tol = 1e-10 * max(1.0, abs(new))
        assert abs(new - c["exponent_new"]) < tol
        q = np.exp(2.0 * (new - old))
        assert abs(q - c["quotient"]) <= 1e-9 * max(1.0, q)
