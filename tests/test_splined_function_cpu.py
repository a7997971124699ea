"""tdvmc::SplinedFunction (include/tdvmc_splined_function.hpp), the pair spline
in the shape of the reference's WFParts::SplinedFunction
(src/WFParts/SplinedFunction.h:10-40): its knot vector, its piece table
splineWeights[s][p][c] and its boundary-condition factors against the
reference's own SplineFactory::GetWeights and SetBoundaryConditions3_1D_OR_2 /
_CO_2 output on the same knots (tests/golden/spline_weights.json), for
uniform knots and for NURBS_GRID knots; and the pieces against the oracle's
basis (partition of unity across the pieces of one interval)."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SC = os.path.join(ROOT, "tests", "cpp", "build", "splined_check")


def _run(P, rc, grid):
    g = [] if grid is None else list(grid)
    inp = f"{P} {rc!r} {len(g)}\n" + " ".join(repr(float(x)) for x in g) + "\n"
    r = subprocess.run([SC], input=inp, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {"w": {}}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        if k == "w":
            out["w"][(int(v[0]), int(v[1]))] = np.array([float(x) for x in v[2:]])
        else:
            out[k] = np.array([float(x) for x in v])
    return out


def _cases(golden_dir):
    with open(os.path.join(golden_dir, "spline_weights.json")) as fh:
        return json.load(fh)["cases"]


def test_view_matches_reference_tables(golden_dir):
    assert os.path.exists(SC), "tests/cpp/build/splined_check is not built"
    for c in _cases(golden_dir):
        out = _run(c["P"], c["rc"], c["grid"])
        nodes = np.array(c["nodes"])
        assert np.max(np.abs(out["nodes"] - nodes)) <= 1e-15 * np.max(np.abs(nodes))
        W = np.array(c["weights"])
        S = W.shape[0]
        assert int(out["np"][4]) == S == c["P"] + 2
        for s in range(S):
            for p in range(4):
                # the reference's closed forms and the recursion here differ by
                # round-off, relative to the largest term of the piece at its knots
                a, b = nodes[s + p], nodes[s + p + 1]
                scale = np.max(np.abs(W[s, p]) * np.max(np.abs([a, b])) ** np.arange(4)) + 1e-300
                assert np.max(np.abs(out["w"][(s, p)] - W[s, p])) <= 1e-11 * scale, (c["P"], s, p)
        assert np.array_equal(out["bcs"], np.array(c["bcFactorsStart"]).ravel())
        assert np.array_equal(out["bce"], np.array(c["bcFactorsEnd"]).ravel())
        np1, np2, np3, nstd = (int(x) for x in out["np"][:4])
        assert (np1, np2, np3, nstd) == (2, c["P"] - 2, c["P"], c["P"] - 4)


def test_view_pieces_are_the_oracle_basis(golden_dir):
    for c in _cases(golden_dir):
        out = _run(c["P"], c["rc"], c["grid"])
        P, rc = c["P"], c["rc"]
        nodes = out["nodes"]
        for r in np.random.default_rng(P).uniform(0.0, rc, 50):
            if c["grid"] is None:
                j, b, _, _ = o.bspline(r, rc / (P - 1), P)
            else:
                j, b, _, _ = o.bspline_nu(r, np.array(c["grid"]), P)
            e = j + 3                                  # extended interval of grid interval j
            for m in range(4):
                s, p = j + m, e - (j + m)              # spline j + m, its piece on interval e
                w = out["w"][(s, p)]
                assert abs(w[0] + r * (w[1] + r * (w[2] + r * w[3])) - b[m]) < 1e-9, (P, r, m)
            assert nodes[e] <= r <= nodes[e + 1] + 1e-15
