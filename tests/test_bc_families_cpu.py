"""Boundary-condition families of the cubic pair spline (SPLINE_BC_ORIGIN /
SPLINE_BC_CUTOFF) on the CPU.

* The factor tables of the oracle (tdo_bc_factors) and of the library
  (tdvmc_amd/csrc/knots.cpp, printed by tests/cpp/knots_check bc) equal the
  reference's own SetBoundaryConditions3_1D_OR_k / _CO_k (src/SplineFactory.cpp:
  377-656, uniform = true as BosonsBulk passes it): golden bc_factors.json,
  and the compiled reference live where oracle/_ref is built.
* The oracle's map is pinned by what each family means: with c = T^T u on
  uniform knots, u'(0) = 0 (family 2), u''(0) = 0 (3), both (4), and the same
  at rc; family 1 passes the first / last three parameters through unchanged.
  T (spline_ops) and T^T (spline_coeffs) are transposes of each other, and
  BosonsBulk's pair (2, 2) gives the map of BosonsBulk.cpp:158-177 that the
  reference goldens pin.
* Config handling: the six pairs that keep N_PARAM parameters (4 boundary
  parameters in all) are accepted, the others refused.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as o

tdvmc = pytest.importorskip("tdvmc_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KC = os.path.join(ROOT, "tests", "cpp", "build", "knots_check")
ROWS = {1: 3, 2: 2, 3: 2, 4: 1}
VALID = [(a, b) for a in ROWS for b in ROWS if ROWS[a] + ROWS[b] == 4]


def _golden(golden_dir):
    with open(os.path.join(golden_dir, "bc_factors.json")) as fh:
        return {k: np.array(v) for k, v in json.load(fh)["factors"].items()}


def test_valid_pairs():
    assert sorted(VALID) == [(1, 4), (2, 2), (2, 3), (3, 2), (3, 3), (4, 1)]


@pytest.mark.parametrize("side", ["origin", "cutoff"])
@pytest.mark.parametrize("fam", [1, 2, 3, 4])
def test_oracle_factors_are_the_reference_factors(golden_dir, side, fam):
    ref = _golden(golden_dir)[f"{side}_{fam}"]
    got = o.bc_factors(fam, side == "cutoff")
    assert got.shape == ref.shape and np.array_equal(got, ref)


def test_library_factors_are_the_reference_factors(golden_dir):
    assert os.path.exists(KC), "tests/cpp/build/knots_check is not built"
    ref = _golden(golden_dir)
    res = subprocess.run([KC, "bc"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    lines = [ln.split() for ln in res.stdout.strip().split("\n")]
    assert len(lines) == 8
    for ln in lines:
        side, fam, rows = ln[0], int(ln[1]), int(ln[2])
        f = np.array([float(x) for x in ln[3:]]).reshape(rows, 3)
        assert np.array_equal(f, ref[f"{side}_{fam}"]), (side, fam)


@pytest.mark.skipif(not o.ref_available(), reason="needs oracle/_ref (built from the reference checkout)")
@pytest.mark.parametrize("P", [6, 13])
def test_factors_live_against_compiled_reference(P):
    r = o.ref()
    r.ref_bc_factors.restype = C.c_int
    dp = C.POINTER(C.c_double)
    r.ref_bc_factors.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, dp]
    rc = 2.7
    nodes = np.array([i * rc / (P - 1.0) for i in range(-3, P + 3)])
    for side in (0, 1):
        for fam in (1, 2, 3, 4):
            f = np.zeros(9)
            n = r.ref_bc_factors(fam, side, nodes.ctypes.data_as(dp), nodes.size, 1,
                                 f.ctypes.data_as(dp))
            assert np.array_equal(f.reshape(3, 3)[:n], o.bc_factors(fam, side == 1))


def _u_and_derivs(s, c, r):
    """u, u', u'' of the spline with coefficients c at r (uniform knots)."""
    h = s.rc / (s.P_s - 1)
    j, b, db, d2b = o.bspline(r, h, s.P_s)
    cc = c[j:j + 4]
    return float(cc @ b), float(cc @ db), float(cc @ d2b)


@pytest.mark.parametrize("pair", VALID)
def test_oracle_map_meets_the_boundary_conditions(pair):
    P, rc = 9, 3.0
    s = o.System(N=2, D=3, L=6.0, P_s=P, rc=rc, bc_origin=pair[0], bc_cutoff=pair[1])
    rng = np.random.default_rng(sum(pair))
    for _ in range(5):
        u = rng.normal(0, 1, P)
        c = o.spline_coeffs(s, u)
        assert c.size == P + 2
        _, d1, d2 = _u_and_derivs(s, c, 0.0)
        # just below rc: interval P_s - 2 near its end (the spline's last piece)
        e = rc * 1e-12
        _, e1, e2 = _u_and_derivs(s, c, rc - e)
        scale = np.max(np.abs(u)) * (P - 1) ** 2 / rc ** 2
        if pair[0] in (2, 4):
            assert abs(d1) < 1e-12 * scale
        if pair[0] in (3, 4):
            assert abs(d2) < 1e-12 * scale
        if pair[0] == 1:
            assert np.array_equal(c[:3], u[:3])
        if pair[1] in (2, 4):
            assert abs(e1) < 1e-9 * scale
        if pair[1] in (3, 4):
            assert abs(e2) < 1e-9 * scale
        if pair[1] == 1:
            assert np.array_equal(c[-3:], u[-3:])
        # the standard parameters in between map one to one
        n0 = ROWS[pair[0]]
        assert np.array_equal(c[3:P + 2 - 3], u[n0:n0 + P - 4])


@pytest.mark.parametrize("pair", VALID)
def test_oracle_map_and_transpose(pair):
    P = 8
    s = o.System(N=2, D=3, L=6.0, P_s=P, bc_origin=pair[0], bc_cutoff=pair[1])
    T = np.array([o.spline_ops(s, e) for e in np.eye(P + 2)]).T        # O = T B
    Tt = np.array([o.spline_coeffs(s, e) for e in np.eye(P)]).T        # c = T^T u
    assert np.array_equal(T.T, Tt)
    # the reference's parameter count: S = P_s + 2 splines, n0 + n1 = 4
    assert T.shape == (P, P + 2)


def test_bosonsbulk_pair_is_the_zero_slope_map():
    P = 7
    s = o.System(N=2, D=3, L=6.0, P_s=P)
    B = np.arange(1.0, P + 3.0) ** 2
    O = o.spline_ops(s, B)
    expect = np.concatenate([[B[1], B[0] + B[2]], B[3:P - 1], [B[P - 1] + B[P + 1], B[P]]])
    assert np.array_equal(O, expect)


BASE = {"N": 16, "DIM": 3, "LBOX": 6.0, "N_PARAM": 8, "SYSTEM_PARAMS": [0.5, 1.0]}


@pytest.mark.parametrize("pair", VALID)
def test_parse_config_accepts(pair):
    info = tdvmc.parse_config(json.dumps(dict(BASE, SPLINE_BC_ORIGIN=pair[0], SPLINE_BC_CUTOFF=pair[1])))
    assert (info.bc_origin, info.bc_cutoff) == pair and info.P == 8


def test_parse_config_default_is_bosonsbulk():
    info = tdvmc.parse_config(json.dumps(BASE))
    assert (info.bc_origin, info.bc_cutoff) == (2, 2)


@pytest.mark.parametrize("pair", [(1, 1), (1, 2), (4, 4), (2, 4), (5, 2), (2, -1)])
def test_parse_config_rejects(pair):
    with pytest.raises(tdvmc.TdvmcError) as e:
        tdvmc.parse_config(json.dumps(dict(BASE, SPLINE_BC_ORIGIN=pair[0], SPLINE_BC_CUTOFF=pair[1])))
    assert e.value.code == tdvmc.ERR_BAD_CONFIG
