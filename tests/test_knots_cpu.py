"""Non-uniform pair-spline knots (USE_NURBS / NURBS_GRID) on the CPU.

* The oracle's B-spline basis on non-uniform knots (tdo_bspline_nu, Cox-de
  Boor) is pinned by mathematics: partition of unity, derivatives against
  central differences, and on a uniform grid the closed-form uniform basis.
  (Against the reference's own NURBS evaluation: eval_nurbs_*.npz in
  test_oracle_golden.py.)
* The library's lookup tables (tdvmc_amd/csrc/knots.cpp), looked up the way
  the kernels do by tests/cpp/knots_check, give the oracle's interval and
  basis values, with the reference's lower_bound tie rule at the knots, at rc
  and beyond it.
* Config handling: accepted grids set rc to the last knot; bad grids are
  rejected with TDVMC_ERR_BAD_CONFIG.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as o

tdvmc = pytest.importorskip("tdvmc_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KC = os.path.join(ROOT, "tests", "cpp", "build", "knots_check")


def _grid(P, rc, rng, a=1.3):
    g = rc * (np.arange(P) / (P - 1)) ** a
    gaps = np.diff(g)
    g[1:-1] += rng.uniform(-0.2, 0.2, P - 2) * np.minimum(gaps[:-1], gaps[1:])
    return g


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_basis_partition_of_unity_and_derivatives(seed):
    rng = np.random.default_rng(seed)
    P, rc = int(rng.integers(5, 16)), 4.0
    g = _grid(P, rc, rng)
    h = 1e-5
    for r in rng.uniform(0.02, rc - 0.02, 60):
        j, b, db, d2b = o.bspline_nu(r, g, P)
        assert abs(b.sum() - 1.0) < 1e-13
        assert np.all(b >= -1e-15)
        assert abs(db.sum()) < 1e-11 and abs(d2b.sum()) < 1e-9
        # central differences of the four pieces of interval j (same interval on both sides)
        if g[j] + h < r < g[j + 1] - h:
            _, bp, dbp, _ = o.bspline_nu(r + h, g, P)
            _, bm, dbm, _ = o.bspline_nu(r - h, g, P)
            assert np.max(np.abs((bp - bm) / (2 * h) - db)) < 1e-6 * max(1.0, np.max(np.abs(db)))
            assert np.max(np.abs((dbp - dbm) / (2 * h) - d2b)) < 1e-5 * max(1.0, np.max(np.abs(d2b)))


def test_oracle_basis_uniform_grid_is_the_uniform_basis():
    P, rc = 11, 5.0
    g = np.linspace(0.0, rc, P)
    for r in np.random.default_rng(4).uniform(0.0, rc, 200):
        j, b, db, d2b = o.bspline_nu(r, g, P)
        ju, bu, dbu, d2bu = o.bspline(r, rc / (P - 1), P)
        if j != ju:
            # a knot hit exactly: the two rules put r in neighbouring intervals
            continue
        assert np.max(np.abs(b - bu)) < 1e-13
        assert np.max(np.abs(db - dbu)) < 1e-11
        assert np.max(np.abs(d2b - d2bu)) < 1e-9


def _run_kc(g, rmax, r):
    f = lambda a: " ".join(repr(float(x)) for x in a)
    inp = f"{len(g)}\n{f(g)}\n{rmax}\n{len(r)}\n{f(r)}\n"
    res = subprocess.run([KC], input=inp, capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    return res.stdout.split("\n")


@pytest.mark.parametrize("seed", [5, 6, 7, 8])
def test_library_tables_vs_oracle_basis(seed):
    assert os.path.exists(KC), "tests/cpp/build/knots_check is not built"
    rng = np.random.default_rng(seed)
    P, rc = int(rng.integers(4, 30)), float(rng.uniform(1.0, 10.0))
    g = _grid(P, rc, rng, a=float(rng.uniform(0.7, 1.8)))
    # random distances, every knot exactly, rc, just beyond rc, far beyond, 0
    r = np.concatenate([rng.uniform(0, rc, 500), g, [rc, np.nextafter(rc, 2 * rc), 1.5 * rc, 0.0]])
    out = _run_kc(g, rc, r)
    assert out[0].startswith("ncell")
    for k, line in enumerate(out[1:1 + r.size]):
        v = line.split()
        j, vals = int(v[0]), np.array([float(x) for x in v[1:]])
        b, db = vals[:4], vals[4:]
        if r[k] > rc:
            assert j == P - 1, (r[k], j)          # beyond rc: the constant outer piece
            continue
        jo, bo, dbo, _ = o.bspline_nu(r[k], g, P)
        assert j == jo, (r[k], j, jo)
        assert np.max(np.abs(b - bo)) < 1e-13, (r[k], b, bo)
        assert np.max(np.abs(db - dbo)) < 1e-12 * max(1.0, np.max(np.abs(dbo))), (r[k], db, dbo)


def test_library_tables_reject_bad_grids():
    rc = 3.0
    assert _run_kc([0.0, 1.0, 0.5, 3.0], rc, [])[0].startswith("error")     # not increasing
    assert _run_kc([0.1, 1.0, 2.0, 3.0], rc, [])[0].startswith("error")     # does not start at 0
    assert _run_kc([0.0, 1.0, 2.0, 3.5], rc, [])[0].startswith("error")     # beyond half the box
    fine = np.concatenate([[0.0, 1e-3], np.linspace(0.5, 3.0, 6)])          # > 511 cells of the spacing
    assert _run_kc(fine, rc, [])[0].startswith("error")


BASE = {"N": 16, "DIM": 3, "LBOX": 6.0, "N_PARAM": 8, "SYSTEM_PARAMS": [0.5, 1.0]}


def test_parse_config_nurbs():
    g = [0.0, 0.2, 0.5, 0.9, 1.4, 2.0, 2.5, 2.8]
    info = tdvmc.parse_config(json.dumps(dict(BASE, USE_NURBS=1, NURBS_GRID=g)))
    assert info.use_nurbs == 1 and info.rc == 2.8 and info.P == 8
    info = tdvmc.parse_config(json.dumps(BASE))
    assert info.use_nurbs == 0 and info.rc == 3.0


@pytest.mark.parametrize("grid", [
    [0.0, 0.2, 0.5],                                   # fewer knots than N_PARAM
    [0.0, 0.2, 0.5, 0.9, 1.4, 1.4, 2.5, 2.8],          # repeated knot
    [0.0, 0.2, 0.5, 0.9, 1.4, 2.0, 2.5, 3.2],          # beyond L/2
    [0.1, 0.2, 0.5, 0.9, 1.4, 2.0, 2.5, 2.8],          # not starting at 0
])
def test_parse_config_nurbs_rejects(grid):
    with pytest.raises(tdvmc.TdvmcError) as e:
        tdvmc.parse_config(json.dumps(dict(BASE, USE_NURBS=1, NURBS_GRID=grid)))
    assert e.value.code == tdvmc.ERR_BAD_CONFIG


def test_parse_config_nurbs_needs_periodic_box():
    cfg = dict(BASE, LBOX=0.0, ONE_BODY=1, TRAP_OMEGA=1.0, SPLINE_CUTOFF=3.0, USE_NURBS=1,
               NURBS_GRID=[0.0, 0.2, 0.5, 0.9, 1.4, 2.0, 2.5, 2.8])
    with pytest.raises(tdvmc.TdvmcError):
        tdvmc.parse_config(json.dumps(cfg))
