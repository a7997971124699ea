"""The reference driver's output files (tdvmc_files_*), byte for byte against
the reference's own writers WriteDataToFile / AppendDataToFile
(src/Utils.cpp:818-1016, compiled into oracle/_ref) on the same call sequence
as src/TDVMC.cpp:3117-3129 (start), :3833-3839 (per step), :3151-3157 (grids)
and :3475-3500 (observables).  The values include signed zeros, subnormals,
huge and tiny magnitudes and infinities.  Host code only."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as o

tdvmc = pytest.importorskip("tdvmc_amd")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


VALUES = np.array([0.0, -0.0, 1.5e-300, -2.25e300, np.pi, 1e-5, -123456.789, 4.9e-324, 7.0, np.inf, -np.inf,
                   -1.0 / 3.0])


def _lib_sequence(d, P, uR, uI, steps, grids, obs):
    L = tdvmc.lib()
    assert L.tdvmc_files_init(d.encode(), _dp(uR), _dp(uI), P) == 0
    for (t, eR, eI, O, other, pR, pI) in steps:
        sd = tdvmc.StepData()
        sd.localEnergyR, sd.localEnergyI = eR, eI
        sd.localOperators, sd.otherExpectationValues = _dp(O), _dp(other)
        assert L.tdvmc_files_append_step(d.encode(), t, C.byref(sd), _dp(pR), _dp(pI), P) == 0
    gg, sg = grids
    assert L.tdvmc_files_write_grids(d.encode(), _dp(gg), len(gg), _dp(sg), len(sg)) == 0
    for (t, gr, sk) in obs:
        assert L.tdvmc_files_append_observables(d.encode(), t, _dp(gr), len(gr), _dp(sk), len(sk)) == 0


def _ref_sequence(d, P, uR, uI, steps, grids, obs):
    r = o.ref()
    for f in ("ref_write_vector", "ref_append_vector", "ref_append_scalar"):
        getattr(r, f).restype = C.c_int
    r.ref_write_vector.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_char_p]
    r.ref_append_vector.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_int]
    r.ref_append_scalar.argtypes = [C.c_char_p, C.c_char_p, C.c_double]
    b = d.encode()
    empty = np.zeros(1)
    r.ref_write_vector(b, b"ParametersR0", _dp(uR), P, b"uR_k...")
    r.ref_write_vector(b, b"ParametersI0", _dp(uI), P, b"uI_k...")
    for name, head in [("LocalEnergyR", "E^R"), ("LocalEnergyI", "E^I"), ("LocalOperators", "<O_k>..."),
                       ("OtherExpectationValues", "kinetic, potential, wf, g(r)"), ("ParametersR", "uR_k..."),
                       ("ParametersI", "uI_k..."), ("times", "t")]:
        r.ref_write_vector(b, name.encode(), _dp(empty), 0, head.encode())
    for (t, eR, eI, O, other, pR, pI) in steps:
        r.ref_append_scalar(b, b"timesSystem", t)
        r.ref_append_scalar(b, b"LocalEnergyR", eR)
        r.ref_append_scalar(b, b"LocalEnergyI", eI)
        r.ref_append_vector(b, b"LocalOperators", _dp(O), P)
        r.ref_append_vector(b, b"OtherExpectationValues", _dp(other), 9)
        r.ref_append_vector(b, b"ParametersR", _dp(pR), P)
        r.ref_append_vector(b, b"ParametersI", _dp(pI), P)
    gg, sg = grids
    r.ref_write_vector(b, b"gr_grid", _dp(gg), len(gg), b"r_ij")
    r.ref_write_vector(b, b"sk_grid", _dp(sg), len(sg), b"k")
    for (t, gr, sk) in obs:
        r.ref_append_scalar(b, b"timesAdditional", t)
        r.ref_append_vector(b, b"gr", _dp(gr), len(gr))
        r.ref_append_vector(b, b"sk", _dp(sk), len(sk))


@pytest.mark.skipif(not o.ref_available(), reason="needs oracle/_ref (built from the reference checkout)")
def test_files_match_reference_writers(tmp_path):
    rng = np.random.default_rng(4)
    P = 7
    uR = np.concatenate([VALUES[:5], rng.normal(0, 1, 2)])
    uI = rng.normal(0, 1e-3, P)
    steps = []
    for s in range(3):
        steps.append((s * 1e-3, float(rng.normal()), float(VALUES[s + 5]), rng.normal(0, 10, P) * 10.0 ** s,
                      np.concatenate([VALUES[s:s + 5], rng.normal(0, 1, 4)]), rng.normal(0, 1, P), rng.normal(0, 1, P)))
    grids = (np.linspace(0.05, 2.45, 25), np.sqrt(np.arange(1, 51)) * 2 * np.pi / 5.0)
    obs = [(0.0, rng.uniform(0, 2, 25), rng.uniform(0, 3, 50)), (0.5, VALUES[:25 % len(VALUES)], VALUES)]
    a, b = str(tmp_path / "lib") + "/", str(tmp_path / "ref") + "/"
    os.makedirs(a)
    os.makedirs(b)
    _lib_sequence(a, P, uR, uI, steps, grids, obs)
    _ref_sequence(b, P, uR, uI, steps, grids, obs)
    names = sorted(os.listdir(b))
    assert sorted(os.listdir(a)) == names and len(names) == 15
    for n in names:
        with open(os.path.join(a, n), "rb") as fa, open(os.path.join(b, n), "rb") as fb:
            assert fa.read() == fb.read(), n


def test_files_refuse_bad_arguments(tmp_path):
    L = tdvmc.lib()
    x = np.zeros(3)
    assert L.tdvmc_files_init(None, _dp(x), _dp(x), 3) == tdvmc.ERR_BAD_ARG
    assert L.tdvmc_files_init((str(tmp_path / "missing") + "/").encode(), _dp(x), _dp(x), 3) == tdvmc.ERR_BAD_ARG
    assert L.tdvmc_files_append_observables(str(tmp_path).encode(), 0.0, None, 2, None, 0) == tdvmc.ERR_BAD_ARG
