"""The reference driver's output files (tdvmc_files_*), byte for byte against
the reference's own writers WriteDataToFile / AppendDataToFile
(src/Utils.cpp:818-1016, compiled into oracle/_ref) on the same call sequence
as src/TDVMC.cpp:3117-3129 (start), :3833-3839 (per step), :3151-3157 (grids)
and :3475-3500 (observables).  The values include signed zeros, subnormals,
huge and tiny magnitudes and infinities.  Host code only."""
import ctypes as C
import json
import os
import re

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


KIND = {"STRING": 0, "INT": 1, "DOUBLE": 2, "ARR_INT": 3, "ARR_DOUBLE": 4}
EXTENSION_KEYS = ["HBAR2_2M", "PAIR_POTENTIAL", "K_VECTORS", "ONE_BODY", "TRAP_OMEGA", "SPLINE_CUTOFF", "N_WALKERS",
                  "SAMPLES_PER_WALKER", "SEED", "PRECISION", "REG_EPS", "MOVE_ORDER", "SPLINE_BC_ORIGIN",
                  "SPLINE_BC_CUTOFF"]


def _ref_json(kind, value):
    r = o.ref()
    r.ref_config_json.restype = C.c_int
    r.ref_config_json.argtypes = [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int,
                                  C.c_char_p, C.c_int]
    buf = C.create_string_buffer(1 << 16)
    if kind == 0:
        n = r.ref_config_json(0, value.encode(), 0, None, None, 0, buf, len(buf))
    elif kind == 1:
        n = r.ref_config_json(1, None, int(value), None, None, 0, buf, len(buf))
    elif kind == 2:
        d = np.array([float(value)])
        n = r.ref_config_json(2, None, 0, _dp(d), None, 1, buf, len(buf))
    elif kind == 3:
        a = np.array(value, dtype=np.int32)
        n = r.ref_config_json(3, None, 0, None, a.ctypes.data_as(C.POINTER(C.c_int)), len(a), buf, len(buf))
    else:
        a = np.array(value, dtype=np.float64)
        n = r.ref_config_json(4, None, 0, _dp(a) if len(a) else None, None, len(a), buf, len(buf))
    assert n >= 0
    return buf.value.decode()


@pytest.mark.skipif(not o.ref_available(), reason="needs oracle/_ref (built from the reference checkout)")
def test_final_config_matches_reference_format(tmp_path, golden_dir):
    # the reference's own config file, its final parameters replaced
    with open(os.path.join(golden_dir, "config_items.json")) as fh:
        ref_items = json.load(fh)["items"]
    src = os.path.join(golden_dir, "config_BosonsBulk1D.json")
    with open(src) as fh:
        cfg = json.load(fh)
    P = int(cfg["N_PARAM"])
    rng = np.random.default_rng(3)
    uR, uI = rng.normal(0, 1, P) * 10.0 ** rng.integers(-8, 8, P), rng.normal(0, 1e-3, P)
    phiR, phiI = -12.3456789123, 4.5e-7
    out = str(tmp_path / "AAFinish_tVMC.config")
    L = tdvmc.lib()
    assert L.tdvmc_write_final_config(src.encode(), _dp(uR), _dp(uI), phiR, phiI, P, out.encode()) == 0
    text = open(out).read()
    assert text.startswith("{\n") and text.endswith("\n}")
    pairs = re.findall(r'\t"([A-Z0-9_]+)" : (.*?)(?:,\n|\n)(?=\t"|\}$)', text[2:], re.S)
    names = [n for n, _ in pairs]
    assert names == [n for n, _ in ref_items] + EXTENSION_KEYS
    final = dict(PARAMS_REAL=list(uR), PARAMS_IMAGINARY=list(uI), PARAM_PHIR=phiR, PARAM_PHII=phiI)
    for (name, value_text), (rname, rtype) in zip(pairs, ref_items):
        kind = KIND[rtype]
        v = final.get(name, cfg.get(name))
        if v is None or (isinstance(v, int) and not isinstance(v, bool) and v == 0 and name not in final):
            v = {0: "", 1: 0, 2: 0.0, 3: [], 4: []}[kind]   # missing key or explicit integer 0 (ConfigItem::setValue)
        assert value_text == _ref_json(kind, v), name
    # and it reads back (integers exact, doubles to the reference's six decimals, parameters to 15 digits)
    info = tdvmc.parse_config(out)
    assert info.N == cfg["N"] and info.DIM == cfg["DIM"] and info.P == P
    back = json.loads(text)
    np.testing.assert_allclose(back["PARAMS_REAL"], uR, rtol=1e-14)
    assert back["PARAM_PHIR"] == float("%f" % phiR)


# ---------------------------------------------------------------- resume files

def test_coordinate_files_vs_reference_writer(tmp_path, golden_dir):
    """The successive-run coordinate files (src/TDVMC.cpp:4007-4022) against the
    text the reference driver's own MoveCoordinatesToFirstCell +
    WriteParticleInputFile wrote (tests/golden/resume_cases.json), including
    coordinates on exactly +-L/2 (the 1e-10 nudge) and far outside the box."""
    cases = json.load(open(os.path.join(golden_dir, "resume_cases.json")))["write"]
    L = tdvmc.lib()
    for i, c in enumerate(cases):
        d = tmp_path / f"case{i}"
        d.mkdir()
        R = np.array(c["R"], dtype=np.float64)
        N, D = c["N"], c["D"]
        assert L.tdvmc_files_write_coordinates((str(d) + "/").encode(), _dp(R), 1, N, D, c["L"]) == 0
        got = (d / "coords" / f"particleconfiguration_{N}_{D}D_0.dat").read_text()
        assert got == c["text"], (got, c["text"])
        assert (d / f"AAFinish_particleconfiguration_{N}_{D}D.dat").read_text() == c["text"]
    # several walkers: one file per walker, as one per MPI rank
    d = tmp_path / "multi"
    d.mkdir()
    R = np.arange(3 * 4 * 2, dtype=np.float64).reshape(3, 4, 2) * 0.37
    assert L.tdvmc_files_write_coordinates((str(d) + "/").encode(), _dp(R), 3, 4, 2, 0.0) == 0
    for w in range(3):
        vals = [float(x) for x in (d / "coords" / f"particleconfiguration_4_2D_{w}.dat").read_text().split()]
        assert np.allclose(vals, R[w].ravel(), rtol=1e-15, atol=0)   # open space: as they are (16 digits)


def test_read_coordinates_vs_reference_loader(golden_dir):
    """LoadLastPositionsFromFile (src/TDVMC.cpp:640-670) on the reference's own
    config/particleconfiguration_400.csv and _64.csv: the same doubles."""
    cases = json.load(open(os.path.join(golden_dir, "resume_cases.json")))["load"]
    for c in cases:
        R = np.zeros(c["N"] * c["D"])
        rc = tdvmc.lib().tdvmc_files_read_coordinates(os.path.join(golden_dir, c["file"]).encode(), _dp(R),
                                                       c["N"], c["D"])
        assert rc == 0
        assert np.array_equal(R, np.array(c["R"]))
    # a file with the wrong count is refused
    R = np.zeros(12)
    assert tdvmc.lib().tdvmc_files_read_coordinates(os.path.join(golden_dir, cases[0]["file"]).encode(), _dp(R),
                                                     4, 3) == tdvmc.ERR_BAD_ARG


def test_rng_state_file_roundtrip(tmp_path):
    L = tdvmc.lib()
    seed, epoch = 0xDEADBEEFCAFEF00D, 123456
    assert L.tdvmc_files_write_rng_state((str(tmp_path) + "/").encode(), seed, epoch, 5) == 0
    s, e = C.c_uint64(), C.c_uint32()
    path = tmp_path / "random" / "state_philox_5.dat"
    assert L.tdvmc_files_read_rng_state(str(path).encode(), C.byref(s), C.byref(e)) == 0
    assert (s.value, e.value) == (seed, epoch)
