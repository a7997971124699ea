"""CPU checks of the C-ABI library: it loads without a GPU, exports every
symbol include/tdvmc.h declares, and parses the reference's config schema with
the ConfigItem semantics (no device calls)."""
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
tdvmc = pytest.importorskip("tdvmc_amd")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tdvmc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|tdvmc_ctx\*)\s+(tdvmc_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = tdvmc.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(tdvmc.SYMBOLS), set(syms) ^ set(tdvmc.SYMBOLS)


def test_parse_reference_config_file(golden_dir):
    # config/BosonsBulk1D.config of the reference: no LBOX key -> L = N / RHO
    info = tdvmc.parse_config(os.path.join(golden_dir, "config_BosonsBulk1D.json"))
    assert info.N == 20 and info.DIM == 1 and info.P == 100 and info.P_pair == 100
    assert info.L == pytest.approx(20.0)
    assert info.imaginary_time == 1
    assert info.therm_moves == 100 and info.init_moves == 200
    assert info.hbar2_2m == 1.0          # src/Constants.h:12
    # MC_NSTEPS = 5000 samples per stage on the reference's one walker per rank
    # (src/TDVMC.cpp:1073-1080): 5000 samples per walker, one walker by default
    assert info.samples_per_walker == 5000 and info.n_walkers == 1


def test_samples_per_walker_from_mc_nsteps():
    base = {"N": 8, "DIM": 3, "LBOX": 4.0, "N_PARAM": 6, "SYSTEM_PARAMS": [0.5, 1.0]}
    assert tdvmc.parse_config(json.dumps(base)).samples_per_walker == 1
    assert tdvmc.parse_config(json.dumps(dict(base, MC_NSTEPS=40))).samples_per_walker == 40
    info = tdvmc.parse_config(json.dumps(dict(base, MC_NSTEPS=40, SAMPLES_PER_WALKER=3)))
    assert info.samples_per_walker == 3


def test_parse_config_item_semantics():
    base = {"N": 8, "DIM": 3, "LBOX": 4.0, "N_PARAM": 6, "SYSTEM_PARAMS": [0.5, 1.0], "MC_NTHERMSTEPS": 7}
    info = tdvmc.parse_config(json.dumps(base))
    assert info.therm_moves == 7 and info.reg_eps == 1e-3
    # an explicit integer 0 leaves the value at its initial 0 as well
    info = tdvmc.parse_config(json.dumps(dict(base, MC_NTHERMSTEPS=0)))
    assert info.therm_moves == 0
    info = tdvmc.parse_config(json.dumps(dict(base, K_VECTORS=26, ONE_BODY=0, REG_EPS=1e-4, PRECISION="MIXED")))
    assert info.P == 6 + 26 and info.K == 26 and info.reg_eps == 1e-4 and info.precision == 1
    info = tdvmc.parse_config(json.dumps(dict(base, LBOX=0.0, ONE_BODY=1, TRAP_OMEGA=1.0, SPLINE_CUTOFF=3.0)))
    assert info.L == 0.0 and info.one_body == 1 and info.P == 7 and info.rc == 3.0


@pytest.mark.parametrize("bad", [
    {"N": 8, "DIM": 4, "LBOX": 4.0, "N_PARAM": 6},
    {"N": 8, "DIM": 3, "LBOX": 4.0, "N_PARAM": 3},
    {"N": 8, "DIM": 3, "LBOX": 0.0, "N_PARAM": 6, "K_VECTORS": 4, "SPLINE_CUTOFF": 2.0},
    {"N": 8, "DIM": 3, "LBOX": 4.0, "N_PARAM": 6, "PAIR_POTENTIAL": "MORSE", "SYSTEM_PARAMS": [1, 1]},
    {"N": 8, "DIM": 3, "LBOX": 4.0, "N_PARAM": 6, "USE_NURBS": 1},
])
def test_parse_config_rejects(bad):
    with pytest.raises(tdvmc.TdvmcError) as e:
        tdvmc.parse_config(json.dumps(bad))
    assert e.value.code == tdvmc.ERR_BAD_CONFIG


def test_parse_config_bad_json():
    with pytest.raises(tdvmc.TdvmcError):
        tdvmc.parse_config("{ \"N\": 3, ")


def test_workloads_are_valid_configs():
    from tdvmc_amd import workloads

    for name, make in workloads.WORKLOADS.items():
        info = tdvmc.parse_config(make())
        assert info.N >= 20
    info = tdvmc.parse_config(workloads.bulk3d_lj343_k256())
    assert (info.N, info.DIM, info.P, info.K, info.n_walkers) == (343, 3, 296, 256, 65536)
    info = tdvmc.parse_config(workloads.bulk2d_gauss400_k512())
    assert (info.N, info.DIM, info.P, info.K) == (400, 2, 560, 512)


def test_cpp_header_compiles_and_check_program_built():
    """include/tdvmc.hpp compiles on its own as C++17, and the C++ check
    program is built (it only prints its usage without arguments)."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = "#include \"tdvmc.hpp\"\nint main() { return 0; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", "-I",
                        os.path.join(root, "include"), "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    binp = os.path.join(root, "tests", "cpp", "build", "adapter_check")
    assert os.path.exists(binp), "run the build first"
    r = subprocess.run([binp], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
