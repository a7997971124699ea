"""Writes the golden vectors under tests/golden/ from the reference's own code.

Run in a container that holds the reference checkout, after
``make -C oracle ref`` has compiled oracle/_ref/libref.so from it:

    python tests/golden/make_golden.py

Outputs (inputs and expected outputs only; no reference source):
  nic_cases.json      -- the 48 known answers of src/test/Tests.h:76-156
                         (inputs and expected results transcribed as data),
                         plus 200 random displacements evaluated by the
                         reference's VectorDisplacementNIC
  eval_<name>.npz     -- fed configurations R, parameters, and the reference
                         BosonsBulk outputs: exponent, O, E_R, E_I, other[9],
                         sD[S][N][D], sD2[S][N]
  move_cases.npz      -- single-particle moves through CalculateWFQuotient
  config_*.json       -- reference config files (data) for the schema loader
  obs_<name>.npz      -- fed configurations and the reference BosonsBulk's
                         additional observables: pair distribution g(r) with
                         its grid, S(k) on the 50 wave-vector shells
  kshells.json        -- the first 50 shells of config/kVectors<D>D.json and
                         config/kNorm<D>D.csv (data) for D = 1, 2, 3
  config_items.json   -- names and types of the reference's config items in
                         registration order (src/TDVMC.cpp:297-349), as data

  tdvp_cases.npz      -- moment sets and the reference driver's right-hand
                         sides, solutions of both linear solvers (with and
                         without scaling) and phi-dot, real and imaginary time
  driver_cases.json   -- AdjustParameters and LJ::GetPotential values
  resume_cases.json   -- the driver's successive-run coordinate files (text)
                         and the coordinates its loader reads from config/
  sample_bulk3d_sw16.json -- <E_R> +- error of 256 reference Metropolis chains
  eval_nurbs_<name>.npz -- as eval_<name>.npz on non-uniform knots (USE_NURBS:
                         BosonsBulk::SetNodes with the knots in `grid`)
  spline_weights.json -- SplineFactory::GetWeights tables and BC factors on
                         uniform and NURBS knot vectors

    python tests/golden/make_golden.py obs    # only the observables fixtures
    python tests/golden/make_golden.py drv    # only the driver fixtures (needs `make -C oracle drv`)
    python tests/golden/make_golden.py sample # only the reference sampling fixture
    python tests/golden/make_golden.py nurbs  # only the non-uniform knot fixtures and spline tables
"""
import ctypes as C
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as o  # noqa: E402

REF = "/root/reference"


def f(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# src/test/Tests.h:76-156 -- (L, D, ri, rj, expected |d|, expected d)
NIC_CASES = [
    (4, 3, (0, 0, 0), (0, 0, 0), 0, (0, 0, 0)),
    (4, 3, (0, 0, 0), (1, 0, 0), 1, (-1, 0, 0)),
    (4, 3, (1, 0, 0), (0, 0, 0), 1, (1, 0, 0)),
    (4, 3, (1, 0, 0), (6, 0, 0), 1, (-1, 0, 0)),
    (4, 3, (1, 0, 0), (3, 0, 0), 2, (2, 0, 0)),
    (4, 3, (1, 0, 0), (-1, 0, 0), 2, (-2, 0, 0)),
    (4, 3, (0, 0, 0), (10, 10, 10), 12 ** 0.5, (2, 2, 2)),
    (4, 3, (4, 4, 4), (10, 10, 10), 12 ** 0.5, (2, 2, 2)),
    (4, 3, (4, 4, 4), (10.1, 0, 0), 1.9, (1.9, 0, 0)),
    (4, 3, (4, 4, 4), (9.9, 0, 0), 1.9, (-1.9, 0, 0)),
    (4, 3, (0.1, 0, 0), (3.9, 0, 0), 0.2, (0.2, 0, 0)),
    (4, 3, (0, 0, 0), (3.9, 3.9, 3.9), (0.03) ** 0.5, (0.1, 0.1, 0.1)),
    (4, 3, (4.5, 2.1, -2.1), (0, 0, 0), (0.25 + 1.9 * 1.9 * 2) ** 0.5, (0.5, -1.9, 1.9)),
    (4, 3, (0.5, 1.5, 2.5), (0, 0, 0), (0.25 + 2.25 * 2) ** 0.5, (0.5, 1.5, -1.5)),
    (4, 3, (3.5, 4.5, 5.5), (0, 0, 0), (0.25 * 2 + 2.25) ** 0.5, (-0.5, 0.5, 1.5)),
    (5, 3, (0, 0, 0), (0, 0, 0), 0, (0, 0, 0)),
    (5, 3, (0, 0, 0), (1, 0, 0), 1, (-1, 0, 0)),
    (5, 3, (1, 0, 0), (0, 0, 0), 1, (1, 0, 0)),
    (5, 3, (1, 0, 0), (6, 0, 0), 0, (0, 0, 0)),
    (5, 3, (1, 0, 0), (3, 0, 0), 2, (-2, 0, 0)),
    (5, 3, (1, 0, 0), (-1, 0, 0), 2, (2, 0, 0)),
    (5, 3, (0, 0, 0), (10, 10, 10), 0, (0, 0, 0)),
    (5, 3, (4, 4, 4), (10, 10, 10), 3 ** 0.5, (-1, -1, -1)),
    (5, 3, (4, 4, 4), (10.1, 0, 0), (1.1 * 1.1 + 2) ** 0.5, (-1.1, -1, -1)),
    (5, 3, (4, 4, 4), (9.9, 0, 0), (2 + 0.81) ** 0.5, (-0.9, -1, -1)),
    (5, 3, (0.1, 0, 0), (3.9, 0, 0), 1.2, (1.2, 0, 0)),
    (5, 3, (0, 0, 0), (4.9, 4.9, 4.9), (0.03) ** 0.5, (0.1, 0.1, 0.1)),
    (5, 3, (1.25, 0, 0), (-1.25, 0, 0), 2.5, (-2.5, 0, 0)),
    (5, 3, (5.9, 2.6, -2.6), (0, 0, 0), (0.81 + 2.4 * 2.4 * 2) ** 0.5, (0.9, -2.4, 2.4)),
    (5, 3, (0, 1, 2), (0, 0, 0), 5 ** 0.5, (0, 1, 2)),
    (5, 3, (3, 4, 5), (0, 0, 0), 5 ** 0.5, (-2, -1, 0)),
    (4, 1, (0,), (0,), 0, (0,)),
    (4, 1, (0,), (1,), 1, (-1,)),
    (4, 1, (1,), (0,), 1, (1,)),
    (4, 1, (1,), (1,), 0, (0,)),
    (4, 1, (0,), (2,), 2, (2,)),
    (4, 1, (0,), (3,), 1, (1,)),
    (4, 1, (0,), (4,), 0, (0,)),
    (4, 1, (0,), (5,), 1, (-1,)),
    (5, 1, (0,), (0,), 0, (0,)),
    (5, 1, (0,), (1,), 1, (-1,)),
    (5, 1, (1,), (0,), 1, (1,)),
    (5, 1, (1,), (1,), 0, (0,)),
    (5, 1, (0,), (2,), 2, (-2,)),
    (5, 1, (0,), (3,), 2, (2,)),
    (5, 1, (0,), (4,), 1, (1,)),
    (5, 1, (0,), (5,), 0, (0,)),
    (5, 1, (0,), (6,), 1, (-1,)),
]


def nic():
    r = o.ref()
    cases = []
    for (L, D, ri, rj, dd, dv) in NIC_CASES:
        cases.append(dict(L=L, D=D, ri=list(ri), rj=list(rj), dist=dd, disp=list(dv)))
    rng = np.random.default_rng(7)
    rand = []
    for _ in range(200):
        D = int(rng.integers(1, 4))
        L = float(rng.uniform(2, 12))
        ri = rng.uniform(-1.5 * L, 1.5 * L, D)
        rj = rng.uniform(-1.5 * L, 1.5 * L, D)
        out = np.zeros(D)
        d = r.ref_displacement_nic(f(ri), f(rj), D, L, f(out))
        rand.append(dict(L=L, D=D, ri=ri.tolist(), rj=rj.tolist(), dist=d, disp=out.tolist()))
    with open(os.path.join(HERE, "nic_cases.json"), "w") as fh:
        json.dump(dict(source="src/test/Tests.h:76-156 (known answers, tolerance 1e-9)", known=cases,
                       random=rand), fh, indent=1)


EVAL_CASES = [
    # name, D, N, L, P, sysparams (square well a, b [, t_q, a', b']), W, seed, time
    ("bulk1d_n20_p16", 1, 20, 20.0, 16, [0.5, 3.0], 6, 11, 0.0),
    ("bulk2d_n25_p12", 2, 25, 6.0, 12, [0.6, 2.0], 6, 12, 0.0),
    ("bulk3d_n32_p10", 3, 32, 4.5, 10, [0.7, 2.0], 6, 13, 0.0),
    ("bulk3d_quench", 3, 16, 3.5, 8, [0.7, 2.0, 0.5, 0.9, 4.0], 4, 14, 1.0),
    ("bulk3d_n64_p32", 3, 64, 5.5971, 32, [0.9, 1.5], 4, 15, 0.0),
]


def eval_cases():
    r = o.ref()
    for (name, D, N, L, P, sp, W, seed, time) in EVAL_CASES:
        rng = np.random.default_rng(seed)
        # lattice-jittered configurations spread over several images
        R = rng.uniform(-L / 2, L / 2, (W, N, D))
        R[W // 2:] += rng.integers(-2, 3, (W - W // 2, N, D)) * L
        uR = rng.normal(0, 0.3, P)
        uI = rng.normal(0, 0.2, P)
        phiR, phiI = 0.05, -0.02
        S = P + 2
        spa = np.array(sp, dtype=np.float64)
        expo, O, ER, EI = np.zeros(W), np.zeros((W, P)), np.zeros(W), np.zeros(W)
        oth, sD, sD2 = np.zeros((W, 9)), np.zeros((W, S, N, D)), np.zeros((W, S, N))
        R = np.ascontiguousarray(R)
        r.ref_bosonsbulk_eval(N, D, L, P, f(spa), len(sp), time, W, f(R), f(uR), f(uI), phiR, phiI,
                              f(expo), f(O), f(ER), f(EI), f(oth), f(sD), f(sD2))
        np.savez_compressed(os.path.join(HERE, f"eval_{name}.npz"), D=D, N=N, L=L, P=P, sysparams=spa, time=time,
                            R=R, uR=uR, uI=uI, phiR=phiR, phiI=phiI, expo=expo, O=O, ER=ER, EI=EI, other=oth,
                            sD=sD, sD2=sD2)


# Large cases (N = 343 in 3-D, N = 400 in 2-D): the reference's sD / sD2 are
# contracted here with the effective spline coefficients c = T^T u (the BC
# map of src/PhysicalSystems/BosonsBulk.cpp:158-177) into the per-particle
# gradient and Laplacian of ln psi, which keeps the fixtures small.
EVAL_LARGE = [
    ("bulk3d_n343_p40", 3, 343, 9.7949, 40, [0.9, 1.5], 2, 17, 0.0),
    ("bulk2d_n400_p48", 2, 400, 20.0, 48, [0.25, 3.0], 2, 16, 0.0),
]


def _spline_coeff(u, P):
    c = np.zeros(P + 2)
    c[1] += u[0]
    c[0] += u[1]
    c[2] += u[1]
    for k in range(2, P - 2):
        c[k + 1] += u[k]
    c[P - 1] += u[P - 2]
    c[P + 1] += u[P - 2]
    c[P] += u[P - 1]
    return c


def eval_large():
    r = o.ref()
    for (name, D, N, L, P, sp, W, seed, time) in EVAL_LARGE:
        rng = np.random.default_rng(seed)
        R = np.ascontiguousarray(rng.uniform(-L / 2, L / 2, (W, N, D)))
        uR = rng.normal(0, 0.3, P)
        uI = rng.normal(0, 0.2, P)
        phiR, phiI = 0.05, -0.02
        S = P + 2
        spa = np.array(sp, dtype=np.float64)
        expo, O, ER, EI = np.zeros(W), np.zeros((W, P)), np.zeros(W), np.zeros(W)
        oth, sD, sD2 = np.zeros((W, 9)), np.zeros((W, S, N, D)), np.zeros((W, S, N))
        r.ref_bosonsbulk_eval(N, D, L, P, f(spa), len(sp), time, W, f(R), f(uR), f(uI), phiR, phiI,
                              f(expo), f(O), f(ER), f(EI), f(oth), f(sD), f(sD2))
        cR, cI = _spline_coeff(uR, P), _spline_coeff(uI, P)
        grad = np.stack([np.einsum("s,wsna->wna", cR, sD), np.einsum("s,wsna->wna", cI, sD)], -1)
        lap = np.stack([np.einsum("s,wsn->wn", cR, sD2), np.einsum("s,wsn->wn", cI, sD2)], -1)
        np.savez_compressed(os.path.join(HERE, f"eval_{name}.npz"), D=D, N=N, L=L, P=P, sysparams=spa, time=time,
                            R=R, uR=uR, uI=uI, phiR=phiR, phiI=phiI, expo=expo, O=O, ER=ER, EI=EI, other=oth,
                            grad=grad, lap=lap)


NURBS_CASES = [
    # name, D, N, L, P, sysparams, W, seed, time, knot exponent (g_k = rc (k / (P-1))^a, then jittered)
    ("bulk3d_n32_p12", 3, 32, 4.5, 12, [0.7, 2.0], 6, 41, 0.0, 1.4),
    ("bulk2d_n25_p10", 2, 25, 6.0, 10, [0.6, 2.0], 6, 42, 0.0, 0.8),
    ("bulk1d_n20_p9", 1, 20, 20.0, 9, [0.5, 3.0], 6, 43, 0.0, 1.2),
    ("bulk3d_quench_p8", 3, 16, 3.5, 8, [0.7, 2.0, 0.5, 0.9, 4.0], 4, 44, 1.0, 1.6),
]


def nurbs_grid(P, rc, a, rng):
    """Strictly increasing knots from 0 to rc: a power-law spacing with the
    interior knots jittered by up to 15 % of their gaps."""
    g = rc * (np.arange(P) / (P - 1)) ** a
    gaps = np.diff(g)
    jit = rng.uniform(-0.15, 0.15, P - 2) * np.minimum(gaps[:-1], gaps[1:])
    g[1:-1] += jit
    return g


def nurbs_cases():
    """BosonsBulk on non-uniform knots (SetNodes, src/PhysicalSystems/BosonsBulk.cpp:35-44),
    evaluated by the reference on fed configurations (ref_bosonsbulk_eval_nurbs)."""
    r = o.ref()
    for (name, D, N, L, P, sp, W, seed, time, a) in NURBS_CASES:
        rng = np.random.default_rng(seed)
        grid = nurbs_grid(P, L / 2.0, a, rng)
        R = rng.uniform(-L / 2, L / 2, (W, N, D))
        R[W // 2:] += rng.integers(-2, 3, (W - W // 2, N, D)) * L
        # one pair per walker on an interior knot: the lower_bound tie rule (a
        # pair at exactly L/2 = rc along an axis is left out: the reference's
        # minimum image moves it 1e-10 inwards, src/Utils.cpp:266-281, where the
        # kernels keep it at L/2)
        for w in range(W):
            d = np.zeros(D)
            d[0] = grid[1 + (w % (P - 2))]
            R[w, 1] = R[w, 0] + d
        uR = rng.normal(0, 0.3, P)
        uI = rng.normal(0, 0.2, P)
        phiR, phiI = 0.05, -0.02
        S = P + 2
        spa = np.array(sp, dtype=np.float64)
        expo, O, ER, EI = np.zeros(W), np.zeros((W, P)), np.zeros(W), np.zeros(W)
        oth, sD, sD2 = np.zeros((W, 9)), np.zeros((W, S, N, D)), np.zeros((W, S, N))
        R = np.ascontiguousarray(R)
        r.ref_bosonsbulk_eval_nurbs(N, D, L, P, f(grid), f(spa), len(sp), time, W, f(R), f(uR), f(uI), phiR, phiI,
                                    f(expo), f(O), f(ER), f(EI), f(oth), f(sD), f(sD2))
        np.savez_compressed(os.path.join(HERE, f"eval_nurbs_{name}.npz"), D=D, N=N, L=L, P=P, sysparams=spa,
                            time=time, grid=grid, R=R, uR=uR, uI=uI, phiR=phiR, phiI=phiI, expo=expo, O=O, ER=ER,
                            EI=EI, other=oth, sD=sD, sD2=sD2)


def spline_weights():
    """SplineFactory::GetWeights and the zero-slope BC factors (uniform = true)
    on the knot vectors BosonsBulk builds: uniform (BosonsBulk.cpp:60-66) and
    from a NURBS_GRID (SetNodes, :35-44)."""
    r = o.ref()
    r.ref_spline_weights.restype = C.c_int
    dp = C.POINTER(C.c_double)
    r.ref_spline_weights.argtypes = [dp, C.c_int, dp, dp, dp]
    cases = []
    rng = np.random.default_rng(51)
    for (P, rc, grid) in ((8, 2.5, None), (12, 4.0, None), (10, 3.0, nurbs_grid(10, 3.0, 1.3, rng)),
                          (16, 7.5, nurbs_grid(16, 7.5, 0.8, rng))):
        if grid is None:
            nodes = np.array([i * rc / (P - 1.0) for i in range(-3, P + 3)])
        else:
            nodes = list(grid)
            for i in range(3):
                nodes.insert(0, -grid[i + 1])
                nodes.append(2.0 * grid[P - 1] - grid[P - 2 - i])
            nodes = np.array(nodes)
        S = nodes.size - 4
        w, bcs, bce = np.zeros(S * 16), np.zeros(6), np.zeros(6)
        assert r.ref_spline_weights(f(nodes), nodes.size, f(w), f(bcs), f(bce)) == S
        cases.append(dict(P=P, rc=rc, grid=None if grid is None else grid.tolist(), nodes=nodes.tolist(),
                          weights=w.reshape(S, 4, 4).tolist(), bcFactorsStart=bcs.reshape(2, 3).tolist(),
                          bcFactorsEnd=bce.reshape(2, 3).tolist()))
    with open(os.path.join(HERE, "spline_weights.json"), "w") as fh:
        json.dump(dict(source="SplineFactory::GetWeights (src/SplineFactory.cpp:54-101), "
                              "SetBoundaryConditions3_1D_OR_2 / _CO_2 (:423-456, :563-596)", cases=cases), fh)


def bc_factors():
    """The reference's cubic boundary-condition factors, every family at both
    ends (SetBoundaryConditions3_1D_OR_1..4 / _CO_1..4, src/SplineFactory.cpp:
    377-656), with uniform = true as BosonsBulk passes it (BosonsBulk.cpp:76-77),
    on a uniform knot vector."""
    r = o.ref()
    r.ref_bc_factors.restype = C.c_int
    dp = C.POINTER(C.c_double)
    r.ref_bc_factors.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, dp]
    P, rc = 10, 3.0
    nodes = np.array([i * rc / (P - 1.0) for i in range(-3, P + 3)])
    out = {}
    for side, name in ((0, "origin"), (1, "cutoff")):
        for fam in (1, 2, 3, 4):
            fv = np.zeros(9)
            n = r.ref_bc_factors(fam, side, f(nodes), nodes.size, 1, f(fv))
            assert n > 0
            out[f"{name}_{fam}"] = fv.reshape(3, 3)[:n].tolist()
    with open(os.path.join(HERE, "bc_factors.json"), "w") as fh:
        json.dump(dict(source="SetBoundaryConditions3_1D_OR_k / _CO_k (src/SplineFactory.cpp:377-656), "
                              "uniform = true", factors=out), fh, indent=1)


def vmcsampler_case():
    """The 64-particle 3-D configuration and the 100 spline parameters held in
    src/VMCSampler.h:18-29 (a prototype sampler the driver does not use; its
    own basis differs), evaluated by the reference's BosonsBulk plug-in in a
    box of L = 4 (the fixture's coordinates lie in [-2, 2], maxDistance = 2):
    the fixture and three jittered copies, square-well potential, real
    parameters; gradient / Laplacian contracted as for the large cases."""
    import re
    r = o.ref()
    text = open(os.path.join(REF, "src", "VMCSampler.h")).read()
    body = text[text.index("vector<vector<double> > R = {"):text.index("vector<double> params")]
    R0 = np.array([float(x) for x in re.findall(r"-?\d+\.\d+(?:[eE][-+]?\d+)?", body)]).reshape(64, 3)
    pb = text[text.index("vector<double> params = {"):text.index("int nAcceptances")]
    uR0 = np.array([float(x) for x in re.findall(r"-?\d+\.\d+(?:[eE][-+]?\d+)?", pb)])
    assert uR0.size == 100
    N, D, L, P, W = 64, 3, 4.0, 100, 4
    rng = np.random.default_rng(19)
    R = np.ascontiguousarray(np.stack([R0] + [R0 + rng.uniform(-0.02, 0.02, R0.shape) for _ in range(W - 1)]))
    uR, uI = uR0.copy(), np.zeros(P)
    phiR, phiI = 0.0, 0.0
    sp = np.array([0.3, 2.0])
    S = P + 2
    expo, O, ER, EI = np.zeros(W), np.zeros((W, P)), np.zeros(W), np.zeros(W)
    oth, sD, sD2 = np.zeros((W, 9)), np.zeros((W, S, N, D)), np.zeros((W, S, N))
    r.ref_bosonsbulk_eval(N, D, L, P, f(sp), 2, 0.0, W, f(R), f(uR), f(uI), phiR, phiI, f(expo), f(O), f(ER), f(EI),
                          f(oth), f(sD), f(sD2))
    cR = _spline_coeff(uR, P)
    grad = np.stack([np.einsum("s,wsna->wna", cR, sD), np.zeros((W, N, D))], -1)
    lap = np.stack([np.einsum("s,wsn->wn", cR, sD2), np.zeros((W, N))], -1)
    np.savez_compressed(os.path.join(HERE, "eval_vmcsampler_n64_p100.npz"), D=D, N=N, L=L, P=P, sysparams=sp,
                        time=0.0, R=R, uR=uR, uI=uI, phiR=phiR, phiI=phiI, expo=expo, O=O, ER=ER, EI=EI, other=oth,
                        grad=grad, lap=lap)


def move_cases():
    r = o.ref()
    rng = np.random.default_rng(21)
    rows = []
    for (D, N, L, P) in [(1, 12, 10.0, 8), (2, 20, 6.0, 10), (3, 27, 4.0, 12)]:
        for _ in range(10):
            R = rng.uniform(-L / 2, L / 2, (N, D))
            uR = rng.normal(0, 0.3, P)
            p = int(rng.integers(0, N))
            newpos = R[p] + rng.normal(0, 0.4, D)
            q = C.c_double(0)
            spa = np.array([0.5, 1.0])
            en = r.ref_bosonsbulk_move(N, D, L, P, f(spa), 2, f(np.ascontiguousarray(R)), f(uR), p,
                                       f(np.ascontiguousarray(newpos)), C.byref(q))
            rows.append(dict(D=D, N=N, L=L, P=P, R=R.ravel().tolist(), uR=uR.tolist(), p=p,
                             newpos=newpos.tolist(), exponent_new=en, quotient=q.value))
    with open(os.path.join(HERE, "move_cases.json"), "w") as fh:
        json.dump(rows, fh)



OBS_CASES = [
    # name, D, N, L, P (pair-spline knots), GR_BIN_COUNT, W, seed
    ("bulk1d_n20", 1, 20, 20.0, 16, 50, 4, 21),
    ("bulk2d_n25", 2, 25, 6.0, 12, 37, 4, 22),
    ("bulk3d_n64", 3, 64, 5.5971, 32, 100, 4, 23),
    ("bulk3d_n343", 3, 343, 9.7949, 40, 200, 2, 24),
]


def obs_cases():
    r = o.ref()
    for (name, D, N, L, P, bins, W, seed) in OBS_CASES:
        rng = np.random.default_rng(seed)
        R = rng.uniform(-L / 2, L / 2, (W, N, D))
        R[W // 2:] += rng.integers(-2, 3, (W - W // 2, N, D)) * L
        R = np.ascontiguousarray(R)
        spa = np.array([0.9, 1.5])
        cap = 4096
        ng = C.c_int()
        gr, sk = np.zeros(W * cap), np.zeros(W * 50)
        gg, sg = np.zeros(cap), np.zeros(50)
        rc = r.ref_bosonsbulk_observables(N, D, L, P, f(spa), 2, W, f(R), bins, cap, C.byref(ng), f(gr), f(sk),
                                          f(gg), f(sg))
        assert rc == 0 and ng.value == bins, (rc, ng.value)
        n = ng.value
        np.savez_compressed(os.path.join(HERE, f"obs_{name}.npz"), D=D, N=N, L=L, P=P, bins=bins, R=R,
                            gr=gr[:W * n].reshape(W, n), sk=sk.reshape(W, 50), gr_grid=gg[:n], sk_grid=sg)


def kshells():
    out = {}
    for D in (1, 2, 3):
        with open(os.path.join(REF, "config", f"kVectors{D}D.json")) as fh:
            data = json.load(fh)["data"][:50]
        with open(os.path.join(REF, "config", f"kNorm{D}D.csv")) as fh:
            norms = [float(x) for x in fh.read().split()][:50]
        out[str(D)] = dict(shells=[[[int(c) for c in v] for v in sh] for sh in data], norms=norms)
    with open(os.path.join(HERE, "kshells.json"), "w") as fh:
        json.dump(dict(source="config/kVectors<D>D.json, config/kNorm<D>D.csv (first 50 shells)", shells=out), fh)


def config_items():
    import re
    with open(os.path.join(REF, "src", "TDVMC.cpp")) as fh:
        text = fh.read()
    items = re.findall(r'configItems\.push_back\(ConfigItem\("([A-Z0-9_]+)",[^,]+, ConfigItemType::([A-Z_]+)', text)
    with open(os.path.join(HERE, "config_items.json"), "w") as fh:
        json.dump(dict(source="src/TDVMC.cpp:297-349 (registration order)", items=items), fh)


def configs():
    for name in ["BosonsBulk1D.config", "InhContactBosonsLinearResponseToPulse.config"]:
        src = os.path.join(REF, "config", name)
        if os.path.exists(src):
            shutil.copyfile(src, os.path.join(HERE, "config_" + name.replace(".config", ".json")))


def drv():
    """The reference driver's own TDVP pieces (oracle/_ref/libref_driver.so,
    built by `make -C oracle drv` from the whole reference source tree)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libref_driver.so")
    if not os.path.exists(path):
        sys.exit("oracle/_ref/libref_driver.so missing: run `make -C oracle drv`")
    return C.CDLL(path)


def _moments_raw(O, ER, EI):
    """The driver's accumulated means (src/TDVMC.cpp:1103-1109): <O>, (<E_R>,
    <E_I>), <O O^T>, <O E_R>, <O E_I>."""
    W = O.shape[0]
    return (O.mean(0), np.array([ER.mean(), EI.mean()]), (O.T @ O) / W, (O * ER[:, None]).mean(0),
            (O * EI[:, None]).mean(0))


def _tdvp_inputs(name):
    if name == "p12":
        # correlated synthetic samples: generic, well-conditioned S
        rng = np.random.default_rng(31)
        W, P = 600, 12
        A = rng.normal(0, 1, (P, P)) / np.sqrt(P) + np.eye(P)
        O = rng.normal(0, 1, (W, P)) @ A + rng.normal(0, 0.3, P)
        ER = 2.0 + O @ rng.normal(0, 0.3, P) + rng.normal(0, 0.5, W)
        EI = O @ rng.normal(0, 0.1, P) + rng.normal(0, 0.2, W)
        return O, ER, EI
    # p296_twins: O, E_L of a 3-D bulk system with a 40-parameter pair spline
    # and 256 k-space terms, evaluated by the oracle on random configurations;
    # O_k = O_-k for every +-k pair, so S is exactly singular
    N, D, L, Ps, K, W = 27, 3, 3.0, 40, 256, 400
    s = o.System(N=N, D=D, L=L, P_s=Ps, kvec=o.kvectors(D, K, L), pot_type=o.POT_GAUSS, pot=(0.5, 2.0),
                 hbar2_2m=0.5)
    rng = np.random.default_rng(32)
    R = rng.uniform(-L / 2, L / 2, (W, N, D))
    uR = rng.normal(0, 0.2, s.P)
    uR[Ps:] *= 0.01
    uI = rng.normal(0, 0.05, s.P)
    uI[Ps:] *= 0.01
    out = o.eval_configs(s, R, uR, uI)
    return out["O"], out["ER"], out["EI"]


def tdvp_cases():
    """Right-hand sides (BuildSystemOfEquationsForParameters), both linear
    solvers with and without the diagonal scaling (SolveForParametersDot,
    CalculatePhiDot), in real and imaginary time, on two moment sets; plus
    CalculatePhiDot alone and AdjustParameters."""
    L = drv()
    out = {}
    for name in ("p12", "p296_twins"):
        O, ER, EI = _tdvp_inputs(name)
        mO, mE, OO, OER, OEI = (np.ascontiguousarray(a) for a in _moments_raw(O, ER, EI))
        P = mO.size
        rec = dict(mO=mO, mE=mE, OO=OO, OER=OER, OEI=OEI)
        for imag in (0, 1):
            S, bR, bI = np.zeros((P, P)), np.zeros(P), np.zeros(P)
            L.refd_build_rhs(P, imag, f(mO), f(mE), f(OO), f(OER), f(OEI), f(S), f(bR), f(bI))
            rec[f"bR_i{imag}"], rec[f"bI_i{imag}"] = bR, bI
            if P <= 64:
                rec[f"S_i{imag}"] = S
            for solver in (0, 1):
                for pre in (0, 1):
                    xR, xI, pd = np.zeros(P), np.zeros(P), np.zeros(2)
                    bad = L.refd_solve(P, imag, solver, pre, f(mO), f(mE), f(OO), f(OER), f(OEI), f(xR), f(xI),
                                       f(pd))
                    key = f"i{imag}_s{solver}_p{pre}"
                    rec["xR_" + key], rec["xI_" + key], rec["phidot_" + key] = xR, xI, pd
                    rec["bad_" + key] = np.array(bad)
            # phi-dot alone on a given alpha-dot
            rng = np.random.default_rng(40 + imag)
            xR, xI, pd = rng.normal(0, 1, P), rng.normal(0, 1, P), np.zeros(2)
            L.refd_phi_dot(P, imag, f(mO), f(mE), f(xR), f(xI), f(pd))
            rec[f"phi_xR_i{imag}"], rec[f"phi_xI_i{imag}"], rec[f"phi_out_i{imag}"] = xR, xI, pd
        out[name] = rec
    np.savez_compressed(os.path.join(HERE, "tdvp_cases.npz"),
                        **{f"{n}__{k}": v for n, rec in out.items() for k, v in rec.items()})
    # AdjustParameters over the first N_PARAM entries of a longer vector
    rng = np.random.default_rng(43)
    n_param, P = 10, 16
    uR, uI = rng.normal(0, 1, P), rng.normal(0, 1, P)
    aR, aI = uR.copy(), uI.copy()
    L.refd_adjust(n_param, f(aR), f(aI), P)
    # Lennard-Jones (LJ::GetPotential) with unit and He-He constants
    r_unit = np.linspace(0.85, 3.0, 40)
    r_he = np.linspace(2.2, 12.5, 40)
    v_unit, v_he = np.zeros(40), np.zeros(40)
    L.refd_lj(0, f(r_unit), 40, f(v_unit))
    L.refd_lj(1, f(r_he), 40, f(v_he))
    with open(os.path.join(HERE, "driver_cases.json"), "w") as fh:
        json.dump(dict(source="src/TDVMC.cpp:2531-2567 (AdjustParameters), src/Potentials/LJ.cpp:16-22 with "
                              "LJ_He_He.cpp:13-14 (GetPotential)",
                       adjust=dict(n_param=n_param, uR=uR.tolist(), uI=uI.tolist(), outR=aR.tolist(),
                                   outI=aI.tolist()),
                       lj=dict(r_unit=r_unit.tolist(), v_unit=v_unit.tolist(), r_he=r_he.tolist(),
                               v_he=v_he.tolist(), sigma_he=2.556, eps_he=10.22)), fh)


def resume_cases():
    """The successive-run coordinate file as the reference driver writes it
    (MoveCoordinatesToFirstCell + WriteParticleInputFile), as text, for walkers
    partly outside the box in 1, 2 and 3 dimensions; and the coordinates its
    LoadLastPositionsFromFile reads from config/particleconfiguration_400.csv
    (2-D, N = 400) and _64.csv (3-D, N = 64)."""
    import tempfile
    L = drv()
    rng = np.random.default_rng(61)
    files = []
    for (n, d, box) in [(5, 1, 6.0), (7, 2, 5.0), (9, 3, 4.0)]:
        R = np.ascontiguousarray(rng.uniform(-1.2 * box, 1.2 * box, (n, d)))
        R[0, 0] = box / 2.0          # exactly on the half-box: the nudge of GetCoordinateNIC
        R[1, 0] = -box / 2.0
        with tempfile.TemporaryDirectory() as tmp:
            L.refd_write_particle_file((tmp + "/").encode(), b"coords_case", f(R), n, d, C.c_double(box), 1)
            text = open(os.path.join(tmp, "coords_case.dat")).read()
        files.append(dict(N=n, D=d, L=box, R=R.ravel().tolist(), text=text))
    loads = []
    for (name, n, d) in [("particleconfiguration_400", 400, 2), ("particleconfiguration_64", 64, 3)]:
        out = np.zeros(n * d)
        rc = L.refd_load_positions((REF + "/config/").encode(), name.encode(), f(out), n, d)
        assert rc == 0
        loads.append(dict(file="refconf_" + name + ".csv", N=n, D=d, R=out.tolist()))
    with open(os.path.join(HERE, "resume_cases.json"), "w") as fh:
        json.dump(dict(source="src/TDVMC.cpp:493-504, :640-670, :787-796 (WriteParticleInputFile, "
                              "LoadLastPositionsFromFile, MoveCoordinatesToFirstCell)", write=files, load=loads), fh)


# Reference Metropolis chains of a 3-D square-well system (BosonsBulk with its
# own incremental move ratio, src/PhysicalSystems/BosonsBulk.cpp:553-665, and
# the move loop of DoMetropolisStep, src/TDVMC.cpp:858-916): independent
# chains from jittered lattices, each thermalised, then n_samples sweeps of N
# moves followed by one measurement of E_R.
SAMPLE_CASE = dict(D=3, N=16, L=3.2, P=8, sysparams=[0.6, 2.0], step=0.45, chains=256, burn=400, samples=3000,
                   seed=51)


def sample_case():
    r = o.ref()
    c = SAMPLE_CASE
    D, N, L, P = c["D"], c["N"], c["L"], c["P"]
    rng = np.random.default_rng(c["seed"])
    uR = np.array([-0.9, -0.6, -0.35, -0.2, -0.1, -0.05, -0.02, 0.0]) + rng.normal(0, 0.02, P)
    spa = np.array(c["sysparams"])
    n = 2
    base = (np.stack(np.meshgrid(*[np.arange(n)] * D, indexing="ij"), -1).reshape(-1, D) + 0.5) * (L / n)
    means = []
    for ch in range(c["chains"]):
        R = np.ascontiguousarray(base.repeat(N // len(base), 0) + rng.uniform(-0.3, 0.3, (N, D)) - L / 2)
        e = C.c_double()
        r.ref_bosonsbulk_sample(N, D, L, P, f(spa), 2, f(R), f(uR), c["step"], 1000 + ch, c["burn"], N,
                                C.byref(e))
        r.ref_bosonsbulk_sample(N, D, L, P, f(spa), 2, f(R), f(uR), c["step"], 5000 + ch, c["samples"], N,
                                C.byref(e))
        means.append(e.value)
    means = np.array(means)
    res = dict(c, uR=uR.tolist(), chain_means=means.tolist(), mean=float(means.mean()),
               sigma=float(means.std(ddof=1) / np.sqrt(len(means))),
               source="ref_bosonsbulk_sample over the reference BosonsBulk plug-in (oracle/ref_harness.cpp)")
    with open(os.path.join(HERE, "sample_bulk3d_sw16.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print("reference <E_R> =", res["mean"], "+-", res["sigma"])


if __name__ == "__main__":
    if not o.ref_available():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` in a container with the reference")
    if sys.argv[1:] == ["drv"]:
        tdvp_cases()
        resume_cases()
        sys.exit(0)
    if sys.argv[1:] == ["eval_large"]:
        eval_large()
        vmcsampler_case()
        sys.exit(0)
    if sys.argv[1:] == ["sample"]:
        sample_case()
        sys.exit(0)
    if sys.argv[1:] == ["nurbs"]:
        nurbs_cases()
        spline_weights()
        sys.exit(0)
    if sys.argv[1:] == ["bc"]:
        bc_factors()
        sys.exit(0)
    if sys.argv[1:] != ["obs"]:
        nic()
        eval_cases()
        nurbs_cases()
        move_cases()
        configs()
    obs_cases()
    kshells()
    config_items()
    print("golden vectors written to", HERE)
