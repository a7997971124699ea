"""CPU oracle for the TDVMC hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``liboracle.so`` (the scalar C restatement in
``tdvmc_oracle.c``) and over ``_ref/libref.so`` (the reference's own
``BosonsBulk`` plug-in, compiled from the reference checkout by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module; the product library never
does.

Parity status of each function is recorded in ``tdvmc_oracle.c``'s header.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


class _CSystem(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("D", C.c_int), ("L", C.c_double), ("P_s", C.c_int), ("rc", C.c_double),
        ("K", C.c_int), ("kvec", _dp), ("one_body", C.c_int), ("hbar2_2m", C.c_double),
        ("pot_type", C.c_int), ("n_pot", C.c_int), ("pot", C.c_double * 5),
        ("pot_cutoff", C.c_double), ("trap_omega", C.c_double), ("grid", _dp),
        ("bc_origin", C.c_int), ("bc_cutoff", C.c_int),
    ]


POT_NONE, POT_SQUARE, POT_GAUSS, POT_LJ = 0, 1, 2, 3


@dataclass
class System:
    """Shape and physics of one system (mirrors the library's system record)."""
    N: int
    D: int
    L: float            # periodic box (<= 0: open space)
    P_s: int
    rc: float = 0.0     # spline cutoff; default L/2
    kvec: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    one_body: int = 0
    hbar2_2m: float = 1.0
    pot_type: int = POT_NONE
    pot: tuple = ()
    pot_cutoff: float = 0.0   # default: rc
    trap_omega: float = 0.0
    grid: np.ndarray = None   # non-uniform knots (NURBS_GRID, P_s values from 0); rc = grid[-1]
    bc_origin: int = 2        # boundary-condition family at r = 0 (1..4; BosonsBulk: 2)
    bc_cutoff: int = 2        # boundary-condition family at r = rc (1..4; BosonsBulk: 2)

    def __post_init__(self):
        if self.grid is not None:
            self.grid = np.ascontiguousarray(np.asarray(self.grid, dtype=np.float64))
            assert self.grid.size == self.P_s
            self.rc = float(self.grid[-1])
        if self.rc <= 0:
            self.rc = self.L / 2.0
        if self.pot_cutoff <= 0:
            self.pot_cutoff = self.rc
        self.kvec = np.ascontiguousarray(np.asarray(self.kvec, dtype=np.float64).reshape(-1, self.D))

    @property
    def K(self):
        return self.kvec.shape[0]

    @property
    def P(self):
        return (1 if self.one_body else 0) + self.P_s + self.K

    def c(self):
        s = _CSystem()
        s.N, s.D, s.L, s.P_s, s.rc = self.N, self.D, self.L, self.P_s, self.rc
        s.K = self.K
        s.kvec = _ptr(self.kvec) if self.K else None
        s.one_body, s.hbar2_2m = self.one_body, self.hbar2_2m
        s.pot_type, s.n_pot = self.pot_type, len(self.pot)
        for i, v in enumerate(self.pot):
            s.pot[i] = v
        s.pot_cutoff, s.trap_omega = self.pot_cutoff, self.trap_omega
        s.grid = _ptr(self.grid) if self.grid is not None else None
        s.bc_origin, s.bc_cutoff = self.bc_origin, self.bc_cutoff
        s._keep = (self.kvec, self.grid)
        return s


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.tdo_nic_coord.restype = C.c_double
        L.tdo_nic_coord.argtypes = [C.c_double, C.c_double]
        L.tdo_displacement.restype = C.c_double
        L.tdo_displacement.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp]
        L.tdo_bspline.restype = C.c_int
        L.tdo_bspline.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
        L.tdo_bspline_nu.restype = C.c_int
        L.tdo_bspline_nu.argtypes = [C.c_double, _dp, C.c_int, _dp, _dp, _dp]
        L.tdo_basis_sums.restype = C.c_int
        L.tdo_basis_sums.argtypes = [C.POINTER(_CSystem), _dp, _dp]
        L.tdo_bc_factors.restype = C.c_int
        L.tdo_bc_factors.argtypes = [C.c_int, C.c_int, _dp]
        L.tdo_spline_coeffs.restype = None
        L.tdo_spline_coeffs.argtypes = [C.POINTER(_CSystem), _dp, _dp]
        L.tdo_spline_ops.restype = None
        L.tdo_spline_ops.argtypes = [C.POINTER(_CSystem), _dp, _dp]
        L.tdo_pair_potential.restype = C.c_double
        L.tdo_pair_potential.argtypes = [C.POINTER(_CSystem), C.c_double, C.c_double]
        L.tdo_eval_batch.restype = C.c_int
        L.tdo_eval_batch.argtypes = [C.POINTER(_CSystem), C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                     _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.tdo_metropolis.restype = C.c_int64
        L.tdo_metropolis.argtypes = [C.POINTER(_CSystem), _dp, _dp, C.c_double, C.c_uint64, C.c_int64,
                                     C.c_int64, _dp]
        L.tdo_moments.restype = None
        L.tdo_moments.argtypes = [C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.tdo_cholesky_ref.restype = C.c_int
        L.tdo_cholesky_ref.argtypes = [C.c_int, _dp]
        L.tdo_chol_solve.restype = None
        L.tdo_chol_solve.argtypes = [C.c_int, _dp, _dp, _dp]
        L.tdo_solve_param_dot.restype = C.c_int
        L.tdo_solve_param_dot.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                          _dp, _dp, _dp, _dp]
        L.tdo_solve_param_dot_qr.restype = C.c_int
        L.tdo_solve_param_dot_qr.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _dp]
        L.tdo_build_rhs.restype = None
        L.tdo_build_rhs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.tdo_phi_dot.restype = None
        L.tdo_phi_dot.argtypes = [C.c_int, C.c_int, _dp, C.c_double, _dp, _dp, _dp, _dp]
        L.tdo_fpqr_solve.restype = C.c_int
        L.tdo_fpqr_solve.argtypes = [C.c_int, _dp, _dp, _dp]
        _ip = C.POINTER(C.c_int)
        L.tdo_gr_grid.restype = C.c_int
        L.tdo_gr_grid.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        L.tdo_kshells.restype = C.c_int
        L.tdo_kshells.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip]
        L.tdo_observables.restype = C.c_int
        L.tdo_observables.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                      _dp, _dp, _dp]
        _LIB = L
    return _LIB


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libref.so is not built (needs the reference checkout)")
        L = C.CDLL(path)
        L.ref_displacement_nic.restype = C.c_double
        L.ref_displacement_nic.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp]
        L.ref_bosonsbulk_eval.restype = C.c_int
        L.ref_bosonsbulk_eval.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int, C.c_double,
                                          C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                          _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_bosonsbulk_eval_nurbs.restype = C.c_int
        L.ref_bosonsbulk_eval_nurbs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp, C.c_int,
                                                C.c_double, C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                                _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_bosonsbulk_move.restype = C.c_double
        L.ref_bosonsbulk_move.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int, _dp, _dp,
                                          C.c_int, _dp, _dp]
        L.ref_bosonsbulk_sample.restype = C.c_int64
        L.ref_bosonsbulk_sample.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int, _dp, _dp,
                                            C.c_double, C.c_uint64, C.c_int64, C.c_int64, _dp]
        L.ref_bosonsbulk_observables.restype = C.c_int
        L.ref_bosonsbulk_observables.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int, C.c_int64,
                                                 _dp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp, _dp]
        _REF = L
    return _REF


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------- geometry

def nic_coord(x, L):
    return lib().tdo_nic_coord(float(x), float(L))


def displacement(ri, rj, L):
    ri, rj = _f64(ri), _f64(rj)
    out = np.zeros_like(ri)
    d = lib().tdo_displacement(_ptr(ri), _ptr(rj), ri.size, float(L), _ptr(out))
    return d, out


def bspline(r, h, P_s):
    b, db, d2b = np.zeros(4), np.zeros(4), np.zeros(4)
    j = lib().tdo_bspline(float(r), float(h), int(P_s), _ptr(b), _ptr(db), _ptr(d2b))
    return j, b, db, d2b


def bspline_nu(r, grid, P_s):
    """The four non-zero cubic B-splines (values, d/dr, d2/dr2) at r on the
    mirrored non-uniform knots of grid (tdo_bspline_nu); returns (j, b, db, d2b)."""
    g = _f64(grid)
    b, db, d2b = np.zeros(4), np.zeros(4), np.zeros(4)
    j = lib().tdo_bspline_nu(float(r), _ptr(g), int(P_s), _ptr(b), _ptr(db), _ptr(d2b))
    return j, b, db, d2b


def bc_factors(family, at_cutoff):
    """Boundary-condition rows of a cubic family 1..4 (tdo_bc_factors):
    array [rows][3] of weights on splines 0..2 (origin) or S-3..S-1 (cutoff)."""
    f = np.zeros(9)
    n = lib().tdo_bc_factors(int(family), 1 if at_cutoff else 0, _ptr(f))
    return f.reshape(3, 3)[:n].copy()


def spline_coeffs(sysd: System, u):
    """Effective spline coefficients c = T^T u (P_s + 2 values) of the pair
    parameters u [P_s] under the system's boundary conditions."""
    u = _f64(u)
    c = np.zeros(sysd.P_s + 2)
    lib().tdo_spline_coeffs(C.byref(sysd.c()), _ptr(u), _ptr(c))
    return c


def spline_ops(sysd: System, B):
    """O = T B of the basis sums B [P_s + 2]."""
    B = _f64(B)
    O = np.zeros(sysd.P_s)
    lib().tdo_spline_ops(C.byref(sysd.c()), _ptr(B), _ptr(O))
    return O


def basis_sums(sysd: System, R):
    """B_s (s = 0 .. P_s + 1) of each configuration R [W][N][D] and the
    number of pairs inside the cutoff (tdo_basis_sums)."""
    R = _f64(R).reshape(-1, sysd.N, sysd.D)
    B = np.zeros((R.shape[0], sysd.P_s + 2))
    n_in = np.zeros(R.shape[0], dtype=np.int64)
    cs = sysd.c()
    for w in range(R.shape[0]):
        Rw = np.ascontiguousarray(R[w])
        Bw = np.zeros(sysd.P_s + 2)
        n_in[w] = lib().tdo_basis_sums(C.byref(cs), _ptr(Rw), _ptr(Bw))
        B[w] = Bw
    return B, n_in


def pair_potential(sysd: System, r, time=0.0):
    cs = sysd.c()
    return lib().tdo_pair_potential(C.byref(cs), float(r), float(time))


# ---------------------------------------------------------------- evaluation

def eval_configs(sysd: System, R, uR, uI=None, phiR=0.0, time=0.0):
    """ln psi, O, E_L, gradient, Laplacian and the 9 'other' values per walker."""
    R = _f64(R).reshape(-1, sysd.N, sysd.D)
    W = R.shape[0]
    P = sysd.P
    uR = _f64(uR)
    uI = np.zeros(P) if uI is None else _f64(uI)
    assert uR.size == P and uI.size == P
    out = dict(expo=np.zeros(W), O=np.zeros((W, P)), ER=np.zeros(W), EI=np.zeros(W),
               grad=np.zeros((W, sysd.N, sysd.D, 2)), lap=np.zeros((W, sysd.N, 2)), other=np.zeros((W, 9)))
    cs = sysd.c()
    rc = lib().tdo_eval_batch(C.byref(cs), W, _ptr(R), _ptr(uR), _ptr(uI), float(phiR), float(time),
                              _ptr(out["expo"]), _ptr(out["O"]), _ptr(out["ER"]), _ptr(out["EI"]),
                              _ptr(out["grad"]), _ptr(out["lap"]), _ptr(out["other"]))
    if rc:
        raise MemoryError("oracle evaluation failed")
    return out


def metropolis(sysd: System, R, uR, step, seed, n_samples, n_therm):
    R = _f64(R).copy()
    uR = _f64(uR)
    rec = np.zeros((n_samples, sysd.N, sysd.D))
    cs = sysd.c()
    acc = lib().tdo_metropolis(C.byref(cs), _ptr(R), _ptr(uR), float(step), int(seed), int(n_samples),
                               int(n_therm), _ptr(rec))
    return R, rec, acc


# ---------------------------------------------------------------- TDVP

def moments(O, ER, EI):
    O = _f64(O)
    W, P = O.shape
    mO, mE = np.zeros(P), np.zeros(2)
    S, FR, FI = np.zeros((P, P)), np.zeros(P), np.zeros(P)
    lib().tdo_moments(W, P, _ptr(O), _ptr(_f64(ER)), _ptr(_f64(EI)), _ptr(mO), _ptr(mE), _ptr(S),
                      _ptr(FR), _ptr(FI))
    return dict(mO=mO, mE=mE, S=S, FR=FR, FI=FI)


def cholesky_ref(A):
    A = _f64(A).copy()
    bad = lib().tdo_cholesky_ref(A.shape[0], _ptr(A))
    return np.tril(A), bad


def solve_param_dot(m, imag_time=1, precondition=0, eps=1e-3, solver_type=0):
    """Solver type 0 (Cholesky with S += eps I) or type 1 (full-pivot QR,
    mean-subtracted solution; eps unused).  'bad' = non-positive pivots
    (type 0) or the rank deficit (type 1)."""
    P = m["mO"].size
    S = _f64(m["S"]).copy()
    xR, xI = np.zeros(P), np.zeros(P)
    pR, pI = C.c_double(0), C.c_double(0)
    # the converted arrays are held in locals for the duration of the call
    # (a pointer into a temporary copy would dangle)
    mO, mE, FR, FI = _f64(m["mO"]), _f64(m["mE"]), _f64(m["FR"]), _f64(m["FI"])
    args = (_ptr(S), _ptr(mO), _ptr(mE), _ptr(FR), _ptr(FI), _ptr(xR), _ptr(xI), C.byref(pR), C.byref(pI))
    if solver_type == 1:
        bad = lib().tdo_solve_param_dot_qr(P, int(imag_time), int(precondition), *args)
    else:
        bad = lib().tdo_solve_param_dot(P, int(imag_time), int(precondition), float(eps), *args)
    return dict(xR=xR, xI=xI, phiDotR=pR.value, phiDotI=pI.value, bad=bad)


def build_rhs(m, imag_time):
    """Right-hand sides (bR, bI) of src/TDVMC.cpp:1506-1537 from centred moments."""
    P = m["mO"].size
    bR, bI = np.zeros(P), np.zeros(P)
    mO, mE, FR, FI = _f64(m["mO"]), _f64(m["mE"]), _f64(m["FR"]), _f64(m["FI"])
    lib().tdo_build_rhs(P, int(imag_time), _ptr(mO), _ptr(mE), _ptr(FR), _ptr(FI), _ptr(bR), _ptr(bI))
    return bR, bI


def phi_dot(m, imag_time, xR, xI):
    """phi-dot (R, I) of src/TDVMC.cpp:1658-1682 on given alpha-dot."""
    P = m["mO"].size
    pR, pI = C.c_double(0), C.c_double(0)
    mO, xR, xI = _f64(m["mO"]), _f64(xR), _f64(xI)
    lib().tdo_phi_dot(P, int(imag_time), _ptr(mO), float(m["mE"][0]), _ptr(xR), _ptr(xI), C.byref(pR),
                      C.byref(pI))
    return pR.value, pI.value


def ode_step(f, x0, ode, dt):
    """One step of ODE type `ode` for x' = f(x) on the packed parameter vector
    (the stage structure of the reference's integrators):
      0  CalculateNextParametersEuler (src/TDVMC.cpp:1834-1853): x0 + dt k1;
      1  CalculateNextParametersPC (:1855-1910): predictor x0 + dt k1, one
         corrector x0 + dt/2 (k1 + f(xp));
      5  CalculateNextParametersPCReuseSamples (:1912-1967): six correctors;
      2, 3  CalculateNextParametersRK4 (:1969-2035) / ...ReuseSamples
         (:2037-2103): classic RK4, x0 + dt/6 (k1 + 2 k2 + 2 k3 + k4) with
         stages at x0 + dt/2 k1, x0 + dt/2 k2, x0 + dt k3.
    Which samples each stage evaluates is the caller's f; the arithmetic of
    the step is this function's."""
    x0 = np.asarray(x0, dtype=np.float64)
    if ode == 0:
        return x0 + dt * f(x0)
    if ode in (2, 3):
        k1 = f(x0)
        k2 = f(x0 + dt / 2 * k1)
        k3 = f(x0 + dt / 2 * k2)
        k4 = f(x0 + dt * k3)
        return x0 + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    if ode in (1, 5):
        k1 = f(x0)
        xp = x0 + dt * k1
        for _ in range(6 if ode == 5 else 1):
            xp = x0 + (k1 + f(xp)) * (dt / 2)
        return xp
    raise ValueError(f"ODE type {ode}")


def fpqr_solve(A, b):
    """Full-pivot Householder QR solve of A x = b; returns (x, rank)."""
    A = _f64(A).copy()
    n = A.shape[0]
    x = np.zeros(n)
    rank = lib().tdo_fpqr_solve(n, _ptr(A), _ptr(_f64(b).copy()), _ptr(x))
    return x, rank


# ---------------------------------------------------------------- g(r), S(k)

def gr_grid(D, P_s, rc, bins):
    """Pair-distribution grid (tdo_gr_grid): (gmax, spacing, centres, scaling)."""
    gmax, sp = C.c_double(), C.c_double()
    n = lib().tdo_gr_grid(D, P_s, rc, bins, C.byref(gmax), C.byref(sp), None, None)
    cen, scl = np.zeros(max(n, 0)), np.zeros(max(n, 0))
    if n > 0:
        lib().tdo_gr_grid(D, P_s, rc, bins, C.byref(gmax), C.byref(sp), _ptr(cen), _ptr(scl))
    return gmax.value, sp.value, cen, scl


def kshells(D, nshell=50):
    """First nshell wave-vector shells (tdo_kshells): (vectors [V][D], counts, |n|^2)."""
    cap = 8192
    vec = np.zeros(cap * D, dtype=np.int32)
    cnt = np.zeros(nshell, dtype=np.int32)
    n2 = np.zeros(nshell, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    V = lib().tdo_kshells(D, nshell, cap, vec.ctypes.data_as(ip), cnt.ctypes.data_as(ip), n2.ctypes.data_as(ip))
    if V < 0:
        raise ValueError("too many vectors")
    return vec[:V * D].reshape(V, D), cnt, n2


def observables(R, L, P_s, rc, bins, nshell=50):
    """g(r) [W][count] and S(k) [W][nshell] of configurations R [W][N][D]
    (tdo_observables, BosonsBulk::CalculateAdditionalSystemProperties)."""
    R = _f64(R)
    W, N, D = R.shape
    _, _, cen, _ = gr_grid(D, P_s, rc, bins)
    gr = np.zeros((W, len(cen)))
    sk = np.zeros((W, nshell))
    for w in range(W):
        Rw = _f64(R[w])
        g = np.zeros(len(cen))
        s = np.zeros(nshell)
        lib().tdo_observables(N, D, L, P_s, rc, bins, nshell, _ptr(Rw), _ptr(g), _ptr(s))
        gr[w], sk[w] = g, s
    return gr, sk


# ---------------------------------------------------------------- helpers

def kvectors(D, K, L):
    """K shortest non-zero reciprocal-lattice vectors 2 pi n / L, ordered by
    (|n|^2, lexicographic n) -- both k and -k are kept."""
    if K == 0:
        return np.zeros((0, D))
    m = 1
    while True:
        vs = [v for v in itertools.product(range(-m, m + 1), repeat=D) if any(v)]
        vs.sort(key=lambda v: (sum(x * x for x in v), v))
        if len(vs) >= K and sum(x * x for x in vs[K - 1]) < (m + 1) ** 2:
            break
        m += 1
    n = np.array(vs[:K], dtype=np.float64)
    return n * (2.0 * np.pi / L)
