"""Pins the parts of the CPU oracle that the reference cannot: mathematics
fixes them.  Each check is chosen so that a plausible slip (a dropped term,
a sign, an index shift, a transposed operand) fails it.

* B-spline basis: Cox-de Boor recursion on the knot sequence, partition of
  unity, finite differences of the first and second derivative;
* boundary-condition map: zero slope at r = 0 and at the cutoff;
* O_k = d ln psi / d alpha_k: ln psi is linear in alpha, so
  ln psi(alpha + e_k) - ln psi(alpha) = O_k;
* gradient / Laplacian of ln psi (pair + k-space + one-body): central finite
  differences in the coordinates;
* rho_k against a direct numpy sum, rho_{-k} = conj(rho_k);
* E_L: u = 0 gives E_L = V; the Gaussian one-body factor at its optimal
  width in a harmonic trap gives E_L = N D omega / 2 for every configuration;
* O_k of a 64-particle 3-D case against a 40-digit evaluation;
* moments against numpy covariance; Cholesky / solves against numpy;
  the reference's non-positive-pivot behaviour;
* Metropolis: two particles on a ring sample |psi|^2 = exp(2 u(r)).
"""
import math

import numpy as np
import pytest

from oracle import oracle as o


def cox_de_boor(knots, s, k, x):
    if k == 0:
        return 1.0 if knots[s] <= x < knots[s + 1] else 0.0
    a = b = 0.0
    if knots[s + k] != knots[s]:
        a = (x - knots[s]) / (knots[s + k] - knots[s]) * cox_de_boor(knots, s, k - 1, x)
    if knots[s + k + 1] != knots[s + 1]:
        b = (knots[s + k + 1] - x) / (knots[s + k + 1] - knots[s + 1]) * cox_de_boor(knots, s + 1, k - 1, x)
    return a + b


def test_bspline_cox_de_boor_and_unity():
    P, rc = 9, 3.7
    h = rc / (P - 1)
    knots = [(i - 3) * h for i in range(P + 6)]
    rng = np.random.default_rng(0)
    for r in np.concatenate([rng.uniform(0, rc, 200), [0.0, h, 2 * h, rc * 0.999999]]):
        j, b, db, d2b = o.bspline(r, h, P)
        full = np.array([cox_de_boor(knots, s, 3, r) for s in range(P + 2)])
        mine = np.zeros(P + 2)
        mine[j:j + 4] = b
        assert np.allclose(mine, full, atol=1e-13), r
        assert abs(b.sum() - 1.0) < 1e-14
        assert abs(db.sum()) < 1e-12 and abs(d2b.sum()) < 1e-10
        eps = 1e-5
        if eps < r < rc - eps and abs(r / h - round(r / h)) > 1e-3:
            _, bp, dbp, _ = o.bspline(r + eps, h, P)
            _, bm, dbm, _ = o.bspline(r - eps, h, P)
            assert np.allclose((bp - bm) / (2 * eps), db, atol=1e-7)
            assert np.allclose((dbp - dbm) / (2 * eps), d2b, atol=1e-5)


def _u_of_r(s, u, r, eps=0.0):
    # pair function through the oracle: two particles at distance r, 1-D box
    R = np.array([[[0.0], [r]]])
    return o.eval_configs(s, R, u)["expo"][0]


def test_boundary_conditions_zero_slope():
    P, L = 10, 12.0
    s = o.System(N=2, D=1, L=L, P_s=P)
    rng = np.random.default_rng(1)
    u = rng.normal(size=P)
    eps = 1e-6
    # slope at r = 0 (symmetric one-sided difference) and at r = rc
    d0 = (_u_of_r(s, u, 2 * eps) - _u_of_r(s, u, eps)) / eps
    drc = (_u_of_r(s, u, L / 2 - eps) - _u_of_r(s, u, L / 2 - 2 * eps)) / eps
    assert abs(d0) < 1e-4 and abs(drc) < 1e-4
    # a generic interior point has non-zero slope
    dmid = (_u_of_r(s, u, 2.0 + eps) - _u_of_r(s, u, 2.0 - eps)) / (2 * eps)
    assert abs(dmid) > 1e-3


def _bulk(D=3, N=9, L=4.0, P=8, K=6, one_body=0, seed=3):
    kv = o.kvectors(D, K, L) if K else np.zeros((0, D))
    s = o.System(N=N, D=D, L=L, P_s=P, kvec=kv, one_body=one_body, hbar2_2m=0.7, pot_type=o.POT_GAUSS,
                 pot=(0.6, 1.3))
    rng = np.random.default_rng(seed)
    R = rng.uniform(-L / 2, L / 2, (2, N, D))
    uR = rng.normal(0, 0.3, s.P)
    uI = rng.normal(0, 0.2, s.P)
    if one_body:
        uR[0], uI[0] = 0.3, 0.05
    return s, R, uR, uI


@pytest.mark.parametrize("D", [1, 2, 3])
def test_O_is_derivative_of_log_psi(D):
    s, R, uR, uI = _bulk(D=D)
    base = o.eval_configs(s, R, uR)
    last_pair = s.P_s - 1
    for k in range(s.P):
        e = np.zeros(s.P)
        e[k] = 1.0
        up = o.eval_configs(s, R, uR + e)
        diff = up["expo"] - base["expo"]
        expect = base["O"][:, k].copy()
        if k == last_pair:
            # the reference's outer-pair term uR[last] * outerSum is not part of O
            outer = diff - expect
            assert np.allclose(outer, np.round(outer), atol=1e-8)
            continue
        assert np.allclose(diff, expect, rtol=1e-9, atol=1e-9), k


@pytest.mark.parametrize("D,one_body,L", [(1, 0, 5.0), (2, 0, 4.0), (3, 0, 4.0), (2, 1, 40.0), (3, 1, 40.0)])
def test_grad_lap_finite_differences(D, one_body, L):
    s, R, uR, uI = _bulk(D=D, one_body=one_body, L=L, K=0 if one_body else 6)
    if one_body:
        # open space: no minimum image, spline cutoff inside the cloud
        s = o.System(N=s.N, D=D, L=0.0, rc=3.0, P_s=s.P_s, one_body=1, hbar2_2m=0.7)
        R = R / 10.0
    out = o.eval_configs(s, R, uR, uI)
    h = 1e-4
    for part, u in ((0, uR), (1, uI)):
        def lnpsi(Rx):
            return o.eval_configs(s, Rx, u)["expo"]
        g = np.zeros((R.shape[0], s.N, D))
        lap = np.zeros((R.shape[0], s.N))
        f0 = lnpsi(R)
        for n in range(s.N):
            for a in range(D):
                Rp, Rm = R.copy(), R.copy()
                Rp[:, n, a] += h
                Rm[:, n, a] -= h
                fp, fm = lnpsi(Rp), lnpsi(Rm)
                g[:, n, a] = (fp - fm) / (2 * h)
                lap[:, n] += (fp - 2 * f0 + fm) / (h * h)
        assert np.allclose(out["grad"][..., part], g, atol=2e-6, rtol=1e-6)
        assert np.allclose(out["lap"][..., part], lap, atol=2e-3, rtol=1e-4)


def test_rho_k_direct_and_conjugate():
    s, R, uR, uI = _bulk(D=3, K=26)
    out = o.eval_configs(s, R, uR)
    off = s.P_s
    for w in range(R.shape[0]):
        rho = np.exp(1j * R[w] @ s.kvec.T).sum(axis=0)
        assert np.allclose(out["O"][w, off:], np.abs(rho) ** 2, rtol=1e-12)
    # the 26 shortest vectors are the 3 shells |n|^2 = 1, 2, 3 and come in +- pairs
    n = np.round(s.kvec * s.L / (2 * np.pi)).astype(int)
    assert {tuple(v) for v in n} == {tuple(-v) for v in n}
    rho = np.exp(1j * R[0] @ s.kvec.T).sum(axis=0)
    for i, v in enumerate(n):
        j = [tuple(x) for x in n].index(tuple(-v))
        assert np.allclose(rho[j], np.conj(rho[i]))


def test_energy_zero_jastrow_equals_potential():
    s, R, uR, uI = _bulk(D=3, K=4)
    z = np.zeros(s.P)
    out = o.eval_configs(s, R, z, z)
    for w in range(R.shape[0]):
        V = 0.0
        for n in range(s.N):
            for i in range(n):
                r, _ = o.displacement(R[w, n], R[w, i], s.L)
                if r <= s.L / 2:
                    V += 1.3 * np.exp(-0.5 * (r / 0.6) ** 2)
        assert abs(out["ER"][w] - V) < 1e-12 * max(1, abs(V))
        assert abs(out["EI"][w]) < 1e-14


@pytest.mark.parametrize("D,omega", [(2, 1.0), (3, 1.7), (1, 0.4)])
def test_trap_closed_form(D, omega):
    N = 7
    s = o.System(N=N, D=D, L=0.0, rc=2.0, P_s=6, one_body=1, hbar2_2m=0.5, trap_omega=omega)
    rng = np.random.default_rng(5)
    R = rng.normal(0, 1.0, (5, N, D))
    u = np.zeros(s.P)
    u[0] = omega / 2.0          # alpha = omega / (4 hbar^2/2m) with hbar^2/2m = 1/2
    out = o.eval_configs(s, R, u)
    assert np.allclose(out["ER"], N * D * omega / 2.0, rtol=1e-12)
    # off the optimum E_L depends on R
    u[0] *= 1.1
    out = o.eval_configs(s, R, u)
    assert np.std(out["ER"]) > 1e-3


def test_O_against_high_precision():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    D, N, L, P = 3, 64, 4.0, 40
    rng = np.random.default_rng(1)
    R = rng.uniform(-L / 2, L / 2, (N, D))
    s = o.System(N=N, D=D, L=L, P_s=P)
    out = o.eval_configs(s, R[None], np.zeros(P))
    h = mpmath.mpf(L) / 2 / (P - 1)
    B = [mpmath.mpf(0)] * (P + 2)
    for n in range(N):
        for i in range(n):
            d = [mpmath.mpf(R[n, a]) - mpmath.mpf(R[i, a]) for a in range(D)]
            d = [x - L * int(x / L + (0.5 if x >= 0 else -0.5)) for x in d]
            rr = mpmath.sqrt(sum(x * x for x in d))
            if rr <= L / 2:
                x = rr / h
                j = min(int(mpmath.floor(x)), P - 2)
                t = x - j
                u = 1 - t
                b = [u ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]
                for m in range(4):
                    B[j + m] += b[m]
    Oh = [B[1], B[0] + B[2]] + [B[k + 1] for k in range(2, P - 2)] + [B[P - 1] + B[P + 1], B[P]]
    Oh = np.array([float(x) for x in Oh])
    assert np.max(np.abs(out["O"][0] - Oh) / np.abs(Oh)) < 1e-13


def test_moments_vs_numpy():
    rng = np.random.default_rng(2)
    W, P = 500, 7
    O = rng.normal(3.0, 1.0, (W, P)) @ rng.normal(size=(P, P))
    ER = rng.normal(size=W) + O[:, 0]
    EI = rng.normal(size=W)
    m = o.moments(O, ER, EI)
    assert np.allclose(m["mO"], O.mean(0))
    assert np.allclose(m["S"], np.cov(O.T, bias=True), rtol=1e-12, atol=1e-12)
    assert np.allclose(m["FR"], ((O - O.mean(0)) * (ER - ER.mean())[:, None]).mean(0), rtol=1e-12)
    assert np.allclose(m["FI"], ((O - O.mean(0)) * (EI - EI.mean())[:, None]).mean(0), rtol=1e-12)
    assert np.allclose(m["S"], m["S"].T)


def test_cholesky_and_solve_vs_numpy():
    rng = np.random.default_rng(3)
    P = 12
    A = rng.normal(size=(P, P))
    S = A @ A.T + 0.1 * np.eye(P)
    Lf, bad = o.cholesky_ref(S)
    assert bad == 0
    assert np.allclose(Lf, np.linalg.cholesky(S), rtol=1e-12, atol=1e-12)
    m = dict(mO=rng.normal(size=P), mE=np.array([1.5, -0.3]), S=S, FR=rng.normal(size=P), FI=rng.normal(size=P))
    for imag in (0, 1):
        for pre in (0, 1):
            res = o.solve_param_dot(m, imag_time=imag, precondition=pre, eps=1e-3)
            if imag == 0:
                bR, bI = m["FI"], -m["FR"]
            else:
                bR, bI = -m["FR"], -(m["FI"] + m["mO"] * m["mE"][1])
            if pre:
                d = np.sqrt(np.diag(S))
                Sp = S / np.outer(d, d) + 1e-3 * np.eye(P)
                yR, yI = np.linalg.solve(Sp, bR / d), np.linalg.solve(Sp, bI / d)
                phR, phI = -m["mO"] @ yR, -m["mO"] @ yI    # phi-dot uses the scaled solution
                xR, xI = yR / d, yI / d
            else:
                Sp = S + 1e-3 * np.eye(P)
                xR, xI = np.linalg.solve(Sp, bR), np.linalg.solve(Sp, bI)
                phR, phI = -m["mO"] @ xR, -m["mO"] @ xI
            if imag == 0:
                phI -= m["mE"][0]
            else:
                phR -= m["mE"][0]
            assert np.allclose(res["xR"], xR, rtol=1e-10, atol=1e-12)
            assert np.allclose(res["xI"], xI, rtol=1e-10, atol=1e-12)
            assert abs(res["phiDotR"] - phR) < 1e-10 * max(1, abs(phR))
            assert abs(res["phiDotI"] - phI) < 1e-10 * max(1, abs(phI))


def test_cholesky_nonpositive_pivot_keeps_diagonal():
    S = np.array([[4.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 9.0]])   # rank-deficient leading block
    Lf, bad = o.cholesky_ref(S)
    assert bad == 1
    assert Lf[0, 0] == 2.0 and Lf[1, 0] == 1.0
    assert Lf[1, 1] == 1.0      # unreduced diagonal kept
    assert Lf[2, 2] == 3.0


def test_metropolis_two_particles_ring():
    L, P = 6.0, 6
    s = o.System(N=2, D=1, L=L, P_s=P)
    u = np.array([-1.2, -0.8, -0.2, 0.1, 0.05, 0.0])
    R0 = np.array([[0.0], [2.0]])
    _, rec, acc = o.metropolis(s, R0, u, step=1.0, seed=9, n_samples=40000, n_therm=5)
    r = np.array([o.displacement(x[0], x[1], L)[0] for x in rec])
    edges = np.linspace(0, L / 2, 13)
    hist, _ = np.histogram(r, bins=edges)
    grid = np.linspace(0, L / 2, 3001)
    ug = np.array([_u_of_r(s, u, x) for x in grid])
    dens = np.exp(2 * ug)
    mass = np.array([np.trapezoid(dens[(grid >= a) & (grid <= b)], grid[(grid >= a) & (grid <= b)])
                     for a, b in zip(edges[:-1], edges[1:])])
    expect = mass / mass.sum() * len(r)
    # successive records are correlated; allow a generous chi-square margin
    chi2 = np.sum((hist - expect) ** 2 / expect)
    assert chi2 < 60, (hist, expect)
    assert 0.2 < acc / (40000 * 5) < 0.95


def test_full_pivot_qr_vs_numpy():
    rng = np.random.default_rng(11)
    for n in (1, 5, 17):
        A = rng.normal(size=(n, n))        # non-symmetric: a wrong row/column un-permutation shows
        b = rng.normal(size=n)
        x, rank = o.fpqr_solve(A, b)
        assert rank == n
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # rank-deficient but consistent: the returned x solves the system
    n = 9
    B = rng.normal(size=(n, n - 2))
    A = B @ rng.normal(size=(n - 2, n))
    b = A @ rng.normal(size=n)
    x, rank = o.fpqr_solve(A, b)
    assert rank == n - 2
    assert np.allclose(A @ x, b, rtol=1e-9, atol=1e-9)


def test_solver_type1_vs_numpy():
    """Solver type 1 (src/TDVMC.cpp:1763-1822): solve, subtract the mean of
    each solution, phi-dot from that vector, then undo the scaling; S gets
    +0.002 I only together with the scaling."""
    rng = np.random.default_rng(5)
    P = 10
    A = rng.normal(size=(P, P))
    S = A @ A.T + 0.1 * np.eye(P)
    m = dict(mO=rng.normal(size=P), mE=np.array([0.7, 0.2]), S=S, FR=rng.normal(size=P), FI=rng.normal(size=P))
    for imag in (0, 1):
        for pre in (0, 1):
            res = o.solve_param_dot(m, imag_time=imag, precondition=pre, solver_type=1)
            if imag == 0:
                bR, bI = m["FI"], -m["FR"]
            else:
                bR, bI = -m["FR"], -(m["FI"] + m["mO"] * m["mE"][1])
            d = np.sqrt(np.diag(S)) if pre else np.ones(P)
            Sp = S / np.outer(d, d) + (0.002 * np.eye(P) if pre else 0.0)
            yR, yI = np.linalg.solve(Sp, bR / d), np.linalg.solve(Sp, bI / d)
            yR, yI = yR - yR.mean(), yI - yI.mean()
            phR, phI = -m["mO"] @ yR, -m["mO"] @ yI
            if imag == 0:
                phI -= m["mE"][0]
            else:
                phR -= m["mE"][0]
            assert res["bad"] == 0
            assert np.allclose(res["xR"], yR / d, rtol=1e-10, atol=1e-12)
            assert np.allclose(res["xI"], yI / d, rtol=1e-10, atol=1e-12)
            assert abs(res["phiDotR"] - phR) < 1e-10 * max(1, abs(phR))
            assert abs(res["phiDotI"] - phI) < 1e-10 * max(1, abs(phI))


def test_full_pivot_qr_duplicate_columns():
    """An S with two identical rows/columns (the O of k and -k): rank
    deficit 1, the system is still solved, and the pivot order (first
    maximum in column-major order) puts the whole weight on one of the pair."""
    rng = np.random.default_rng(2)
    n = 8
    B = rng.normal(size=(n - 1, 200))
    O = np.vstack([B, B[-1:]])                 # rows n-2 and n-1 identical
    S = O @ O.T / 200
    b = S @ rng.normal(size=n)
    x, rank = o.fpqr_solve(S, b)
    assert rank == n - 1
    assert np.allclose(S @ x, b, rtol=1e-9, atol=1e-9)
    assert x[n - 1] == 0.0 or x[n - 2] == 0.0


# ---------------------------------------------------------------- right-hand-side signs
def _two_body_quadrature(s, uR, uI, M=20000):
    """Exact expectations of the 1-D two-particle system on a ring: psi
    depends on the minimum-image distance only, so <A> = int |psi|^2 A dr /
    int |psi|^2 dr over r in (-L/2, L/2) (the centre of mass integrates out).
    Midpoint rule on a periodic, C^2 integrand."""
    L = s.L
    r = (np.arange(M) + 0.5) * (L / M) - L / 2
    R = np.zeros((M, 2, 1))
    R[:, 1, 0] = r
    out = o.eval_configs(s, R, uR, uI)
    w = np.exp(2.0 * (out["expo"] - out["expo"].max()))
    w /= w.sum()
    return out, w


def _energy_and_moments(s, uR, uI):
    out, w = _two_body_quadrature(s, uR, uI)
    O, ER, EI = out["O"], out["ER"], out["EI"]
    mO, mER, mEI = w @ O, w @ ER, w @ EI
    Oc = O - mO
    m = dict(mO=mO, mE=np.array([mER, mEI]), S=(Oc * w[:, None]).T @ Oc, FR=w @ (Oc * (ER - mER)[:, None]),
             FI=w @ (Oc * (EI - mEI)[:, None]))
    return mER, m, w @ (O * EI[:, None])


@pytest.mark.parametrize("imag", [0, 1])
def test_rhs_signs_by_energy_identities(imag):
    """Pins the signs and centring of the right-hand sides (src/TDVMC.cpp:
    1506-1537) by physics, independently of their formulas: for a hermitian
    H, d<E>/d uR_k = 2 cov(O_k, E_R) and d<E>/d uI_k = 2 <O_k E_I> (exact
    expectations; here a two-body system integrated by quadrature).  The
    TDVP direction then must lower <E> in imaginary time, separately along its
    real and its imaginary part, and keep <E> constant to first order in real
    time.  A flipped sign or a dropped term in b_R or b_I breaks one of these."""
    s = o.System(N=2, D=1, L=6.0, P_s=6, pot_type=o.POT_GAUSS, pot=(0.5, 1.0), hbar2_2m=0.5)
    rng = np.random.default_rng(17)
    uR = np.array([-0.8, -0.5, -0.25, -0.1, -0.03, 0.0]) + rng.normal(0, 0.05, 6)
    uI = rng.normal(0, 0.15, 6)
    E0, m, g = _energy_and_moments(s, uR, uI)
    assert abs(m["mE"][1]) < 1e-7 * abs(E0)     # <E_I> = 0 for hermitian H (quadrature error ~1e-9)
    r = o.solve_param_dot(m, imag_time=imag, precondition=0, eps=1e-3)
    xR, xI = r["xR"], r["xI"]

    def dE(dR, dI, eps=1e-5):
        ep = _energy_and_moments(s, uR + eps * dR, uI + eps * dI)[0]
        em = _energy_and_moments(s, uR - eps * dR, uI - eps * dI)[0]
        return (ep - em) / (2 * eps)

    gR, gI = dE(xR, 0 * xI), dE(0 * xR, xI)
    # the gradient formula on both parts (checks the moments' meaning)
    scale = np.abs(m["FR"]) @ np.abs(xR) + np.abs(g) @ np.abs(xI)
    assert abs(gR - 2 * m["FR"] @ xR) < 1e-6 * scale
    assert abs(gI - 2 * g @ xI) < 1e-6 * scale
    if imag:
        assert gR < -1e-3 * scale and gI < -1e-6 * scale       # descent on both parts
    else:
        assert abs(gR + gI) < 1e-6 * scale                    # energy conserved
        assert abs(gR) > 1e-3 * scale                          # (each part alone is not zero)


def test_ode_steps_on_a_linear_ode():
    """oracle.ode_step on x' = A x, where each scheme has a closed form:
    Euler (I + hA), RK4 the degree-4 Taylor polynomial of exp(hA), the
    predictor-corrector with one corrector Heun's (I + hA + (hA)^2/2), and
    with six correctors the sixth iterate of y = c + (hA/2) y,
    c = (I + hA/2) x0, from y0 = (I + hA) x0."""
    rng = np.random.default_rng(8)
    n, h = 7, 0.13
    A = rng.normal(0, 1, (n, n))
    x0 = rng.normal(0, 1, n)
    f = lambda x: A @ x  # noqa: E731
    I = np.eye(n)
    hA = h * A
    taylor = lambda k: sum(np.linalg.matrix_power(hA, j) / math.factorial(j) for j in range(k + 1))  # noqa: E731
    assert np.allclose(o.ode_step(f, x0, 0, h), (I + hA) @ x0, rtol=1e-14, atol=1e-14)
    for ode in (2, 3):
        assert np.allclose(o.ode_step(f, x0, ode, h), taylor(4) @ x0, rtol=1e-13, atol=1e-13)
    assert np.allclose(o.ode_step(f, x0, 1, h), taylor(2) @ x0, rtol=1e-13, atol=1e-13)
    Mh = hA / 2
    c = (I + Mh) @ x0
    y = (I + hA) @ x0
    for _ in range(6):
        y = c + Mh @ y
    assert np.allclose(o.ode_step(f, x0, 5, h), y, rtol=1e-13, atol=1e-13)
    # and that is the trapezoidal rule up to (hA/2)^6
    trap = np.linalg.solve(I - Mh, (I + Mh) @ x0)
    assert np.linalg.norm(y - trap) < 2 * np.linalg.norm(np.linalg.matrix_power(Mh, 6), 2) * np.linalg.norm(x0) * 10
