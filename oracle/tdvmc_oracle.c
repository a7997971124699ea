/*
 * tdvmc_oracle.c -- CPU restatement (FP64, scalar C) of the TDVMC sampling and
 * wavefunction-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against; it is never linked into, called by, or shipped as part of
 * the product path (tdvmc_amd / libtdvmc.so).  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it.
 *
 * Every function names the reference code it restates (paths relative to the
 * reference repository root).  How each function is pinned:
 *   tdo_nic_coord / tdo_displacement  -> 48 known answers of src/test/Tests.h:76-156
 *   tdo_bspline                        -> Cox-de Boor recursion, partition of unity,
 *                                         finite differences (tests/test_oracle_math.py)
 *   tdo_bspline_nu (non-uniform knots) -> golden vectors of the compiled reference's
 *                                         BosonsBulk on NURBS_GRID knots
 *                                         (tests/golden/eval_nurbs_*.npz), partition of
 *                                         unity, finite differences, and the uniform
 *                                         basis on a uniform grid (tests/test_knots_cpu.py)
 *   tdo_bc_factors / tdo_spline_coeffs -> the reference's SetBoundaryConditions3_1D_OR_k /
 *   / tdo_spline_ops                      _CO_k rows (tests/golden/bc_factors.json and
 *                                         oracle/_ref live), the conditions each family
 *                                         imposes on u (u' = 0, u'' = 0 at 0 and rc), T and
 *                                         T^T transposes (tests/test_bc_families_cpu.py);
 *                                         family pair (2, 2) through the eval goldens
 *   tdo_basis_sums                     -> partition of unity (sum_s B_s = pairs within
 *                                         rc) and the BC map onto tdo_eval's O
 *                                         (tests/test_gpu_nurbs.py)
 *   tdo_eval                           -> golden vectors written by the compiled
 *                                         reference (oracle/_ref, BosonsBulk) plus
 *                                         linearity of ln(psi) in alpha, finite
 *                                         differences of grad/lap, closed-form E_L
 *                                         of the Gaussian trap, brute-force rho_k
 *   tdo_moments                        -> numpy covariance on the same samples
 *   tdo_cholesky_ref / tdo_chol_solve  -> numpy.linalg.cholesky / solve, and the
 *                                         reference's non-positive-pivot behaviour
 *   tdo_build_rhs / tdo_phi_dot /      -> golden vectors of the reference driver's own
 *   tdo_solve_param_dot(_qr) /            BuildSystemOfEquationsForParameters,
 *   tdo_fpqr_solve                        SolveForParametersDot (Cholesky and Eigen
 *                                         FullPivHouseholderQR, with and without scaling)
 *                                         and CalculatePhiDot (oracle/_ref/libref_driver.so,
 *                                         tests/golden/tdvp_cases.npz); the signs of the
 *                                         right-hand sides also by the energy identities of
 *                                         a two-body system integrated by quadrature
 *                                         (tests/test_oracle_math.py)
 *   tdo_pair_potential                 -> square well / quench: reference eval goldens;
 *                                         Lennard-Jones: LJ::GetPotential values
 *                                         (tests/golden/driver_cases.json)
 *   tdo_metropolis                     -> exact two-particle distribution exp(2u(r));
 *                                         <E_L> against 256 chains of the compiled
 *                                         reference (tests/golden/sample_bulk3d_sw16.json)
 *   (oracle.py) ode_step               -> exact solution of a linear ODE for every ODE
 *                                         type (tests/test_oracle_math.py); the gauge shift
 *                                         of AdjustParameters against the reference driver
 *                                         (tests/golden/driver_cases.json)
 *   tdo_gr_grid / tdo_kshells /        -> golden vectors of the compiled reference's
 *   tdo_observables                       CalculateAdditionalSystemProperties, the
 *                                         reference's kVectors / kNorm files, S(k) of a
 *                                         lattice in closed form, g(r) pair counts
 *                                         (tests/test_oracle_observables.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int N;              /* particles */
    int D;              /* dimension 1..3 */
    double L;           /* periodic box length; <= 0 -> open space (no minimum image) */
    int P_s;            /* pair-spline parameters (>= 4) */
    double rc;          /* spline cutoff (BosonsBulk maxDistance) */
    int K;              /* k-space Jastrow terms */
    const double* kvec; /* [K][D] wave vectors */
    int one_body;       /* 1: parameter 0 multiplies O = -sum_n |r_n|^2 */
    double hbar2_2m;    /* hbar^2 / 2m */
    int pot_type;       /* 0 none, 1 square well, 2 Gaussian, 3 Lennard-Jones */
    int n_pot;          /* number of valid entries in pot */
    double pot[5];      /* square/Gaussian: a, b, t_quench, a', b';  LJ: sigma, eps */
    double pot_cutoff;  /* pair potential applies for r <= pot_cutoff */
    double trap_omega;  /* external 0.5*omega^2*|r|^2 per particle (0 = none) */
    const double* grid; /* NULL: uniform knots; else the P_s knots of NURBS_GRID
                           (0 = g_0 < ... < g_{P_s-1}, rc = g_{P_s-1}) */
    int bc_origin;      /* boundary-condition family at r = 0: 1..4 (0 -> 2, BosonsBulk's) */
    int bc_cutoff;      /* boundary-condition family at r = rc: 1..4 (0 -> 2) */
} tdo_system;

/* ------------------------------------------------------------------ */
/* minimum image                                                       */
/* ------------------------------------------------------------------ */

/* Follows src/Utils.cpp:266-281 (GetCoordinateNIC): round half away from zero,
 * nudge an exact +-L/2 result inwards by 1e-10. */
double tdo_nic_coord(double x, double L)
{
    double Lr = 1.0 / L;
    int k = (int)(x * Lr + ((x >= 0.0) ? 0.5 : -0.5));
    double res = x - k * L;
    if (res == L / 2.0)
        res -= 1e-10;
    else if (res == -L / 2.0)
        res += 1e-10;
    return res;
}

/* Follows src/Utils.cpp:313-374 (VectorDiffNIC_*, VectorDisplacementNIC_*):
 * out = NIC(ri - rj), returns |out|.  L <= 0 means plain Euclidean. */
double tdo_displacement(const double* ri, const double* rj, int D, double L, double* out)
{
    double s = 0.0;
    for (int a = 0; a < D; a++) {
        double d = ri[a] - rj[a];
        if (L > 0.0)
            d = tdo_nic_coord(d, L);
        out[a] = d;
        s += d * d;
    }
    return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* uniform cubic B-splines                                             */
/* ------------------------------------------------------------------ */

/* Uniform knots t_i = (i-3) h, i = 0..P_s+5, h = rc/(P_s-1) as in
 * src/PhysicalSystems/BosonsBulk.cpp:61-67; spline s lives on knots s..s+4.
 * For r in knot interval j ([j h, (j+1) h], j = 0..P_s-2) the non-zero splines
 * are s = j..j+3; b[m] is the value of spline j+m, db/d2b its first/second
 * derivative with respect to r.  Returns j.
 * The interval of r = rc is the last one (j = P_s-2, t = 1). */
int tdo_bspline(double r, double h, int P_s, double* b, double* db, double* d2b)
{
    double x = r / h;
    int j = (int)floor(x);
    if (j > P_s - 2) j = P_s - 2;
    if (j < 0) j = 0;
    double t = x - j;
    double u = 1.0 - t;
    double t2 = t * t, t3 = t2 * t;
    b[0] = u * u * u / 6.0;
    b[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    b[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    b[3] = t3 / 6.0;
    if (db) {
        double ih = 1.0 / h;
        db[0] = -0.5 * u * u * ih;
        db[1] = (1.5 * t2 - 2.0 * t) * ih;
        db[2] = (-1.5 * t2 + t + 0.5) * ih;
        db[3] = 0.5 * t2 * ih;
    }
    if (d2b) {
        double ih2 = 1.0 / (h * h);
        d2b[0] = u * ih2;
        d2b[1] = (3.0 * t - 2.0) * ih2;
        d2b[2] = (-3.0 * t + 1.0) * ih2;
        d2b[3] = t * ih2;
    }
    return j;
}

/* Non-uniform knots (BosonsBulk::SetNodes, src/PhysicalSystems/BosonsBulk.cpp:35-44):
 * the P_s grid values g, extended by three mirrored knots at each end,
 *   n = (-g_3, -g_2, -g_1, g_0, ..., g_{P-1}, 2g_{P-1}-g_{P-2}, 2g_{P-1}-g_{P-3}, 2g_{P-1}-g_{P-4}),
 * cubic B-spline s on n_s..n_{s+4}.  r lies in grid interval j (the knot
 * interval (g_j, g_{j+1}] of lower_bound - 1, src/PhysicalSystems/BosonsBulk.cpp:197,
 * clamped to 0..P-2), where splines s = j..j+3 are non-zero.  The values come
 * from the Cox-de Boor recursion on that interval (C. de Boor, "A Practical
 * Guide to Splines", ch. IX: B_{i,k} from B_{i,k-1} and B_{i+1,k-1}), the
 * derivatives from B'_{i,3} = 3 (B_{i,2}/(n_{i+3}-n_i) - B_{i+1,2}/(n_{i+4}-n_{i+1}))
 * and its analogue for B''.  b/db/d2b as in tdo_bspline.  Returns j. */
static double tdo_ext_knot(const double* g, int P, int i)
{
    if (i < 3) return -g[3 - i];
    if (i < P + 3) return g[i - 3];
    return 2.0 * g[P - 1] - g[P - 1 - (i - (P + 2))];
}

int tdo_bspline_nu(double r, const double* g, int P_s, double* b, double* db, double* d2b)
{
    int k = 0;
    while (k < P_s && g[k] < r) k++;          /* lower_bound */
    int j = k - 1;
    if (j < 0) j = 0;
    if (j > P_s - 2) j = P_s - 2;
    const int e = j + 3;                      /* extended interval n_e .. n_{e+1} */
    double n[8];
    for (int q = 0; q < 8; q++) n[q] = tdo_ext_knot(g, P_s, e - 3 + q);   /* n_{e-3} .. n_{e+4} */
    /* B[o][m]: order-o spline index e - o + m (m = 0..o), non-zero on the interval */
    double B[4][4] = {{0}};
    B[0][0] = 1.0;
    for (int o = 1; o <= 3; o++) {
        for (int m = 0; m <= o; m++) {
            const int i = e - o + m;          /* spline index, knots n_i .. n_{i+o+1} */
            const int li = i - (e - 3);       /* local index into n[] */
            double v = 0.0;
            if (m > 0) v += (r - n[li]) / (n[li + o] - n[li]) * B[o - 1][m - 1];
            if (m < o) v += (n[li + o + 1] - r) / (n[li + o + 1] - n[li + 1]) * B[o - 1][m];
            B[o][m] = v;
        }
    }
    for (int m = 0; m < 4; m++) b[m] = B[3][m];
    if (db || d2b) {
        /* first derivatives of the order-3 splines from the order-2 values, and
         * the order-2 splines' derivatives from the order-1 values */
        double d2[3] = {0, 0, 0};             /* B'_{i,2}, i = e-2 .. e */
        for (int m = 0; m <= 2; m++) {
            const int i = e - 2 + m, li = i - (e - 3);
            double v = 0.0;
            if (m > 0) v += 2.0 * B[1][m - 1] / (n[li + 2] - n[li]);
            if (m < 2) v -= 2.0 * B[1][m] / (n[li + 3] - n[li + 1]);
            d2[m] = v;
        }
        for (int m = 0; m < 4; m++) {
            const int i = e - 3 + m, li = i - (e - 3);
            double v1 = 0.0, v2 = 0.0;
            if (m > 0) {
                v1 += 3.0 * B[2][m - 1] / (n[li + 3] - n[li]);
                v2 += 3.0 * d2[m - 1] / (n[li + 3] - n[li]);
            }
            if (m < 3) {
                v1 -= 3.0 * B[2][m] / (n[li + 4] - n[li + 1]);
                v2 -= 3.0 * d2[m] / (n[li + 4] - n[li + 1]);
            }
            if (db) db[m] = v1;
            if (d2b) d2b[m] = v2;
        }
    }
    return j;
}

/* Boundary-condition factors of the cubic pair spline, the uniform branch
 * of src/SplineFactory.cpp (BosonsBulk passes uniform = true for uniform and
 * NURBS knots alike, src/PhysicalSystems/BosonsBulk.cpp:76-77).  Row i gives
 * parameter i's weights on splines 0, 1, 2 (origin, SetBoundaryConditions3_1D_OR_k,
 * :377-515) or on splines S-3, S-2, S-1 (cutoff, ..._CO_k, :517-656):
 *   1: no condition (identity, three rows)      2: u' = 0 (two rows)
 *   3: u'' = 0 (two rows)                       4: u' = u'' = 0 (one row)
 * Returns the row count, or 0 for an unknown family. */
int tdo_bc_factors(int family, int at_cutoff, double f[3][3])
{
    memset(f, 0, 9 * sizeof(double));
    switch (family) {
    case 1:
        f[0][0] = f[1][1] = f[2][2] = 1.0;
        return 3;
    case 2:
        if (at_cutoff) {
            f[0][0] = f[0][2] = 1.0;
            f[1][1] = 1.0;
        } else {
            f[0][1] = 1.0;
            f[1][0] = f[1][2] = 1.0;
        }
        return 2;
    case 3:
        if (at_cutoff) {
            f[0][0] = 1.0; f[0][2] = -1.0;
            f[1][1] = 1.0; f[1][2] = 2.0;
        } else {
            f[0][0] = -1.0; f[0][2] = 1.0;
            f[1][0] = 2.0; f[1][1] = 1.0;
        }
        return 2;
    case 4:
        f[0][0] = f[0][1] = f[0][2] = 1.0;
        return 1;
    default:
        return 0;
    }
}

/* The parameter -> spline map T of src/PhysicalSystems/BosonsBulk.cpp:158-177
 * (RefreshLocalOperators): the first n0 parameters are the origin rows over
 * B_0..B_2, the next P_s - n0 - n1 are B_3, B_4, ..., the last n1 the cutoff
 * rows over B_{S-3}..B_{S-1}, S = P_s + 2 (n0 + n1 = 4 keeps the reference's
 * parameter count, :84-91).  With BosonsBulk's families (2, 2):
 *   O_0 = B_1, O_1 = B_0 + B_2, O_k = B_{k+1} (2 <= k <= P_s-3),
 *   O_{P_s-2} = B_{P_s-1} + B_{P_s+1}, O_{P_s-1} = B_{P_s}.
 * tdo_spline_coeffs returns c = T^T u (effective spline coefficients). */
static int bc_rows(const tdo_system* s, double f0[3][3], double f1[3][3], int* n1)
{
    int n0 = tdo_bc_factors(s->bc_origin ? s->bc_origin : 2, 0, f0);
    *n1 = tdo_bc_factors(s->bc_cutoff ? s->bc_cutoff : 2, 1, f1);
    return n0;
}

void tdo_spline_coeffs(const tdo_system* s, const double* u, double* c)
{
    const int P_s = s->P_s, S = P_s + 2;
    double f0[3][3], f1[3][3];
    int n1;
    const int n0 = bc_rows(s, f0, f1, &n1);
    for (int k = 0; k < S; k++) c[k] = 0.0;
    for (int i = 0; i < n0; i++)
        for (int m = 0; m < 3; m++) c[m] += f0[i][m] * u[i];
    for (int i = n0; i < P_s - n1; i++) c[3 + i - n0] += u[i];
    for (int i = 0; i < n1; i++)
        for (int m = 0; m < 3; m++) c[S - 3 + m] += f1[i][m] * u[P_s - n1 + i];
}

void tdo_spline_ops(const tdo_system* s, const double* B, double* O)
{
    const int P_s = s->P_s, S = P_s + 2;
    double f0[3][3], f1[3][3];
    int n1;
    const int n0 = bc_rows(s, f0, f1, &n1);
    for (int i = 0; i < n0; i++) O[i] = f0[i][0] * B[0] + f0[i][1] * B[1] + f0[i][2] * B[2];
    for (int i = n0; i < P_s - n1; i++) O[i] = B[3 + i - n0];
    for (int i = 0; i < n1; i++)
        O[P_s - n1 + i] = f1[i][0] * B[S - 3] + f1[i][1] * B[S - 2] + f1[i][2] * B[S - 1];
}

/* ------------------------------------------------------------------ */
/* pair potentials                                                     */
/* ------------------------------------------------------------------ */

/* square well: src/PhysicalSystems/BosonsBulk.cpp:236-271 (b if r < a);
 * Gaussian: src/PhysicalSystems/BosonsBulkDamped.cpp:237-264 (b exp(-(r/a)^2/2));
 * both switch to (a', b') once time >= t_quench.
 * Lennard-Jones: src/Potentials/LJ.cpp:16-22 with (sigma, eps): 4 eps ((sigma/r)^12 - (sigma/r)^6). */
double tdo_pair_potential(const tdo_system* s, double r, double time)
{
    double a = s->pot[0], b = s->pot[1];
    if (s->pot_type == 1 || s->pot_type == 2) {
        if (s->n_pot > 2 && time >= s->pot[2]) {
            a = s->pot[3];
            b = s->pot[4];
        }
    }
    switch (s->pot_type) {
    case 1:
        return (r < a) ? b : 0.0;
    case 2: {
        double x = r / a;
        return b * exp(-(x * x) / 2.0);
    }
    case 3: {
        double sr6 = pow(a / r, 6);
        return 4.0 * b * sr6 * (sr6 - 1.0);
    }
    default:
        return 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* evaluation of one configuration                                     */
/* ------------------------------------------------------------------ */

static int tdo_num_params(const tdo_system* s) { return (s->one_body ? 1 : 0) + s->P_s + s->K; }

/* Follows src/PhysicalSystems/BosonsBulk.cpp:179-218 (basis sums, outerSum),
 * :220-336 (sD, sD2, potential), :349-458 (gradient/Laplacian contraction,
 * E_L, otherExpectationValues) and :522-539 (exponent, including the
 * uR[last pair parameter] * outerSum term).  Adds the k-space Jastrow
 * sum_k alpha_k |rho_k|^2 and the one-body factor, which the reference lacks.
 *
 * Outputs (any may be NULL except expo):
 *   O[P], grad[N*D*2] (re, im interleaved per component), lap[N*2],
 *   other[9] in the order of BosonsBulk.cpp:449-457.
 * Returns 0, or -1 on allocation failure. */
int tdo_eval(const tdo_system* s, const double* R, const double* uR, const double* uI,
             double phiR, double time, double* expo, double* O, double* ER, double* EI,
             double* grad, double* lap, double* other)
{
    const int N = s->N, D = s->D, Ps = s->P_s, S = Ps + 2, K = s->K;
    const int off_pair = s->one_body ? 1 : 0, off_k = off_pair + Ps;
    const int P = tdo_num_params(s);
    const double h = s->rc / (Ps - 1);

    double* cR = calloc(S, sizeof(double));
    double* cI = calloc(S, sizeof(double));
    double* B = calloc(S, sizeof(double));
    double* Ol = calloc(P, sizeof(double));
    double* gR = calloc((size_t)N * D, sizeof(double));
    double* gI = calloc((size_t)N * D, sizeof(double));
    double* lR = calloc(N, sizeof(double));
    double* lI = calloc(N, sizeof(double));
    double* rho = calloc(2 * (size_t)(K > 0 ? K : 1), sizeof(double));
    if (!cR || !cI || !B || !Ol || !gR || !gI || !lR || !lI || !rho) {
        free(cR); free(cI); free(B); free(Ol); free(gR); free(gI); free(lR); free(lI); free(rho);
        return -1;
    }
    tdo_spline_coeffs(s, uR + off_pair, cR);
    tdo_spline_coeffs(s, uI + off_pair, cI);

    double outer = 0.0, Vint = 0.0, Vext = 0.0;
    double d[3], b[4], db[4], d2b[4];
    for (int n = 0; n < N; n++) {
        for (int i = 0; i < n; i++) {
            double r = tdo_displacement(R + n * D, R + i * D, D, s->L, d);
            if (r <= s->pot_cutoff)
                Vint += tdo_pair_potential(s, r, time);
            if (r <= s->rc) {
                int j = s->grid ? tdo_bspline_nu(r, s->grid, Ps, b, db, d2b) : tdo_bspline(r, h, Ps, b, db, d2b);
                double upR = 0, upI = 0, uppR = 0, uppI = 0;
                for (int m = 0; m < 4; m++) {
                    B[j + m] += b[m];
                    upR += cR[j + m] * db[m];
                    upI += cI[j + m] * db[m];
                    uppR += cR[j + m] * d2b[m];
                    uppI += cI[j + m] * d2b[m];
                }
                for (int a = 0; a < D; a++) {
                    double e = d[a] / r;
                    gR[n * D + a] += upR * e;
                    gI[n * D + a] += upI * e;
                    gR[i * D + a] -= upR * e;
                    gI[i * D + a] -= upI * e;
                }
                double lfR = uppR + (D - 1.0) / r * upR;
                double lfI = uppI + (D - 1.0) / r * upI;
                lR[n] += lfR; lR[i] += lfR;
                lI[n] += lfI; lI[i] += lfI;
            } else {
                outer += 1.0;
            }
        }
    }
    tdo_spline_ops(s, B, Ol + off_pair);

    /* one-body Gaussian factor: ln psi += alpha * (-sum |r|^2) */
    if (s->one_body) {
        double sr2 = 0.0;
        for (int n = 0; n < N; n++)
            for (int a = 0; a < D; a++) {
                double x = R[n * D + a];
                sr2 += x * x;
                gR[n * D + a] += -2.0 * uR[0] * x;
                gI[n * D + a] += -2.0 * uI[0] * x;
            }
        for (int n = 0; n < N; n++) {
            lR[n] += -2.0 * D * uR[0];
            lI[n] += -2.0 * D * uI[0];
        }
        Ol[0] = -sr2;
    }

    /* external trap */
    if (s->trap_omega != 0.0) {
        for (int n = 0; n < N; n++)
            for (int a = 0; a < D; a++) {
                double x = R[n * D + a];
                Vext += 0.5 * s->trap_omega * s->trap_omega * x * x;
            }
    }

    /* k-space Jastrow: O_k = |rho_k|^2, rho_k = sum_n exp(i k.r_n) */
    for (int k = 0; k < K; k++) {
        const double* kv = s->kvec + k * D;
        double re = 0, im = 0;
        for (int n = 0; n < N; n++) {
            double ph = 0;
            for (int a = 0; a < D; a++) ph += kv[a] * R[n * D + a];
            re += cos(ph);
            im += sin(ph);
        }
        rho[2 * k] = re;
        rho[2 * k + 1] = im;
        Ol[off_k + k] = re * re + im * im;
    }
    for (int n = 0; n < N; n++) {
        for (int k = 0; k < K; k++) {
            const double* kv = s->kvec + k * D;
            double ph = 0, k2 = 0;
            for (int a = 0; a < D; a++) {
                ph += kv[a] * R[n * D + a];
                k2 += kv[a] * kv[a];
            }
            double c = cos(ph), sn = sin(ph);
            /* e^{i k r} * conj(rho) */
            double pre = c * rho[2 * k] + sn * rho[2 * k + 1];
            double pim = sn * rho[2 * k] - c * rho[2 * k + 1];
            double aR = uR[off_k + k], aI = uI[off_k + k];
            for (int a = 0; a < D; a++) {
                double gk = -2.0 * kv[a] * pim;
                gR[n * D + a] += aR * gk;
                gI[n * D + a] += aI * gk;
            }
            double lk = 2.0 * k2 * (1.0 - pre);
            lR[n] += aR * lk;
            lI[n] += aI * lk;
        }
    }

    /* exponent (BosonsBulk.cpp:522-539): boundaryValue = last pair parameter */
    double ex = 0.0;
    for (int k = 0; k < P; k++) ex += uR[k] * Ol[k];
    ex += uR[off_pair + Ps - 1] * outer;

    /* kinetic sums (BosonsBulk.cpp:370-427) */
    double sR1 = 0, sI1 = 0, sR1I1 = 0, sR2 = 0, sI2 = 0;
    for (int n = 0; n < N; n++) {
        double dRI = 0, nR = 0, nI = 0;
        for (int a = 0; a < D; a++) {
            double x = gR[n * D + a], y = gI[n * D + a];
            dRI += x * y;
            nR += x * x;
            nI += y * y;
        }
        sR1I1 += 2.0 * dRI;
        sR1 += nR;
        sI1 += nI;
        sR2 += lR[n];
        sI2 += lI[n];
    }
    double kinR = -(sR1 - sI1 + sR2) * s->hbar2_2m;
    double kinI = -(sR1I1 + sI2) * s->hbar2_2m;

    *expo = ex;
    if (ER) *ER = kinR + Vint + Vext;
    if (EI) *EI = kinI;
    if (O) memcpy(O, Ol, sizeof(double) * P);
    if (grad)
        for (int n = 0; n < N * D; n++) {
            grad[2 * n] = gR[n];
            grad[2 * n + 1] = gI[n];
        }
    if (lap)
        for (int n = 0; n < N; n++) {
            lap[2 * n] = lR[n];
            lap[2 * n + 1] = lI[n];
        }
    if (other) {
        other[0] = kinR;
        other[1] = Vint;
        other[2] = exp(ex + phiR);
        other[3] = ex;
        other[4] = sR1;
        other[5] = sI1;
        other[6] = sR2;
        other[7] = sI2;
        other[8] = sR1I1;
    }
    free(cR); free(cI); free(B); free(Ol); free(gR); free(gI); free(lR); free(lI); free(rho);
    return 0;
}

/* Batched wrapper over W configurations R[W][N][D]. */
int tdo_eval_batch(const tdo_system* s, int64_t W, const double* R, const double* uR, const double* uI,
                   double phiR, double time, double* expo, double* O, double* ER, double* EI,
                   double* grad, double* lap, double* other)
{
    const int N = s->N, D = s->D, P = tdo_num_params(s);
    for (int64_t w = 0; w < W; w++) {
        int rc = tdo_eval(s, R + w * N * D, uR, uI, phiR, time, expo + w, O ? O + w * P : NULL,
                          ER ? ER + w : NULL, EI ? EI + w : NULL,
                          grad ? grad + w * N * D * 2 : NULL, lap ? lap + w * N * 2 : NULL,
                          other ? other + w * 9 : NULL);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* log|psi| change of a single-particle move (for the CPU sampler)     */
/* ------------------------------------------------------------------ */

/* Real part of ln psi for configuration R (pair + one-body + k-space),
 * with the reference's outerSum convention. */
static double tdo_log_psi_re(const tdo_system* s, const double* R, const double* uR, const double* cR)
{
    const int N = s->N, D = s->D, Ps = s->P_s, K = s->K;
    const int off_pair = s->one_body ? 1 : 0, off_k = off_pair + Ps;
    const double h = s->rc / (Ps - 1);
    double d[3], b[4], sum = 0.0;
    for (int n = 0; n < N; n++)
        for (int i = 0; i < n; i++) {
            double r = tdo_displacement(R + n * D, R + i * D, D, s->L, d);
            if (r <= s->rc) {
                int j = s->grid ? tdo_bspline_nu(r, s->grid, Ps, b, NULL, NULL) : tdo_bspline(r, h, Ps, b, NULL, NULL);
                for (int m = 0; m < 4; m++) sum += cR[j + m] * b[m];
            } else {
                sum += uR[off_pair + Ps - 1];
            }
        }
    if (s->one_body) {
        double sr2 = 0;
        for (int n = 0; n < N * D; n++) sr2 += R[n] * R[n];
        sum += -uR[0] * sr2;
    }
    for (int k = 0; k < K; k++) {
        const double* kv = s->kvec + k * D;
        double re = 0, im = 0;
        for (int n = 0; n < N; n++) {
            double ph = 0;
            for (int a = 0; a < D; a++) ph += kv[a] * R[n * D + a];
            re += cos(ph);
            im += sin(ph);
        }
        sum += uR[off_k + k] * (re * re + im * im);
    }
    return sum;
}

/* splitmix64 / xoshiro256** -- the oracle's own generator, independent of the
 * library's Philox stream (sampled trajectories are compared statistically). */
typedef struct { uint64_t s[4]; } tdo_rng;
static uint64_t tdo_splitmix(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t tdo_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t tdo_next(tdo_rng* r)
{
    uint64_t* s = r->s;
    uint64_t res = tdo_rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = tdo_rotl(s[3], 45);
    return res;
}
static double tdo_u01(tdo_rng* r) { return (tdo_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double tdo_normal(tdo_rng* r)
{
    double u1 = tdo_u01(r), u2 = tdo_u01(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Metropolis walk following src/TDVMC.cpp:858-916 (DoMetropolisStep): pick a
 * random particle, displace every component by N(0, step), accept with
 * probability min(1, exp(2 (lnpsi_new - lnpsi_old))), restore on reject.
 * Runs n_samples * (n_therm moves, then one record of the configuration) for
 * one walker; records[n_samples][N][D].  Returns accepted moves.  Recomputes
 * ln psi from scratch per move (slow; for small test systems only). */
int64_t tdo_metropolis(const tdo_system* s, double* R, const double* uR, double step, uint64_t seed,
                       int64_t n_samples, int64_t n_therm, double* records)
{
    const int N = s->N, D = s->D, Ps = s->P_s;
    const int off_pair = s->one_body ? 1 : 0;
    double* cR = calloc(Ps + 2, sizeof(double));
    tdo_spline_coeffs(s, uR + off_pair, cR);
    tdo_rng rng;
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) rng.s[i] = tdo_splitmix(&x);
    double cur = tdo_log_psi_re(s, R, uR, cR);
    int64_t acc = 0;
    double old[3];
    for (int64_t smp = 0; smp < n_samples; smp++) {
        for (int64_t t = 0; t < n_therm; t++) {
            int p = (int)(tdo_u01(&rng) * N);
            if (p >= N) p = N - 1;
            for (int a = 0; a < D; a++) {
                old[a] = R[p * D + a];
                R[p * D + a] += step * tdo_normal(&rng);
            }
            double nw = tdo_log_psi_re(s, R, uR, cR);
            double q = exp(2.0 * (nw - cur));
            double u = tdo_u01(&rng);
            if (isfinite(q) && q >= u) {
                cur = nw;
                acc++;
            } else {
                for (int a = 0; a < D; a++) R[p * D + a] = old[a];
            }
        }
        if (records) memcpy(records + smp * N * D, R, sizeof(double) * N * D);
    }
    free(cR);
    return acc;
}

/* ------------------------------------------------------------------ */
/* TDVP moments, linear system, Cholesky, ODE helpers                  */
/* ------------------------------------------------------------------ */

/* Sample means of src/TDVMC.cpp:1103-1109 in centred form:
 *   mO[P] = <O>, mE[2] = <E_R>, <E_I>,
 *   S[P][P] = <O O^T> - <O><O>^T, FR[P] = <O E_R> - <O><E_R>, FI likewise.
 * Two-pass (mean first) in FP64. */
void tdo_moments(int64_t W, int P, const double* O, const double* ER, const double* EI,
                 double* mO, double* mE, double* S, double* FR, double* FI)
{
    for (int k = 0; k < P; k++) mO[k] = 0.0;
    double eR = 0, eI = 0;
    for (int64_t w = 0; w < W; w++) {
        for (int k = 0; k < P; k++) mO[k] += O[w * P + k];
        eR += ER[w];
        eI += EI[w];
    }
    for (int k = 0; k < P; k++) mO[k] /= (double)W;
    eR /= (double)W;
    eI /= (double)W;
    mE[0] = eR;
    mE[1] = eI;
    memset(S, 0, sizeof(double) * P * P);
    memset(FR, 0, sizeof(double) * P);
    memset(FI, 0, sizeof(double) * P);
    double* c = malloc(sizeof(double) * P);
    for (int64_t w = 0; w < W; w++) {
        for (int k = 0; k < P; k++) c[k] = O[w * P + k] - mO[k];
        double dR = ER[w] - eR, dI = EI[w] - eI;
        for (int i = 0; i < P; i++) {
            for (int j = 0; j <= i; j++) S[i * P + j] += c[i] * c[j];
            FR[i] += c[i] * dR;
            FI[i] += c[i] * dI;
        }
    }
    for (int i = 0; i < P; i++) {
        for (int j = 0; j <= i; j++) {
            S[i * P + j] /= (double)W;
            S[j * P + i] = S[i * P + j];
        }
        FR[i] /= (double)W;
        FI[i] /= (double)W;
    }
    free(c);
}

/* Right-hand sides of src/TDVMC.cpp:1506-1537 from centred moments:
 *   real time (imag_time == 0): bR = F_I, bI = -F_R
 *   imaginary time:             bR = -F_R, bI = -<O E_I> (not centred)
 * mOEI_raw = <O E_I> = F_I + <O><E_I>. */
void tdo_build_rhs(int P, int imag_time, const double* mO, const double* mE, const double* FR,
                   const double* FI, double* bR, double* bI)
{
    for (int i = 0; i < P; i++) {
        if (imag_time == 0) {
            bR[i] = FI[i];
            bI[i] = -FR[i];
        } else {
            bR[i] = -FR[i];
            bI[i] = -(FI[i] + mO[i] * mE[1]);
        }
    }
}

/* In-place lower Cholesky following src/TDVMC.cpp:1560-1592: on a
 * non-positive pivot the diagonal entry keeps its (unreduced) value and the
 * step is flagged.  Returns the number of non-positive pivots. */
int tdo_cholesky_ref(int P, double* A)
{
    int bad = 0;
    for (int i = 0; i < P; i++)
        for (int j = 0; j <= i; j++) {
            double sum = A[i * P + j];
            for (int k = 0; k < j; k++) sum -= A[i * P + k] * A[j * P + k];
            if (i > j)
                A[i * P + j] = sum / A[j * P + j];
            else if (sum > 0)
                A[i * P + i] = sqrt(sum);
            else
                bad++;
        }
    return bad;
}

/* Forward/back substitution with the lower factor (src/TDVMC.cpp:1594-1622). */
void tdo_chol_solve(int P, const double* Lf, const double* b, double* x)
{
    double* t = malloc(sizeof(double) * P);
    for (int i = 0; i < P; i++) {
        double sum = 0;
        for (int j = 0; j < i; j++) sum += Lf[i * P + j] * t[j];
        t[i] = (b[i] - sum) / Lf[i * P + i];
    }
    for (int i = P - 1; i >= 0; i--) {
        double sum = 0;
        for (int j = P - 1; j > i; j--) sum += Lf[j * P + i] * x[j];
        x[i] = (t[i] - sum) / Lf[i * P + i];
    }
    free(t);
}

/* phi-dot of src/TDVMC.cpp:1658-1682 (imag_time 0 or 1). */
void tdo_phi_dot(int P, int imag_time, const double* mO, double mER, const double* xR, const double* xI,
                 double* phiDotR, double* phiDotI)
{
    double pr = 0, pi = 0;
    for (int i = 0; i < P; i++) {
        pr -= mO[i] * xR[i];
        pi -= mO[i] * xI[i];
    }
    if (imag_time == 0)
        pi -= mER;
    else
        pr -= mER;
    *phiDotR = pr;
    *phiDotI = pi;
}

/* Solver type 0 of src/TDVMC.cpp:1713-1762: optional diagonal scaling
 * (PreconditionEquationSystemByScaling :1684-1701), S += eps I
 * (RegularizeEquationSystem :1703-1711), Cholesky, two solves, phi-dot,
 * undo the scaling on alpha-dot.  S is overwritten.  Returns pivot failures. */
int tdo_solve_param_dot(int P, int imag_time, int precondition, double eps, double* S,
                        const double* mO, const double* mE, const double* FR, const double* FI,
                        double* xR, double* xI, double* phiDotR, double* phiDotI)
{
    double* bR = malloc(sizeof(double) * P);
    double* bI = malloc(sizeof(double) * P);
    double* sc = malloc(sizeof(double) * P);
    tdo_build_rhs(P, imag_time, mO, mE, FR, FI, bR, bI);
    if (precondition) {
        for (int i = 0; i < P; i++) sc[i] = sqrt(S[i * P + i]);
        for (int i = 0; i < P; i++) {
            for (int j = 0; j < P; j++) S[i * P + j] /= (sc[i] * sc[j]);
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
    }
    for (int i = 0; i < P; i++) S[i * P + i] += eps;
    int bad = tdo_cholesky_ref(P, S);
    tdo_chol_solve(P, S, bR, xR);
    tdo_chol_solve(P, S, bI, xI);
    tdo_phi_dot(P, imag_time, mO, mE[0], xR, xI, phiDotR, phiDotI);
    if (precondition)
        for (int i = 0; i < P; i++) {
            xR[i] /= sc[i];
            xI[i] /= sc[i];
        }
    free(bR); free(bI); free(sc);
    return bad;
}

/* Householder QR with full (row and column) pivoting, the decomposition
 * Eigen::FullPivHouseholderQR (Eigen 3.3, vendored under resources/) applies
 * in solver type 1 (src/TDVMC.cpp:1786-1796); algorithm of Golub & Van Loan,
 * "Matrix Computations", 4th ed., sec. 5.4.2, with Eigen's conventions:
 *  - pivot = entry of largest magnitude of the trailing block, the first one
 *    in column-major order on ties;
 *  - the factorisation stops when that entry is <= n * eps * (first pivot);
 *  - the row swap exchanges the trailing part of the two rows, the column
 *    swap whole columns;
 *  - reflector H = I - tau v v^T, v = (1, essential), beta = -sign(c0) |x|
 *    (c0 >= 0 -> negative), essential = tail / (c0 - beta),
 *    tau = (beta - c0) / beta; a zero tail gives tau = 0;
 *  - rank = number of |R_kk| above n * eps * max_k |R_kk|; the solve applies
 *    the first rank row swaps and reflectors to b, back-substitutes on the
 *    leading rank x rank block of R, sets the rest to zero and undoes the
 *    column swaps.
 * A is n x n row-major (unchanged).  Returns the rank. */
int tdo_fpqr_solve(int n, const double* A_in, const double* b_in, double* x)
{
    const double eps = 2.220446049250313e-16, prec = n * eps;
    double* A = malloc(sizeof(double) * n * n);
    double* c = malloc(sizeof(double) * n);
    double* tau = malloc(sizeof(double) * n);
    int* rt = malloc(sizeof(int) * n);
    int* ct = malloc(sizeof(int) * n);
    memcpy(A, A_in, sizeof(double) * n * n);
    double biggest = 0.0, maxpivot = 0.0;
    int nz = n;
    for (int k = 0; k < n; k++) {
        int bi = k, bj = k;
        double big = -1.0;
        for (int j = k; j < n; j++)          /* column-major scan, strict > */
            for (int i = k; i < n; i++)
                if (fabs(A[i * n + j]) > big) {
                    big = fabs(A[i * n + j]);
                    bi = i;
                    bj = j;
                }
        if (k == 0) biggest = big;
        if (big <= biggest * prec) {
            nz = k;
            for (int i = k; i < n; i++) {
                rt[i] = i;
                ct[i] = i;
                tau[i] = 0.0;
            }
            break;
        }
        rt[k] = bi;
        ct[k] = bj;
        if (bi != k)
            for (int j = k; j < n; j++) {
                double t = A[k * n + j]; A[k * n + j] = A[bi * n + j]; A[bi * n + j] = t;
            }
        if (bj != k)
            for (int i = 0; i < n; i++) {
                double t = A[i * n + k]; A[i * n + k] = A[i * n + bj]; A[i * n + bj] = t;
            }
        double tail = 0.0;
        for (int i = k + 1; i < n; i++) tail += A[i * n + k] * A[i * n + k];
        const double c0 = A[k * n + k];
        double beta, tk;
        if (tail <= 2.2250738585072014e-308) {
            tk = 0.0;
            beta = c0;
            for (int i = k + 1; i < n; i++) A[i * n + k] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
            for (int i = k + 1; i < n; i++) A[i * n + k] /= (c0 - beta);
            tk = (beta - c0) / beta;
        }
        A[k * n + k] = beta;
        tau[k] = tk;
        if (fabs(beta) > maxpivot) maxpivot = fabs(beta);
        /* apply H to the trailing columns: w = row k + essential^T bottom */
        for (int j = k + 1; j < n; j++) {
            double w = A[k * n + j];
            for (int i = k + 1; i < n; i++) w += A[i * n + k] * A[i * n + j];
            A[k * n + j] -= tk * w;
            for (int i = k + 1; i < n; i++) A[i * n + j] -= (tk * A[i * n + k]) * w;
        }
    }
    int rank = 0;
    for (int i = 0; i < nz; i++) rank += fabs(A[i * n + i]) > maxpivot * prec;
    memcpy(c, b_in, sizeof(double) * n);
    for (int k = 0; k < rank; k++) {
        double t = c[k]; c[k] = c[rt[k]]; c[rt[k]] = t;
        double w = c[k];
        for (int i = k + 1; i < n; i++) w += A[i * n + k] * c[i];
        c[k] -= tau[k] * w;
        for (int i = k + 1; i < n; i++) c[i] -= (tau[k] * A[i * n + k]) * w;
    }
    for (int i = rank - 1; i >= 0; i--) {
        double s = c[i];
        for (int j = i + 1; j < rank; j++) s -= A[i * n + j] * c[j];
        c[i] = s / A[i * n + i];
    }
    for (int i = rank; i < n; i++) c[i] = 0.0;
    /* x = Pi y with Pi = t_0 t_1 ... t_{n-1}: apply the transpositions last first */
    for (int k = n - 1; k >= 0; k--) {
        double t = c[k]; c[k] = c[ct[k]]; c[ct[k]] = t;
    }
    memcpy(x, c, sizeof(double) * n);
    free(A); free(c); free(tau); free(rt); free(ct);
    return rank;
}

/* Solver type 1 of src/TDVMC.cpp:1763-1822: optional diagonal scaling with
 * S += 0.002 I (only then), full-pivot QR solves, subtraction of the mean of
 * each solution vector (:1800-1809), phi-dot on that vector, then the scaling
 * is undone.  S is overwritten.  Returns P - rank. */
int tdo_solve_param_dot_qr(int P, int imag_time, int precondition, double* S, const double* mO, const double* mE,
                           const double* FR, const double* FI, double* xR, double* xI, double* phiDotR,
                           double* phiDotI)
{
    double* bR = malloc(sizeof(double) * P);
    double* bI = malloc(sizeof(double) * P);
    double* sc = malloc(sizeof(double) * P);
    tdo_build_rhs(P, imag_time, mO, mE, FR, FI, bR, bI);
    if (precondition) {
        for (int i = 0; i < P; i++) sc[i] = sqrt(S[i * P + i]);
        for (int i = 0; i < P; i++) {
            for (int j = 0; j < P; j++) S[i * P + j] /= (sc[i] * sc[j]);
            bR[i] /= sc[i];
            bI[i] /= sc[i];
        }
        for (int i = 0; i < P; i++) S[i * P + i] += 0.002;
    }
    int rank = tdo_fpqr_solve(P, S, bR, xR);
    tdo_fpqr_solve(P, S, bI, xI);
    double mr = 0.0, mi = 0.0;
    for (int i = 0; i < P; i++) {
        mr += xR[i];
        mi += xI[i];
    }
    mr /= P;
    mi /= P;
    for (int i = 0; i < P; i++) {
        xR[i] -= mr;
        xI[i] -= mi;
    }
    tdo_phi_dot(P, imag_time, mO, mE[0], xR, xI, phiDotR, phiDotI);
    if (precondition)
        for (int i = 0; i < P; i++) {
            xR[i] /= sc[i];
            xI[i] /= sc[i];
        }
    free(bR); free(bI); free(sc);
    return P - rank;
}

/* ------------------------------------------------------------------ */
/* additional observables: pair distribution g(r) and S(k)            */
/* ------------------------------------------------------------------ */

/* Pair-distribution grid of BosonsBulk.  The grid ends at the last knot,
 * nodes[P_s + 2] = ((P_s - 1) rc) / (P_s - 1) (src/PhysicalSystems/BosonsBulk.cpp:61-72
 * with halfLength = rc), spacing = gmax / bins (:145), count = (int)(gmax / spacing)
 * and lower edges accumulated by repeated addition (src/Observables/Grid.cpp:17-49),
 * centres as in Grid::InitGridCenterPoints, scaling = bin width (1-D) or the
 * area / volume between consecutive circles / spheres
 * (src/Observables/ObservableVsOnGridWithScaling.cpp:19-42).
 * Returns count; centres and scaling (may be NULL) need count entries. */
int tdo_gr_grid(int D, int P_s, double rc, int bins, double* gmax, double* spacing, double* centres,
                double* scaling)
{
    const double mx = ((double)(P_s - 1) * rc) / ((double)P_s - 1);
    const double sp = mx / bins;
    const int count = (int)((mx - 0.0) / sp);
    *gmax = mx;
    *spacing = sp;
    if (count <= 0) return count;
    if (centres) {
        double val = 0.0, prev = 0.0;
        for (int i = 0; i < count; i++) {
            if (i > 0) centres[i - 1] = (val + prev) / 2.0;
            prev = val;
            val += sp;
        }
        centres[count - 1] = (mx + prev) / 2.0;
    }
    if (scaling) {
        for (int i = 0; i < count; i++) {
            if (D == 1)
                scaling[i] = sp;
            else if (D == 3)
                scaling[i] = 4.0 * M_PI * pow(sp * (i + 1), 3.0) / 3.0;
            else
                scaling[i] = M_PI * pow(sp * (i + 1), 2.0);
        }
        if (D != 1)
            for (int i = count - 1; i > 0; i--) scaling[i] = scaling[i] - scaling[i - 1];
    }
    return count;
}

/* The first nshell shells of wave vectors, in the layout of
 * config/kVectors<D>D.json as read by src/Utils.cpp:1020-1049 and used by
 * src/PhysicalSystems/BosonsBulk.cpp:124-137 (50 shells, :55): integer n != 0
 * with non-negative components, grouped by |n|^2 in increasing order,
 * lexicographic within a shell.  vec receives up to max_vec vectors [D] each,
 * count[s] the vectors of shell s, n2[s] its |n|^2.  Returns the total number
 * of vectors (or -1 if max_vec is too small). */
int tdo_kshells(int D, int nshell, int max_vec, int* vec, int* count, int* n2)
{
    int total = 0, s = 0;
    for (int q = 1; s < nshell; q++) {
        int lim = 1;
        while (lim * lim <= q) lim++;
        int c = 0;
        const int l1 = lim, l2 = D > 1 ? lim : 1, l3 = D > 2 ? lim : 1;
        for (int x = 0; x < l1; x++)
            for (int y = 0; y < l2; y++)
                for (int z = 0; z < l3; z++) {
                    if (x * x + y * y + z * z != q) continue;
                    if (total >= max_vec) return -1;
                    const int v[3] = {x, y, z};
                    for (int a = 0; a < D; a++) vec[total * D + a] = v[a];
                    total++;
                    c++;
                }
        if (c == 0) continue;
        count[s] = c;
        n2[s] = q;
        s++;
    }
    return total;
}

/* BosonsBulk::CalculateAdditionalSystemProperties for one configuration
 * (src/PhysicalSystems/BosonsBulk.cpp:474-520): every pair i > j with
 * minimum-image distance r < gmax adds (D / (N - 1)) / scaling[bin] to bin
 * floor(r / spacing) (ObservableVsOnGridWithScaling::AddToHistogram,
 * :45-50, with Grid::GetIndex, Grid.cpp:51-60); the reference writes past
 * the grid if that bin is >= count, which is skipped here.  S(k) per shell:
 * (C^2 + S^2) / (N n_k) with C, S the sums over particles and the shell's
 * vectors of cos / sin(k . r_i), k = n 2 pi / L, summed particle-major as
 * there.  gr needs tdo_gr_grid's count entries, sk nshell.  Returns count. */
int tdo_observables(int N, int D, double L, int P_s, double rc, int bins, int nshell, const double* R, double* gr,
                    double* sk)
{
    double gmax, sp;
    const int count = tdo_gr_grid(D, P_s, rc, bins, &gmax, &sp, NULL, NULL);
    if (count <= 0) return count;
    double* scaling = (double*)malloc(sizeof(double) * count);
    tdo_gr_grid(D, P_s, rc, bins, &gmax, &sp, NULL, scaling);
    for (int b = 0; b < count; b++) gr[b] = 0.0;
    const double weight = 1.0 / ((double)(N - 1)) * D;
    double dv[3];
    for (int i = 0; i < N; i++)
        for (int j = 0; j < i; j++) {
            const double r = tdo_displacement(R + i * D, R + j * D, D, L, dv);
            if (r < gmax) {
                const int bin = (int)floor(r / sp);
                if (bin < count) gr[bin] += weight / scaling[bin];
            }
        }
    free(scaling);

    const int max_vec = 4096;
    int* vec = (int*)malloc(sizeof(int) * max_vec * D);
    int* cnt = (int*)malloc(sizeof(int) * nshell);
    int* n2 = (int*)malloc(sizeof(int) * nshell);
    tdo_kshells(D, nshell, max_vec, vec, cnt, n2);
    double* sc = (double*)calloc(nshell, sizeof(double));
    double* ss = (double*)calloc(nshell, sizeof(double));
    const double kf = 2 * M_PI / L;
    for (int i = 0; i < N; i++) {
        int v0 = 0;
        for (int k = 0; k < nshell; k++) {
            for (int kn = 0; kn < cnt[k]; kn++) {
                const int* n = vec + (v0 + kn) * D;
                double ph = (n[0] * kf) * R[i * D];
                if (D > 1) ph = ph + (n[1] * kf) * R[i * D + 1];
                if (D > 2) ph = ph + (n[2] * kf) * R[i * D + 2];
                sc[k] += cos(ph);
                ss[k] += sin(ph);
            }
            v0 += cnt[k];
        }
    }
    for (int k = 0; k < nshell; k++) sk[k] = (sc[k] * sc[k] + ss[k] * ss[k]) / ((double)(N * cnt[k]));
    free(vec);
    free(cnt);
    free(n2);
    free(sc);
    free(ss);
    return count;
}

/* Basis sums B_s = sum over pairs r_ij <= rc of B_s(r_ij), s = 0 .. P_s + 1,
 * before the boundary-condition map (BosonsBulk::CalculateLocalOperators,
 * src/PhysicalSystems/BosonsBulk.cpp:179-218: splineSums), for one
 * configuration R [N][D].  B [P_s + 2].  Returns the number of pairs inside. */
int tdo_basis_sums(const tdo_system* s, const double* R, double* B)
{
    const int N = s->N, D = s->D, Ps = s->P_s;
    const double h = s->rc / (Ps - 1);
    double b[4], d[3];
    int inside = 0;
    for (int k = 0; k < Ps + 2; k++) B[k] = 0.0;
    for (int n = 0; n < N; n++)
        for (int i = 0; i < n; i++) {
            const double r = tdo_displacement(R + n * D, R + i * D, D, s->L, d);
            if (r > s->rc) continue;
            const int j = s->grid ? tdo_bspline_nu(r, s->grid, Ps, b, NULL, NULL) : tdo_bspline(r, h, Ps, b, NULL, NULL);
            for (int m = 0; m < 4; m++)
                if (j + m < Ps + 2) B[j + m] += b[m];
            inside++;
        }
    return inside;
}
