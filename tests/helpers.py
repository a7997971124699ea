"""Shared helpers of the GPU parity tests: the same system described once as a
config dictionary (for the library) and as an oracle.System (for the checker)."""
import numpy as np

from oracle import oracle as o

POT = {"SQUARE_WELL": o.POT_SQUARE, "GAUSSIAN": o.POT_GAUSS, "LJ": o.POT_LJ, "NONE": o.POT_NONE}


def make_cfg(N, D, L, P_s, pot="SQUARE_WELL", sysparams=(0.7, 2.0), K=0, one_body=0, hbar2_2m=1.0,
             trap_omega=0.0, rc=0.0, W=1, step=0.4, therm=0, init=0, precision="FP64", **extra):
    cfg = {
        "CONFIG_VERSION": "0.24", "SYSTEM_TYPE": "BosonsBulk", "N": N, "DIM": D, "LBOX": L, "N_PARAM": P_s,
        "SYSTEM_PARAMS": list(sysparams), "PAIR_POTENTIAL": pot, "K_VECTORS": K, "ONE_BODY": one_body,
        "HBAR2_2M": hbar2_2m, "TRAP_OMEGA": trap_omega, "SPLINE_CUTOFF": rc, "N_WALKERS": W, "MC_STEP": step,
        "MC_NTHERMSTEPS": therm, "MC_NINITIALIZATIONSTEPS": init, "PRECISION": precision,
    }
    cfg.update(extra)
    return cfg


def oracle_system(cfg):
    L = float(cfg.get("LBOX", 0.0))
    D = int(cfg["DIM"])
    K = int(cfg.get("K_VECTORS", 0))
    rc = float(cfg.get("SPLINE_CUTOFF", 0.0)) or L / 2.0
    grid = np.array(cfg["NURBS_GRID"], dtype=np.float64) if cfg.get("USE_NURBS") else None
    kv = o.kvectors(D, K, L) if K else np.zeros((0, D))
    sp = tuple(cfg.get("SYSTEM_PARAMS", ()))
    return o.System(N=int(cfg["N"]), D=D, L=L, P_s=int(cfg["N_PARAM"]), rc=rc, kvec=kv,
                    one_body=int(cfg.get("ONE_BODY", 0)), hbar2_2m=float(cfg.get("HBAR2_2M", 1.0)) or 1.0,
                    pot_type=POT[cfg.get("PAIR_POTENTIAL", "SQUARE_WELL")], pot=sp,
                    pot_cutoff=(rc if grid is None else float(grid[-1])) if L > 0 else 1e300,
                    trap_omega=float(cfg.get("TRAP_OMEGA", 0.0)), grid=grid,
                    bc_origin=int(cfg.get("SPLINE_BC_ORIGIN", 0)) or 2,
                    bc_cutoff=int(cfg.get("SPLINE_BC_CUTOFF", 0)) or 2)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def energy_scale(out):
    """Scale of E_L's cancelling parts: |kinetic sums| + |V| per walker."""
    oth = out["other"]
    return np.abs(oth[:, 4]) + np.abs(oth[:, 5]) + np.abs(oth[:, 6]) + np.abs(oth[:, 7]) + np.abs(oth[:, 8]) + \
        np.abs(oth[:, 1]) + 1e-300


def adjust_pair_block(uR, uI, off, n):
    """AdjustParameters (src/TDVMC.cpp:2531-2567) on the pair-spline block
    [off, off + n): every entry minus the block's last entry (taken before the
    loop; the last one becomes 0), other entries unchanged."""
    aR, aI = np.array(uR, dtype=np.float64), np.array(uI, dtype=np.float64)
    lr, li = aR[off + n - 1], aI[off + n - 1]
    aR[off:off + n] -= lr
    aI[off:off + n] -= li
    return aR, aI


def ref_golden_tol(P):
    """Tolerance against the compiled reference's evaluation goldens.  The
    reference evaluates each cubic piece in the monomial basis of the absolute
    distance (src/SplineFactory.cpp:54-101), whose coefficients grow like
    (knot index)^3 and cancel: its relative round-off is ~1e-11 up to P = 16,
    6e-9 per weight at P = 40 (checked against a 40-digit evaluation in
    tests/test_oracle_math.py) and grows like P^3 beyond (measured 6.8e-7 on O
    at P = 100)."""
    if P <= 16:
        return 1e-10
    return 2e-7 * max(1.0, (P / 40.0) ** 3)
