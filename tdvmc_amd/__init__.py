"""tdvmc_amd -- Python binding of the MI355X (gfx950) TDVMC hot-path library.

The compute path is ``libtdvmc.so`` (hand-written HIP kernels behind the C ABI
declared in ``include/tdvmc.h``); this module only marshals numpy arrays
through ctypes.  There is no CPU fallback: if the library is missing, or no
GPU is present, every call raises.

    from tdvmc_amd import Context
    ctx = Context(config_dict_or_json_or_path)
    out = ctx.eval_configs(R)                  # ln psi, O_k, E_L, ...
    ctx.sweep(n_moves, seed)
    step = ctx.sample(seed)                    # SimulationStepData-shaped dict
    ctx.tdvp_step(dt)
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdvmc.so")

FP64 = 0
MIXED = 1
DEFAULT = -1

OK, ERR_BAD_ARG, ERR_BAD_CONFIG, ERR_HIP, ERR_NOT_PD, ERR_NONFINITE, ERR_STATE, ERR_REJECTED = range(8)

_dp = C.POINTER(C.c_double)
_lib = None


class TdvmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tdvmc error {code}: {msg}")
        self.code = code


class Info(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("DIM", C.c_int), ("P", C.c_int), ("P_pair", C.c_int), ("K", C.c_int),
        ("one_body", C.c_int), ("n_walkers", C.c_int64), ("L", C.c_double), ("rc", C.c_double),
        ("hbar2_2m", C.c_double), ("imaginary_time", C.c_int), ("ode_solver_type", C.c_int),
        ("precision", C.c_int), ("samples_per_walker", C.c_int), ("timestep", C.c_double),
        ("mc_step", C.c_double), ("therm_moves", C.c_int64), ("init_moves", C.c_int64), ("time", C.c_double),
        ("reg_eps", C.c_double), ("precondition", C.c_int), ("K_terms", C.c_int), ("gr_bin_count", C.c_int),
        ("use_nurbs", C.c_int), ("bc_origin", C.c_int), ("bc_cutoff", C.c_int),
    ]


class StepData(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("step", C.c_int), ("n_samples", C.c_int64),
        ("localEnergyR", C.c_double), ("localEnergyI", C.c_double),
        ("localOperators", _dp), ("localOperatorsMatrix", _dp),
        ("localOperatorlocalEnergyR", _dp), ("localOperatorlocalEnergyI", _dp),
        ("S", _dp), ("FR", _dp), ("FI", _dp), ("otherExpectationValues", _dp),
    ]


SYMBOLS = {
    "tdvmc_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "tdvmc_destroy": (C.c_int, [C.c_void_p]),
    "tdvmc_parse_config": (C.c_int, [C.c_char_p, C.POINTER(Info)]),
    "tdvmc_last_error": (C.c_char_p, [C.c_void_p]),
    "tdvmc_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "tdvmc_set_walker_count": (C.c_int, [C.c_void_p, C.c_int64]),
    "tdvmc_set_params": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int]),
    "tdvmc_get_params": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int]),
    "tdvmc_set_time": (C.c_int, [C.c_void_p, C.c_double]),
    "tdvmc_set_walkers": (C.c_int, [C.c_void_p, _dp, C.c_int64]),
    "tdvmc_get_walkers": (C.c_int, [C.c_void_p, _dp, C.c_int64]),
    "tdvmc_init_walkers": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tdvmc_sweep": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_int]),
    "tdvmc_get_acceptance": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "tdvmc_sample": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(StepData)]),
    "tdvmc_measure": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(StepData)]),
    "tdvmc_eval_configs": (C.c_int, [C.c_void_p, _dp, C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "tdvmc_eval_configs_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tdvmc_sample_dev": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "tdvmc_measure_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tdvmc_set_walkers_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "tdvmc_get_walkers_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "tdvmc_observables_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp]),
    "tdvmc_observables_configs": (C.c_int, [C.c_void_p, _dp, C.c_int64, C.c_int, _dp, _dp]),
    "tdvmc_additional_properties": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_int, _dp, _dp]),
    "tdvmc_files_init": (C.c_int, [C.c_char_p, _dp, _dp, C.c_int]),
    "tdvmc_files_append_step": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(StepData), _dp, _dp, C.c_int]),
    "tdvmc_files_write_grids": (C.c_int, [C.c_char_p, _dp, C.c_int, _dp, C.c_int]),
    "tdvmc_files_append_observables": (C.c_int, [C.c_char_p, C.c_double, _dp, C.c_int, _dp, C.c_int]),
    "tdvmc_write_final_config": (C.c_int, [C.c_char_p, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_char_p]),
    "tdvmc_files_write_coordinates": (C.c_int, [C.c_char_p, _dp, C.c_int64, C.c_int, C.c_int, C.c_double]),
    "tdvmc_files_read_coordinates": (C.c_int, [C.c_char_p, _dp, C.c_int, C.c_int]),
    "tdvmc_files_write_rng_state": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_int]),
    "tdvmc_files_read_rng_state": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "tdvmc_get_rng_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "tdvmc_set_rng_state": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32]),
    "tdvmc_moments": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int64, C.c_int, C.c_int, _dp, _dp, _dp]),
    "tdvmc_solve": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp,
                              C.POINTER(C.c_int)]),
    "tdvmc_get_step_status": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tdvmc_tdvp_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]),
    "tdvmc_adjoint": (C.c_int, [C.c_void_p, _dp, C.c_int64, _dp, C.c_int, _dp, _dp]),
    "tdvmc_basis_sums": (C.c_int, [C.c_void_p, _dp, C.c_int64, C.c_int, _dp]),
    "tdvmc_spline_bc": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int), _dp, C.POINTER(C.c_int)]),
    "tdvmc_adjoint_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_void_p]),
    "tdvmc_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "tdvmc_profile_get": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "tdvmc_set_sweep_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "tdvmc_synchronize": (C.c_int, [C.c_void_p]),
    "tdvmc_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "tdvmc_batch_create": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_void_p)]),
    "tdvmc_batch_destroy": (C.c_int, [C.c_void_p]),
    "tdvmc_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "tdvmc_batch_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tdvmc_batch_system": (C.c_void_p, [C.c_void_p, C.c_int]),
    "tdvmc_batch_set_sweep_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "tdvmc_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tdvmc_batch_sweep": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_int]),
    "tdvmc_batch_measure": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "tdvmc_batch_tdvp_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]),
}


def lib():
    """Loads libtdvmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m tdvmc_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def config_text(cfg) -> str:
    if isinstance(cfg, dict):
        return json.dumps(cfg)
    return str(cfg)


def parse_config(config) -> Info:
    """Parses and validates a config on the host (no GPU needed)."""
    info = Info()
    rc = lib().tdvmc_parse_config(config_text(config).encode(), C.byref(info))
    if rc != OK:
        raise TdvmcError(rc, lib().tdvmc_last_error(None).decode())
    return info


class Context:
    """One system: its walkers, parameters and device buffers (a tdvmc_ctx)."""

    def __init__(self, config, _handle=None, _owner=None):
        L = lib()
        if _handle is not None:
            # a member of a Batch: the batch owns (and destroys) it
            self._h = C.c_void_p(_handle)
            self._owner = _owner
        else:
            h = C.c_void_p()
            rc = L.tdvmc_create(config_text(config).encode(), C.byref(h))
            if rc != OK:
                raise TdvmcError(rc, L.tdvmc_last_error(None).decode())
            self._h = h
            self._owner = None
        self.info = self.get_info()

    # ------------------------------------------------------------ plumbing
    def _check(self, rc):
        if rc != OK:
            raise TdvmcError(rc, lib().tdvmc_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                lib().tdvmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_info(self) -> Info:
        info = Info()
        self._check(lib().tdvmc_get_info(self._h, C.byref(info)))
        return info

    def spline_bc(self):
        """Boundary-condition rows (bcFactorsStart [n0][3], bcFactorsEnd [n1][3])
        of the pair spline (tdvmc_spline_bc)."""
        a, b = np.zeros(9), np.zeros(9)
        n0, n1 = C.c_int(), C.c_int()
        self._check(lib().tdvmc_spline_bc(self._h, _ptr(a), C.byref(n0), _ptr(b), C.byref(n1)))
        return a.reshape(3, 3)[:n0.value].copy(), b.reshape(3, 3)[:n1.value].copy()

    @property
    def N(self):
        return self.info.N

    @property
    def D(self):
        return self.info.DIM

    @property
    def P(self):
        return self.info.P

    @property
    def W(self):
        return self.get_info().n_walkers

    # ------------------------------------------------------------ state
    def set_walker_count(self, W: int):
        self._check(lib().tdvmc_set_walker_count(self._h, int(W)))

    def set_params(self, uR, uI=None, phiR=0.0, phiI=0.0):
        uR = _f64(uR)
        uI = np.zeros(self.P) if uI is None else _f64(uI)
        self._check(lib().tdvmc_set_params(self._h, _ptr(uR), _ptr(uI), float(phiR), float(phiI), uR.size))

    def get_params(self):
        uR, uI = np.zeros(self.P), np.zeros(self.P)
        pR, pI = C.c_double(), C.c_double()
        self._check(lib().tdvmc_get_params(self._h, _ptr(uR), _ptr(uI), C.byref(pR), C.byref(pI), self.P))
        return uR, uI, pR.value, pI.value

    def set_time(self, t: float):
        self._check(lib().tdvmc_set_time(self._h, float(t)))

    def set_walkers(self, R):
        R = _f64(R).reshape(-1, self.N, self.D)
        self._check(lib().tdvmc_set_walkers(self._h, _ptr(R), R.shape[0]))

    def get_walkers(self):
        W = self.W
        R = np.zeros((W, self.N, self.D))
        self._check(lib().tdvmc_get_walkers(self._h, _ptr(R), W))
        return R

    def init_walkers(self, seed: int):
        self._check(lib().tdvmc_init_walkers(self._h, int(seed)))

    def set_sweep_lanes(self, lanes: int):
        self._check(lib().tdvmc_set_sweep_lanes(self._h, int(lanes)))

    # ------------------------------------------------------------ compute
    def sweep(self, n_moves: int, seed: int, precision: int = DEFAULT):
        self._check(lib().tdvmc_sweep(self._h, int(n_moves), int(seed), int(precision)))

    def acceptance(self):
        a, t = C.c_int64(), C.c_int64()
        self._check(lib().tdvmc_get_acceptance(self._h, C.byref(a), C.byref(t)))
        return a.value, t.value

    def _step(self, fn, *args):
        P = self.P
        arrs = dict(localOperators=np.zeros(P), localOperatorsMatrix=np.zeros((P, P)),
                    localOperatorlocalEnergyR=np.zeros(P), localOperatorlocalEnergyI=np.zeros(P),
                    S=np.zeros((P, P)), FR=np.zeros(P), FI=np.zeros(P), otherExpectationValues=np.zeros(9))
        sd = StepData()
        for k, v in arrs.items():
            setattr(sd, k, _ptr(v))
        self._check(fn(self._h, *args, C.byref(sd)))
        arrs.update(time=sd.time, step=sd.step, n_samples=sd.n_samples, localEnergyR=sd.localEnergyR,
                    localEnergyI=sd.localEnergyI)
        return arrs

    def sample(self, seed: int, precision: int = DEFAULT):
        return self._step(lib().tdvmc_sample, int(seed), int(precision))

    def measure(self, precision: int = DEFAULT):
        return self._step(lib().tdvmc_measure, int(precision))

    def eval_configs(self, R, precision: int = DEFAULT, per_particle: bool = False):
        R = _f64(R).reshape(-1, self.N, self.D)
        W = R.shape[0]
        out = dict(expo=np.zeros(W), ER=np.zeros(W), EI=np.zeros(W), O=np.zeros((W, self.P)),
                   other=np.zeros((W, 9)))
        g = l = None
        if per_particle:
            g = np.zeros((W, self.N, self.D, 2))
            l = np.zeros((W, self.N, 2))
            out.update(grad=g, lap=l)
        self._check(lib().tdvmc_eval_configs(self._h, _ptr(R), W, int(precision), _ptr(out["expo"]),
                                             _ptr(out["ER"]), _ptr(out["EI"]), _ptr(out["O"]), _ptr(g), _ptr(l),
                                             _ptr(out["other"])))
        return out

    def observables_info(self):
        """(gr_grid bin centres, sk_grid |k| per shell) of the pair distribution and S(k)."""
        ng, ns = C.c_int(), C.c_int()
        self._check(lib().tdvmc_observables_info(self._h, C.byref(ng), C.byref(ns), None, None))
        gg, sg = np.zeros(ng.value), np.zeros(ns.value)
        self._check(lib().tdvmc_observables_info(self._h, None, None, _ptr(gg), _ptr(sg)))
        return gg, sg

    def observables_configs(self, R, precision: int = DEFAULT):
        """g(r) [W][n_gr] and S(k) [W][n_sk] of fed configurations
        (BosonsBulk::CalculateAdditionalSystemProperties)."""
        R = _f64(R).reshape(-1, self.N, self.D)
        W = R.shape[0]
        gg, sg = self.observables_info()
        gr, sk = np.zeros((W, len(gg))), np.zeros((W, len(sg)))
        self._check(lib().tdvmc_observables_configs(self._h, _ptr(R), W, int(precision), _ptr(gr), _ptr(sk)))
        return gr, sk

    def additional_properties(self, seed: int, n_samples: int = 0, precision: int = DEFAULT):
        """g(r) and S(k) averaged over n_samples observations of every walker
        (CalculateAdditionalSystemProperties + its average over ranks)."""
        gg, sg = self.observables_info()
        gr, sk = np.zeros(len(gg)), np.zeros(len(sg))
        self._check(lib().tdvmc_additional_properties(self._h, int(n_samples), int(seed), int(precision), _ptr(gr),
                                                      _ptr(sk)))
        return gr, sk

    def moments(self, O, ER, EI, precision: int = DEFAULT):
        O = _f64(O)
        W, P = O.shape
        mean, S, F = np.zeros(P + 2), np.zeros((P, P)), np.zeros((2, P))
        self._check(lib().tdvmc_moments(self._h, _ptr(O), _ptr(_f64(ER)), _ptr(_f64(EI)), W, P, int(precision),
                                        _ptr(mean), _ptr(S), _ptr(F)))
        return dict(mO=mean[:P], mE=mean[P:], S=S, FR=F[0], FI=F[1])

    def solve(self, mO, mE, S, FR, FI, imag_time=1, precondition=0, eps=1e-3, check=True, solver_type=0):
        P = len(mO)
        mean = _f64(np.concatenate([mO, mE]))
        F = _f64(np.stack([FR, FI]))
        xdot = np.zeros(2 * P + 2)
        bad = C.c_int()
        rc = lib().tdvmc_solve(self._h, _ptr(mean), _ptr(_f64(S)), _ptr(F), P, int(imag_time), int(precondition),
                               float(eps), int(solver_type), _ptr(xdot), C.byref(bad))
        if rc not in (OK, ERR_NOT_PD) or (check and rc != OK):
            self._check(rc)
        return dict(xR=xdot[:P], xI=xdot[P:2 * P], phiDotR=xdot[2 * P], phiDotI=xdot[2 * P + 1], bad=bad.value)

    def tdvp_step(self, dt: float, ode_type: int = -1, precision: int = DEFAULT, use_graph: bool = False):
        self._check(lib().tdvmc_tdvp_step(self._h, float(dt), int(ode_type), int(precision), int(use_graph)))

    def step_status(self) -> int:
        """Non-positive Cholesky pivots met by the last TDVP step."""
        bad = C.c_int()
        self._check(lib().tdvmc_get_step_status(self._h, C.byref(bad)))
        return bad.value

    def adjoint(self, R, gbar, precision: int = DEFAULT):
        R = _f64(R).reshape(-1, self.N, self.D)
        W = R.shape[0]
        dR, da = np.zeros_like(R), np.zeros(self.P)
        self._check(lib().tdvmc_adjoint(self._h, _ptr(R), W, _ptr(_f64(gbar)), int(precision), _ptr(dR), _ptr(da)))
        return dR, da

    def autotune(self, precision: int = DEFAULT) -> int:
        """Times the sweep's lanes-per-walker choices and keeps the fastest."""
        lanes = C.c_int()
        self._check(lib().tdvmc_autotune(self._h, int(precision), C.byref(lanes)))
        return lanes.value

    def profile_enable(self, on: bool = True):
        self._check(lib().tdvmc_profile_enable(self._h, int(on)))

    def profile_get(self, name: str):
        ms, n = C.c_double(), C.c_int64()
        self._check(lib().tdvmc_profile_get(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        self._check(lib().tdvmc_synchronize(self._h))

    def get_rng_state(self):
        """(seed, epoch) of the context's counter-based random stream."""
        s, e = C.c_uint64(), C.c_uint32()
        self._check(lib().tdvmc_get_rng_state(self._h, C.byref(s), C.byref(e)))
        return s.value, e.value

    def set_rng_state(self, seed: int, epoch: int):
        self._check(lib().tdvmc_set_rng_state(self._h, int(seed), int(epoch)))


def batch_system_seed(seed: int, s: int) -> int:
    """Seed of system s in a batched sweep drawn with `seed`."""
    return (seed ^ ((0x9E3779B97F4A7C15 * (s + 1)) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF


class Batch:
    """Independent systems sampled, measured and advanced together
    (tdvmc_batch): one sweep and one measurement launch for all of them."""

    def __init__(self, configs):
        L = lib()
        arr = (C.c_char_p * len(configs))(*[config_text(c).encode() for c in configs])
        h = C.c_void_p()
        rc = L.tdvmc_batch_create(arr, len(configs), C.byref(h))
        if rc != OK:
            raise TdvmcError(rc, L.tdvmc_batch_last_error(None).decode())
        self._h = h
        self.systems = [Context(None, _handle=L.tdvmc_batch_system(h, s), _owner=self) for s in range(len(configs))]

    def _check(self, rc):
        if rc != OK:
            raise TdvmcError(rc, lib().tdvmc_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            for c in self.systems:
                c._h = None
            lib().tdvmc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.systems)

    def set_sweep_lanes(self, lanes: int):
        self._check(lib().tdvmc_batch_set_sweep_lanes(self._h, int(lanes)))

    def autotune(self, precision: int = DEFAULT):
        """Times every system's measurement with 1, 2 and 4 wavefronts per
        walker and (no k-space terms) its sweep with 8 to 64 lanes per walker,
        and keeps the fastest; returns (waves, lanes) per system."""
        waves = (C.c_int * len(self.systems))()
        lanes = (C.c_int * len(self.systems))()
        self._check(lib().tdvmc_batch_autotune(self._h, int(precision), C.cast(waves, C.c_void_p),
                                               C.cast(lanes, C.c_void_p)))
        return list(waves), list(lanes)

    def sweep(self, n_moves: int = 0, seed: int = 1, precision: int = DEFAULT):
        self._check(lib().tdvmc_batch_sweep(self._h, int(n_moves), int(seed), int(precision)))

    def measure(self, precision: int = DEFAULT):
        """Per system: dict(mO, mE, S, FR, FI) of the current walkers."""
        outs, keep = (StepData * len(self.systems))(), []
        for i, c in enumerate(self.systems):
            P = c.P
            arrs = dict(mO=np.zeros(P), S=np.zeros((P, P)), FR=np.zeros(P), FI=np.zeros(P))
            keep.append(arrs)
            outs[i].localOperators = _ptr(arrs["mO"])
            outs[i].S = _ptr(arrs["S"])
            outs[i].FR = _ptr(arrs["FR"])
            outs[i].FI = _ptr(arrs["FI"])
        self._check(lib().tdvmc_batch_measure(self._h, int(precision), C.cast(outs, C.c_void_p)))
        res = []
        for i, arrs in enumerate(keep):
            arrs["mE"] = np.array([outs[i].localEnergyR, outs[i].localEnergyI])
            res.append(arrs)
        return res

    def tdvp_step(self, dt: float, ode_type: int = -1, precision: int = DEFAULT, use_graph: bool = False):
        self._check(lib().tdvmc_batch_tdvp_step(self._h, float(dt), int(ode_type), int(precision), int(use_graph)))
