"""ctypes binding of the CPU oracle (oracle/z2_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke() entry point and by
bench.py's cpu_baseline leg, always as the checker of the GPU library -- never
as a fallback for it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libz2oracle.so")
_lib = None


class Params(C.Structure):
    """paramC_sets.txt fields (reference src/calc_basic.f90:53-61)."""

    _fields_ = [
        ("RT", C.c_double), ("J", C.c_double), ("h", C.c_double), ("mu", C.c_double),
        ("Q", C.c_double),
        ("Nlx", C.c_int32), ("Nly", C.c_int32), ("Ltrot", C.c_int32), ("Beta", C.c_double),
        ("NlxTherm", C.c_int32), ("NlyTherm", C.c_int32), ("LtrotTherm", C.c_int32),
        ("Nwrap", C.c_int32), ("Nbin", C.c_int32), ("Nsweep", C.c_int32),
        ("is_tau", C.c_int32), ("Nthermal", C.c_int32),
        ("is_warm", C.c_int32), ("Nwarm", C.c_int32),
        ("is_global", C.c_int32), ("Nglobal", C.c_int32),
        ("absolute", C.c_int32),
    ]


class ObsEqual(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("flux_avg", "chi_flux", "pair_sc_q0", "cdw_pi", "cdw_shift", "cdw_ratio",
                 "density")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def make_params(**kw) -> Params:
    """Params with the reference defaults; therm sizes default to the run sizes."""
    d = dict(RT=1.0, J=1.0, h=0.5, mu=0.0, Q=1.0, Nlx=4, Nly=4, Ltrot=40, Beta=4.0,
             Nwrap=10, Nbin=1, Nsweep=1, is_tau=0, Nthermal=0, is_warm=0, Nwarm=0,
             is_global=0, Nglobal=0, absolute=1)
    d.update(kw)
    d.setdefault("NlxTherm", d["Nlx"])
    d.setdefault("NlyTherm", d["Nly"])
    d.setdefault("LtrotTherm", d["Ltrot"])
    p = Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        D, I32, I64, V = C.c_double, C.c_int32, C.c_int64, C.c_void_p
        PD, PI = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        sig = {
            "z2o_ranf": (D, [PI]),
            "z2o_nranf": (I32, [PI, I32]),
            "z2o_model_new": (V, [C.POINTER(Params)]),
            "z2o_model_free": (None, [V]),
            "z2o_model_ndim": (I32, [V]),
            "z2o_lattice_tables": (None, [V, PI, PI, PI, PI, PI, PI, PI]),
            "z2o_kin_mmult_R": (None, [V, PD, PD, C.c_int, C.c_int]),
            "z2o_kin_mmult_L": (None, [V, PD, PD, C.c_int, C.c_int]),
            "z2o_kin_delta": (None, [V, D, D, PD]),
            "z2o_bosonratio": (D, [V, PD, PD, C.c_int, C.c_int]),
            "z2o_mu_factors": (None, [V, PD, PD]),
            "z2o_flip_update": (D, [V, PD, C.c_int, D]),
            "z2o_qdrp": (None, [C.c_int, PD, PD, PI, PD]),
            "z2o_stab_green": (None, [C.c_int, C.c_int, C.c_int, PD, PD, PD, PD, PD, PD, PD]),
            "z2o_chain_new": (V, [V, I32]),
            "z2o_chain_new_conf": (V, [V, I32, PD, PD]),
            "z2o_chain_free": (None, [V]),
            "z2o_therm": (None, [V]),
            "z2o_pre": (None, [V]),
            "z2o_sweep": (None, [V, C.c_int]),
            "z2o_global_sweep": (None, [V]),
            "z2o_sweep_L_partial": (None, [V, C.c_int]),
            "z2o_pass_L": (None, [V]),
            "z2o_pass_R": (None, [V, C.c_int]),
            "z2o_dyn_measure": (None, [V]),
            "z2o_slice_L": (None, [V, C.c_int]),
            "z2o_slice_R": (None, [V, C.c_int]),
            "z2o_get_seed": (I32, [V]),
            "z2o_set_seed": (None, [V, I32]),
            "z2o_get_sigma": (None, [V, PD]),
            "z2o_set_sigma": (None, [V, PD]),
            "z2o_get_lambda": (None, [V, PD]),
            "z2o_get_green": (None, [V, PD]),
            "z2o_set_green": (None, [V, PD]),
            "z2o_get_prop_mat": (None, [V, C.c_int, PD]),
            "z2o_get_prop_d": (None, [V, C.c_int, PD]),
            "z2o_get_is_beta": (C.c_int, [V]),
            "z2o_set_is_beta": (None, [V, C.c_int]),
            "z2o_get_drift": (None, [V, PD, PD]),
            "z2o_get_counters": (None, [V, PD]),
            "z2o_get_acc": (None, [V, PD]),
            "z2o_get_obs_equal": (None, [V, C.POINTER(ObsEqual)]),
            "z2o_get_obs_tau": (None, [V, PD]),
            "z2o_obs_equal_terms": (None, [V, C.c_int, PD]),
            "z2o_get_dyn_drift": (None, [V, PD, PD]),
            "z2o_get_min_gap": (D, [V]),
            "z2o_reset_min_gap": (None, [V]),
            "z2o_get_gmin_gap": (D, [V]),
            "z2o_set_decision_log": (None, [V, C.POINTER(C.c_uint32), I64]),
            "z2o_get_decision_count": (I64, [V]),
            "z2o_run_program": (None, [V, C.POINTER(ObsEqual)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _pd(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def ranf_stream(seed: int, n: int):
    """n successive ranf draws and the final seed."""
    s = C.c_int32(seed)
    out = np.empty(n)
    L = lib()
    for i in range(n):
        out[i] = L.z2o_ranf(C.byref(s))
    return out, s.value


class Model:
    def __init__(self, params: Params):
        self.params = params
        self.ptr = lib().z2o_model_new(C.byref(params))
        if not self.ptr:
            raise ValueError("invalid parameters (Ltrot must be a multiple of Nwrap)")
        self.N = params.Nlx * params.Nly
        self.T = params.Ltrot

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().z2o_model_free(self.ptr)
            self.ptr = None

    def tables(self):
        N = self.N
        bl = np.zeros(4 * N, np.int32)
        glen = np.zeros(4, np.int32)
        groups = np.zeros(2 * N, np.int32)
        bb = np.zeros(14 * N, np.int32)
        imj = np.zeros(N * N, np.int32)
        lb = np.zeros(5 * N, np.int32)
        nl = np.zeros(2 * N, np.int32)
        lib().z2o_lattice_tables(self.ptr, _pi(bl), _pi(groups), _pi(glen), _pi(bb), _pi(imj),
                                 _pi(lb), _pi(nl))
        g, off = [], 0
        for k in range(4):
            g.append(groups[off:off + glen[k]].copy())
            off += glen[k]
        return dict(bond_list=bl.reshape(2 * N, 2), groups=g, bond_bonds=bb.reshape(2 * N, 7),
                    imj=imj.reshape(N, N, order="F"), L_bonds=lb.reshape(N, 5),
                    n_list=nl.reshape(N, 2))

    def kin_R(self, M: np.ndarray, sigma: np.ndarray, nt: int, nflag: int) -> np.ndarray:
        A = np.asfortranarray(M, dtype=np.float64).copy(order="F")
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        lib().z2o_kin_mmult_R(self.ptr, _pd(A), _pd(s), nt, nflag)
        return A

    def kin_L(self, M: np.ndarray, sigma: np.ndarray, nt: int, nflag: int) -> np.ndarray:
        A = np.asfortranarray(M, dtype=np.float64).copy(order="F")
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        lib().z2o_kin_mmult_L(self.ptr, _pd(A), _pd(s), nt, nflag)
        return A

    def delta(self, s_old: float, s_new: float) -> np.ndarray:
        D = np.zeros(4)
        lib().z2o_kin_delta(self.ptr, s_old, s_new, _pd(D))
        return D.reshape(2, 2, order="F")

    def bosonratio(self, sigma: np.ndarray, sigma_new: np.ndarray, ii: int, nt: int) -> float:
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        sn = np.ascontiguousarray(sigma_new, dtype=np.float64)
        return lib().z2o_bosonratio(self.ptr, _pd(s), _pd(sn), ii, nt)

    def flip_update(self, G: np.ndarray, ii: int, s_old: float):
        A = np.asfortranarray(G, dtype=np.float64).copy(order="F")
        det = lib().z2o_flip_update(self.ptr, _pd(A), ii, s_old)
        return A, det

    def mu_factors(self):
        a, b = C.c_double(), C.c_double()
        lib().z2o_mu_factors(self.ptr, C.byref(a), C.byref(b))
        return a.value, b.value


def qdrp(A: np.ndarray):
    n = A.shape[0]
    M = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    D = np.zeros(n)
    ipvt = np.zeros(n, np.int32)
    tau = np.zeros(n)
    lib().z2o_qdrp(n, _pd(M), _pd(D), _pi(ipvt), _pd(tau))
    return M, D, ipvt, tau


def stab_green(Ltrot, nt, UUR, DUR, VUR, UUL, DUL, VUL):
    n = UUR.shape[0]
    f = lambda X: np.asfortranarray(X, dtype=np.float64)
    G = np.zeros((n, n), order="F")
    args = [f(UUR), np.ascontiguousarray(DUR, dtype=np.float64), f(VUR), f(UUL),
            np.ascontiguousarray(DUL, dtype=np.float64), f(VUL)]
    lib().z2o_stab_green(n, Ltrot, nt, *[_pd(a) for a in args], _pd(G))
    return G


class Chain:
    """One Markov chain = one MPI rank of the reference program."""

    def __init__(self, model: Model, seed: int, sigma_therm=None, lambda_therm=None):
        self.model = model
        L = lib()
        if sigma_therm is None:
            self.ptr = L.z2o_chain_new(model.ptr, seed)
        else:
            s = np.ascontiguousarray(sigma_therm, dtype=np.float64)
            if lambda_therm is None:
                self.ptr = L.z2o_chain_new_conf(model.ptr, seed, _pd(s), None)
            else:
                lt = np.ascontiguousarray(lambda_therm, dtype=np.float64)
                self.ptr = L.z2o_chain_new_conf(model.ptr, seed, _pd(s), _pd(lt))
        self.N, self.T = model.N, model.T

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().z2o_chain_free(self.ptr)
            self.ptr = None

    # drivers
    def therm(self):
        lib().z2o_therm(self.ptr)

    def pre(self):
        lib().z2o_pre(self.ptr)

    def sweep(self, toggle=False):
        lib().z2o_sweep(self.ptr, int(toggle))

    def pass_L(self):
        lib().z2o_pass_L(self.ptr)

    def pass_R(self, toggle=False):
        lib().z2o_pass_R(self.ptr, int(toggle))

    def dyn_measure(self):
        lib().z2o_dyn_measure(self.ptr)

    def sweep_L_partial(self, nslices: int):
        lib().z2o_sweep_L_partial(self.ptr, nslices)

    def global_sweep(self):
        lib().z2o_global_sweep(self.ptr)

    def slice_L(self, nt):
        lib().z2o_slice_L(self.ptr, nt)

    def slice_R(self, nt):
        lib().z2o_slice_R(self.ptr, nt)

    def run_program(self):
        nb = self.model.params.Nbin
        arr = (ObsEqual * nb)()
        lib().z2o_run_program(self.ptr, arr)
        return [a.as_dict() for a in arr]

    # state
    @property
    def seed(self):
        return lib().z2o_get_seed(self.ptr)

    @seed.setter
    def seed(self, v):
        lib().z2o_set_seed(self.ptr, v)

    @property
    def is_beta(self):
        return bool(lib().z2o_get_is_beta(self.ptr))

    @is_beta.setter
    def is_beta(self, v):
        lib().z2o_set_is_beta(self.ptr, int(v))

    def sigma(self) -> np.ndarray:
        """sigma[nt-1, ii-1] (row = slice)."""
        s = np.zeros(2 * self.N * self.T)
        lib().z2o_get_sigma(self.ptr, _pd(s))
        return s.reshape(self.T, 2 * self.N)

    def set_sigma(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64).reshape(-1)
        lib().z2o_set_sigma(self.ptr, _pd(a))

    def lam(self):
        s = np.zeros(self.N)
        lib().z2o_get_lambda(self.ptr, _pd(s))
        return s

    def green(self) -> np.ndarray:
        G = np.zeros((self.N, self.N), order="F")
        lib().z2o_get_green(self.ptr, _pd(G))
        return G

    def set_green(self, G):
        A = np.asfortranarray(G, dtype=np.float64)
        lib().z2o_set_green(self.ptr, _pd(A))

    def prop_mat(self, which: str) -> np.ndarray:
        idx = {"UUR": 0, "VUR": 1, "UUL": 2, "VUL": 3}[which]
        G = np.zeros((self.N, self.N), order="F")
        lib().z2o_get_prop_mat(self.ptr, idx, _pd(G))
        return G

    def prop_d(self, which: str) -> np.ndarray:
        d = np.zeros(self.N)
        lib().z2o_get_prop_d(self.ptr, {"DUR": 0, "DUL": 1}[which], _pd(d))
        return d

    def drift(self):
        a, b = C.c_double(), C.c_double()
        lib().z2o_get_drift(self.ptr, C.byref(a), C.byref(b))
        return a.value, b.value

    def counters(self):
        o = np.zeros(7)
        lib().z2o_get_counters(self.ptr, _pd(o))
        keys = ("local_acc", "local_prop", "glob_acc", "glob_prop", "therm_acc", "therm_prop",
                "cluster_sum")
        return dict(zip(keys, o))

    def acc(self):
        o = np.zeros(4)
        lib().z2o_get_acc(self.ptr, _pd(o))
        return dict(zip(("Kl", "Kg", "Kt", "size_cluster"), o))

    def obs_equal(self):
        o = ObsEqual()
        lib().z2o_get_obs_equal(self.ptr, C.byref(o))
        return o.as_dict()

    def obs_equal_terms(self, nt):
        """the sums Obs_equal_calc adds for slice nt on the current G:
        flux_tau, pair_sum, cdw_sum, cdw_shift_sum, density_slice"""
        o = np.zeros(5)
        lib().z2o_obs_equal_terms(self.ptr, nt, _pd(o))
        return dict(zip(("flux_tau", "pair_sum", "cdw_sum", "cdw_shift_sum", "density_slice"), o))

    def obs_tau(self):
        o = np.zeros(7 * self.N * self.T)
        lib().z2o_get_obs_tau(self.ptr, _pd(o))
        names = ("spin", "pair_s", "pair_d", "charge_s", "charge_d", "single", "curxx")
        o = o.reshape(7, self.T, self.N)
        return {n: o[k].T.copy() for k, n in enumerate(names)}

    def dyn_drift(self):
        a, b = np.zeros(3), np.zeros(3)
        lib().z2o_get_dyn_drift(self.ptr, _pd(a), _pd(b))
        return a, b

    @property
    def min_gap(self):
        return lib().z2o_get_min_gap(self.ptr)

    def reset_min_gap(self):
        lib().z2o_reset_min_gap(self.ptr)

    @property
    def global_min_gap(self):
        """smallest |min(log ratio, 0) - log(r)| over the global accept/reject decisions"""
        return lib().z2o_get_gmin_gap(self.ptr)

    def set_decision_log(self, cap: int):
        self._log = np.zeros(cap, np.uint32)
        lib().z2o_set_decision_log(self.ptr, self._log.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   cap)
        return self._log

    def decision_count(self):
        return lib().z2o_get_decision_count(self.ptr)
