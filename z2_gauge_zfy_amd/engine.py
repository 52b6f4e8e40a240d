"""Python view of one engine context (a batch of chains sharing L, Ltrot, Nwrap).

Thin plumbing over the C ABI; all work happens in the HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (Z2BatchChain, Z2ChainDesc, Z2Params, Z2_BETA_PER_CHAIN, Z2_EXEC_EAGER, Z2_PREC_FP64, check,
                   lib)

OBS_NAMES = ("flux_avg", "chi_flux", "pair_sc_q0", "cdw_pi", "cdw_shift", "cdw_ratio", "density")
TAU_NAMES = ("spin", "pair_s", "pair_d", "charge_s", "charge_d", "single", "curxx")


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class ChainSpec:
    """Per-chain couplings (reference paramC_sets.txt line 1 + Beta) and seed."""
    RT: float = 1.0
    J: float = 1.0
    h: float = 0.5
    mu: float = 0.0
    Q: float = 1.0
    Beta: float = 4.0
    seed: int = 12345
    absolute: int = 1


class Engine:
    def __init__(self, Nlx: int, Nly: int, Ltrot: int, Nwrap: int, chains, Nsweep: int = 1,
                 precision: int = Z2_PREC_FP64):
        self.params = Z2Params(Nlx, Nly, Ltrot, Nwrap, Nsweep, precision)
        self.chains = list(chains)
        arr = (Z2ChainDesc * len(self.chains))()
        for k, s in enumerate(self.chains):
            arr[k] = Z2ChainDesc(s.RT, s.J, s.h, s.mu, s.Q, s.Beta, int(s.seed), int(s.absolute))
        self._ptr = C.c_void_p()
        check(lib().z2_ctx_create(C.byref(self.params), len(self.chains), arr, C.byref(self._ptr)))
        self.N = Nlx * Nly
        self.T = Ltrot
        self.nch = len(self.chains)

    def close(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            lib().z2_ctx_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ptr(self):
        return self._ptr

    # ---- drivers
    def init_fields(self):
        check(lib().z2_init_fields(self._ptr))

    def init_fields_therm(self, NlxTherm: int, NlyTherm: int, LtrotTherm: int):
        """fresh start at the thermalisation size, tiled onto the lattice (conf_transfer)"""
        check(lib().z2_init_fields_therm(self._ptr, NlxTherm, NlyTherm, LtrotTherm))

    def set_sigma(self, chain: int, sigma: np.ndarray, seed: int):
        s = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1)
        assert s.size == 2 * self.N * self.T
        check(lib().z2_set_sigma(self._ptr, chain, _pd(s), int(seed)))

    def pre(self):
        check(lib().z2_pre(self._ptr))

    def sweep(self, is_beta=True, toggle: bool = False):
        """LocalSweep%sweep; toggle adds the time-displaced measurement (needs enable_tau).
        is_beta=None runs every chain in its own direction (after global updates)."""
        b = Z2_BETA_PER_CHAIN if is_beta is None else int(bool(is_beta))
        if toggle:
            check(lib().z2_sweep_tau(self._ptr, b, 1))
        else:
            check(lib().z2_sweep(self._ptr, b))

    def global_sweep(self, nglobal: int):
        """Global_sweep: Nglobal Wolff trials per chain in the chain's own direction."""
        check(lib().z2_global_sweep(self._ptr, int(nglobal)))

    def therm(self):
        """one LocalK_therm (boson-only) sweep"""
        check(lib().z2_therm(self._ptr))

    def lam(self, chain: int) -> np.ndarray:
        out = np.zeros(self.N)
        check(lib().z2_get_lambda(self._ptr, chain, _pd(out)))
        return out

    def is_beta(self, chain: int) -> bool:
        v = C.c_int32()
        check(lib().z2_get_is_beta(self._ptr, chain, C.byref(v)))
        return bool(v.value)

    def set_is_beta(self, chain: int, v: bool):
        check(lib().z2_set_is_beta(self._ptr, chain, int(bool(v))))

    def global_counters(self, chain: int):
        """(accepted, proposed, sum of cluster sizes), wolff_size_cluster"""
        o = (C.c_uint64 * 3)()
        z = C.c_double()
        check(lib().z2_get_global_counters(self._ptr, chain, o, C.byref(z)))
        return (int(o[0]), int(o[1]), int(o[2])), z.value

    def therm_counters(self, chain: int):
        o = (C.c_uint64 * 2)()
        check(lib().z2_get_therm_counters(self._ptr, chain, o))
        return int(o[0]), int(o[1])

    def enable_tau(self):
        check(lib().z2_enable_tau(self._ptr))

    def dyn_measure(self):
        """Dyn%reset + Dyn%sweep_R from the current state (accumulators zeroed, not normalised)."""
        check(lib().z2_dyn_measure(self._ptr))

    def obs_tau(self, chain: int) -> np.ndarray:
        """[7, Ltrot, N]: spin, pair_s, pair_d, charge_s, charge_d, single, curxx."""
        o = np.zeros(7 * self.T * self.N)
        check(lib().z2_get_obs_tau(self._ptr, chain, _pd(o)))
        return o.reshape(7, self.T, self.N)

    def dyn_drift(self, chain: int):
        a, b = np.zeros(3), np.zeros(3)
        check(lib().z2_get_dyn_drift(self._ptr, chain, _pd(a), _pd(b)))
        return a, b

    def sweep_shadow(self, ref: "Engine", is_beta: bool = True, toggle: bool = False) -> np.ndarray:
        """One sweep in lockstep with an FP64 reference engine, replaying its
        decisions (toggle: with the time-displaced measurement on both); returns
        max|G_self - G_ref| per chain before each recompute."""
        out = np.zeros(self.nch)
        check(lib().z2_sweep_shadow_tau(self._ptr, ref.ptr, int(is_beta), int(bool(toggle)), _pd(out)))
        return out

    def sweep_shadow_trace(self, ref: "Engine", is_beta: bool = True):
        """one shadow sweep; per step: (kind, nt, max over chains of max|G_self - G_ref|)"""
        cap = 8 * self.T + 64
        tr = np.zeros(cap)
        code = np.zeros(cap, np.int32)
        n = C.c_int32()
        check(lib().z2_sweep_shadow_trace(self._ptr, ref.ptr, int(is_beta), _pd(tr),
                                          code.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n)))
        return [(int(c) >> 16, int(c) & 0xFFFF, float(t)) for c, t in zip(code[:n.value], tr[:n.value])]

    # ---- fine-grained steps (reference procedure granularity), all chains at slice nt
    def _step(self, name, *args):
        check(getattr(lib(), name)(self._ptr, *[int(x) for x in args]))

    def sweep_reset(self, toggle=False):
        self._step("z2_sweep_reset", bool(toggle))

    def prop_K_pre(self, nt):
        self._step("z2_prop_K_pre", nt)

    def prop_T_pre(self, nt):
        self._step("z2_prop_T_pre", nt)

    def prop_T_L(self, nt):
        self._step("z2_prop_T_L", nt)

    def prop_T_R(self, nt):
        self._step("z2_prop_T_R", nt)

    def propgr_K_R(self, nt):
        self._step("z2_propgr_K_R", nt)

    def propgr_T_R(self, nt):
        self._step("z2_propgr_T_R", nt)

    def localK_prop_L(self, nt):
        self._step("z2_localK_prop_L", nt)

    def localK_prop_R(self, nt):
        self._step("z2_localK_prop_R", nt)

    def obs_equal_calc(self, nt):
        self._step("z2_obs_equal_calc", nt)

    def obs_equal_ave(self, nobs):
        self._step("z2_obs_equal_ave", nobs)

    def dyn_reset(self):
        self._step("z2_dyn_reset")

    def obs_tau_reset(self):
        self._step("z2_obs_tau_reset")

    def obs_tau_calc(self, nt):
        self._step("z2_obs_tau_calc", nt)

    def obs_tau_ave(self, nobst):
        self._step("z2_obs_tau_ave", nobst)

    def wrap_tau(self, nt):
        self._step("z2_wrap_tau", nt)

    def wrap_pre(self, nt):
        check(lib().z2_wrap_pre(self._ptr, nt))

    def wrap_L(self, nt, flag=1):
        check(lib().z2_wrap_L(self._ptr, nt, flag))

    def wrap_R(self, nt, flag=1):
        check(lib().z2_wrap_R(self._ptr, nt, flag))

    def slice_L(self, nt):
        check(lib().z2_slice_L(self._ptr, nt))

    def slice_R(self, nt):
        check(lib().z2_slice_R(self._ptr, nt))

    def sync(self):
        check(lib().z2_sync(self._ptr))

    # ---- state
    def sigma(self, chain: int) -> np.ndarray:
        s = np.zeros(2 * self.N * self.T)
        check(lib().z2_get_sigma(self._ptr, chain, _pd(s)))
        return s.reshape(self.T, 2 * self.N)

    def seed(self, chain: int) -> int:
        v = C.c_int32()
        check(lib().z2_get_seed(self._ptr, chain, C.byref(v)))
        return v.value

    def green(self, chain: int) -> np.ndarray:
        G = np.zeros((self.N, self.N), order="F")
        check(lib().z2_get_green(self._ptr, chain, _pd(G)))
        return G

    def set_green(self, chain: int, G: np.ndarray):
        A = np.asfortranarray(G, dtype=np.float64)
        check(lib().z2_set_green(self._ptr, chain, _pd(A)))

    def drift(self, chain: int):
        a, b = C.c_double(), C.c_double()
        check(lib().z2_get_drift(self._ptr, chain, C.byref(a), C.byref(b)))
        return a.value, b.value

    def obs_equal(self, chain: int) -> dict:
        o = np.zeros(7)
        check(lib().z2_get_obs_equal(self._ptr, chain, _pd(o)))
        return dict(zip(OBS_NAMES, o.tolist()))

    def counters(self, chain: int):
        o = (C.c_uint64 * 2)()
        check(lib().z2_get_counters(self._ptr, chain, o))
        return int(o[0]), int(o[1])

    def set_delay(self, k: int):
        check(lib().z2_set_delay(self._ptr, int(k)))

    def set_flip_variant(self, v: int):
        check(lib().z2_set_flip_variant(self._ptr, int(v)))

    def set_u_precision(self, bits: int):
        check(lib().z2_set_u_precision(self._ptr, int(bits)))

    def set_kin_tiles(self, rows_cb: int, cols_rb: int):
        check(lib().z2_set_kin_tiles(self._ptr, int(rows_cb), int(cols_rb)))

    def set_stab_mode(self, mode: int):
        """Z2_STAB_REFERENCE (default) or Z2_STAB_DSPLIT (D-split G recompute, opt-in)"""
        check(lib().z2_set_stab_mode(self._ptr, int(mode)))

    def set_qr_mode(self, mode: int):
        """Z2_QR_BLOCKED (columns pre-ordered by norm, blocked) or Z2_QR_PIVOTED (ZGEQP3 pivoting)."""
        check(lib().z2_set_qr_mode(self._ptr, int(mode)))

    def profile_kernel(self, name: str):
        from ._lib import KERNEL_IDS
        check(lib().z2_profile_kernel(self._ptr, KERNEL_IDS[name] if name else 0))

    def profile_read(self):
        t, n = C.c_double(), C.c_int64()
        check(lib().z2_profile_read(self._ptr, C.byref(t), C.byref(n)))
        return t.value, n.value

    def profile_spans(self, ref=None, cap: int = 1 << 20):
        """(start, end) in ms of the timed launches since ref's profile_kernel (ref: an
        engine view on the same device, profiled first; default self)"""
        st, en, n = np.zeros(cap), np.zeros(cap), C.c_int64()
        check(lib().z2_profile_spans(self._ptr, (ref or self).ptr, _pd(st), _pd(en), cap, C.byref(n)))
        return st[:n.value], en[:n.value]

    def profile_work(self) -> float:
        """algorithmic flops (or bytes) of the launches timed since profile_kernel"""
        w = C.c_double()
        check(lib().z2_profile_work(self._ptr, C.byref(w)))
        return w.value

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().z2_get_stream(self._ptr, C.byref(s)))
        return s.value or 0


class _CtxView(Engine):
    """an Engine view of a context owned by a Batch (not destroyed by the view)"""

    def __init__(self, ptr, N, T, nch):
        self._ptr = C.c_void_p(ptr)
        self._owned = False
        self.N, self.T, self.nch = N, T, nch

    def close(self):
        self._ptr = C.c_void_p()


@dataclass
class BatchChain:
    """one chain of a batch: its own lattice, Ltrot, Nwrap, couplings and seed"""
    L: int
    Ltrot: int
    Nwrap: int
    spec: ChainSpec = field(default_factory=ChainSpec)


class Batch:
    """Chains of uneven L, Ltrot, Nwrap and couplings, one context per shape group
    (z2_batch_* in include/z2dqmc.h)."""

    def __init__(self, chains, precision: int = Z2_PREC_FP64, Nsweep: int = 1, max_group: int = 0):
        self.chains = list(chains)
        arr = (Z2BatchChain * len(self.chains))()
        for k, c in enumerate(self.chains):
            s = c.spec
            arr[k] = Z2BatchChain(c.L, c.L, c.Ltrot, c.Nwrap,
                                  Z2ChainDesc(s.RT, s.J, s.h, s.mu, s.Q, s.Beta, int(s.seed), int(s.absolute)))
        self._ptr = C.c_void_p()
        check(lib().z2_batch_create(C.cast(arr, C.c_void_p), len(self.chains), precision, Nsweep, max_group,
                                    C.byref(self._ptr)))

    def close(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            lib().z2_batch_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, g = C.c_int32(), C.c_int32()
        check(lib().z2_batch_info(self._ptr, C.byref(n), C.byref(g)))
        return n.value, g.value

    def where(self, k: int):
        """(engine view of the chain's group context, index of the chain in it)"""
        p, i = C.c_void_p(), C.c_int32()
        check(lib().z2_batch_chain_ctx(self._ptr, k, C.byref(p), C.byref(i)))
        c = self.chains[k]
        N, nch, prec = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().z2_ctx_info(p, C.byref(N), C.byref(nch), C.byref(prec)))
        return _CtxView(p.value, c.L * c.L, c.Ltrot, nch.value), i.value

    def group(self, g: int) -> Engine:
        p = C.c_void_p()
        check(lib().z2_batch_group_ctx(self._ptr, g, C.byref(p)))
        N, nch, prec = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().z2_ctx_info(p, C.byref(N), C.byref(nch), C.byref(prec)))
        e = _CtxView(p.value, N.value, 0, nch.value)
        return e

    def init_fields(self):
        check(lib().z2_batch_init_fields(self._ptr))

    def enable_tau(self):
        check(lib().z2_batch_enable_tau(self._ptr))

    def pre(self):
        check(lib().z2_batch_pre(self._ptr))

    def sweep(self, is_beta=True, toggle=False):
        b = Z2_BETA_PER_CHAIN if is_beta is None else int(bool(is_beta))
        check(lib().z2_batch_sweep(self._ptr, b, int(bool(toggle))))

    def sync(self):
        check(lib().z2_batch_sync(self._ptr))

    def set_exec(self, mode: int):
        check(lib().z2_batch_set_exec(self._ptr, int(mode)))

    def exec_mode(self) -> int:
        m = C.c_int32()
        check(lib().z2_batch_get_exec(self._ptr, C.byref(m)))
        return m.value

    def autotune(self, warm=1, sweeps=1, drift_tol=5e-3, path=None):
        check(lib().z2_batch_autotune(self._ptr, int(warm), int(sweeps), float(drift_tol),
                                      path.encode() if path else None))

    def load_tuning(self, path):
        check(lib().z2_batch_load_tuning(self._ptr, path.encode()))

    # per-chain results through the chain's group context
    def _call(self, k, name, *a):
        e, i = self.where(k)
        return getattr(e, name)(i, *a)

    def sigma(self, k):
        return self._call(k, "sigma")

    def seed(self, k):
        return self._call(k, "seed")

    def green(self, k):
        return self._call(k, "green")

    def counters(self, k):
        return self._call(k, "counters")

    def obs_equal(self, k):
        return self._call(k, "obs_equal")

    def drift(self, k):
        return self._call(k, "drift")

