"""Time-displaced measurement on the GPU vs the CPU oracle (FP64 path).

Between recomputes G(t,0), G(0,t), G(t,t) are propagated with the same
checkerboard arithmetic as the oracle (bit-identical); the Wrap_tau recompute
inverts the 2N x 2N block matrix with QR/triangular solves where the oracle
uses LU inverses, so values agree to rounding: observables within 1e-9 of the
largest entry of each correlator array."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("spin", "pair_s", "pair_d", "charge_s", "charge_d", "single", "curxx")


def _setup(L, T, beta, w, specs, Nsweep=1):
    import z2_gauge_zfy_amd as z
    e = z.Engine(L, L, T, w, specs, Nsweep=Nsweep)
    e.enable_tau()
    e.init_fields()
    chains = []
    for s in specs:
        p = O.make_params(RT=s.RT, J=s.J, h=s.h, mu=s.mu, Q=s.Q, Nlx=L, Nly=L, Ltrot=T, Beta=beta,
                          Nwrap=w, absolute=s.absolute, is_tau=1, Nsweep=Nsweep)
        m = O.Model(p)
        chains.append((m, O.Chain(m, s.seed)))
    return e, chains


def _gpu_pass_L(e, T, w):
    for nt in range(T, 0, -1):
        if nt % w == 0:
            e.wrap_L(nt)
        e.slice_L(nt)
    e.wrap_L(0)


def _compare_tau(got, want, tol=1e-9):
    for k, n in enumerate(NAMES):
        ref = want[n].T
        scale = max(1.0, np.abs(ref).max())
        err = np.abs(got[k] - ref).max() / scale
        assert err < tol, (n, err)


def test_dyn_measure_after_left_pass():
    import z2_gauge_zfy_amd as z
    L, T, beta, w = 4, 20, 2.0, 5
    specs = [z.ChainSpec(h=0.5, mu=0.2, Beta=beta, seed=2468),
             z.ChainSpec(h=0.3, mu=0.0, J=1.2, Beta=beta, seed=97531)]
    e, chains = _setup(L, T, beta, w, specs)
    e.pre()
    _gpu_pass_L(e, T, w)
    e.dyn_measure()
    for k, (m, c) in enumerate(chains):
        c.pre()
        c.pass_L()
        assert np.array_equal(e.sigma(k), c.sigma()), "field history"
        c.dyn_measure()
        print(f"chain {k}", flush=True)
        _compare_tau(e.obs_tau(k), c.obs_tau())
        xg, _ = e.dyn_drift(k)
        xo, _ = c.dyn_drift()
        assert np.all(xg < 1e-9) and np.all(xo < 1e-9), (xg, xo)


@pytest.mark.parametrize("L,T,w", [(4, 40, 10), (6, 24, 6)])
def test_sweeps_with_toggle(L, T, w):
    import z2_gauge_zfy_amd as z
    beta = T / 10.0
    specs = [z.ChainSpec(h=0.5, mu=0.1 * k, Beta=beta, seed=12345 + 7919 * k) for k in range(2)]
    e, chains = _setup(L, T, beta, w, specs)
    e.pre()
    for _, c in chains:
        c.pre()
    for sw in range(3):
        e.sweep(True, toggle=True)
        for k, (m, c) in enumerate(chains):
            c.sweep(toggle=True)
            print(f"chain {k} sweep {sw}: min gap {c.min_gap:.3e}", flush=True)
            if not np.array_equal(e.sigma(k), c.sigma()):
                assert c.min_gap < 1e-12, f"field history diverged at sweep {sw} (chain {k})"
                return
            _compare_tau(e.obs_tau(k), c.obs_tau())
            xg, mg = e.dyn_drift(k)
            xo, mo = c.dyn_drift()
            assert np.all(xg < 1e-6), xg
            np.testing.assert_allclose(xg, xo, rtol=0.5, atol=1e-12)


def test_tau_requires_enable():
    import z2_gauge_zfy_amd as z
    e = z.Engine(4, 4, 10, 5, [z.ChainSpec(Beta=1.0)])
    e.init_fields()
    e.pre()
    with pytest.raises(z.Z2Error, match="not enabled"):
        e.sweep(True, toggle=True)
    with pytest.raises(z.Z2Error, match="not enabled"):
        e.obs_tau(0)


def test_mixed_precision_tau_close_to_fp64():
    """Mixed mode runs the time-displaced path in FP64 from the FP32-propagated G;
    after a sweep in which the mixed engine replays the FP64 engine's decisions
    (z2_sweep_shadow_tau, the measurement at the start of the right pass), its
    correlators stay close to the FP64 run."""
    import z2_gauge_zfy_amd as z
    L, T, w, beta = 4, 20, 5, 2.0
    specs = [z.ChainSpec(h=0.5, mu=0.2, Beta=beta, seed=4242)]
    e64 = z.Engine(L, L, T, w, specs)
    e32 = z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_MIXED)
    for e in (e64, e32):
        e.enable_tau()
        e.init_fields()
        e.pre()
    e32.sweep_shadow(e64, True, toggle=True)
    assert np.array_equal(e64.sigma(0), e32.sigma(0)), "the replayed decisions must give the same field"
    assert e64.seed(0) == e32.seed(0)
    a, b = e64.obs_tau(0), e32.obs_tau(0)
    for k in range(7):
        scale = max(1.0, np.abs(a[k]).max())
        assert np.abs(a[k] - b[k]).max() / scale < 1e-3, NAMES[k]
