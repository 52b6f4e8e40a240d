"""The reference's procedures one call each (Multiply_mod, LocalK_mod,
ObserEqual/ObserTau calc and ave, Dyn%reset, Wrap_tau, Wrap_pre/L/R), composed
in the order of src/local_sweep.f90:87-195 and src/dynamics.f90:41-62, against
the batched z2_pre + z2_sweep_tau on a second context with the same chains: the
fields, seeds, Green's functions, equal-time and time-displaced observables
agree bit for bit, in FP64 and in the mixed mode (mu != 0 on some chains, so
the exp(+-dtau mu) steps run with real factors)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fine_pre(e, T, w):
    """Local_sweep_pre, src/local_sweep.f90:87-104"""
    e.sweep_reset(False)
    e.wrap_pre(0)
    for nt in range(1, T + 1):
        e.prop_K_pre(nt)
        e.prop_T_pre(nt)
        if nt % w == 0:
            e.wrap_pre(nt)


def fine_sweep(e, T, w, toggle):
    """Local_sweep with is_beta (src/local_sweep.f90:152-195): Local_sweep_L, Local_sweep_R"""
    e.sweep_reset(toggle)
    nobs = nobst = 0
    for nt in range(T, 0, -1):                      # Local_sweep_L, :106-124
        if nt % w == 0:
            e.wrap_L(nt)
        e.obs_equal_calc(nt)
        nobs += 1
        e.prop_T_L(nt)
        e.localK_prop_L(nt)
    e.wrap_L(0)
    if toggle:                                      # Dyn%reset + Dyn%sweep_R, src/dynamics.f90:41-62
        e.dyn_reset()
        for nt in range(1, T + 1):
            e.obs_tau_calc(nt)
            e.propgr_K_R(nt)
            e.propgr_T_R(nt)
            if nt % w == 0:
                e.wrap_tau(nt)
        nobst += 1
    e.wrap_R(0)                                     # Local_sweep_R, :126-150
    for nt in range(1, T + 1):
        e.localK_prop_R(nt)
        e.prop_T_R(nt)
        if nt % w == 0:
            e.wrap_R(nt)
        e.obs_equal_calc(nt)
        nobs += 1
    e.obs_equal_ave(nobs)
    if toggle:
        e.obs_tau_ave(nobst)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("L, T, w", [(4, 20, 5), (8, 24, 6), (16, 20, 10)])
def test_fine_grained_sweep_matches_batched(precision, L, T, w):
    import z2_gauge_zfy_amd as z
    prec = z.Z2_PREC_FP64 if precision == "fp64" else z.Z2_PREC_MIXED
    specs = [z.ChainSpec(h=0.4 + 0.3 * k, mu=0.25 * k, Beta=2.0, seed=777 + 2 * 7919 * k) for k in range(3)]
    a = z.Engine(L, L, T, w, specs, precision=prec)
    b = z.Engine(L, L, T, w, specs, precision=prec)
    for e in (a, b):
        e.enable_tau()
        e.init_fields()
    fine_pre(a, T, w)
    b.pre()
    for k in range(len(specs)):
        assert np.array_equal(a.green(k), b.green(k)), "pre"
    for sw in range(3):
        toggle = sw == 1
        fine_sweep(a, T, w, toggle)
        b.sweep(True, toggle=toggle)
        for k in range(len(specs)):
            tag = f"{precision} L={L} sweep {sw} chain {k}"
            assert np.array_equal(a.sigma(k), b.sigma(k)), tag
            assert a.seed(k) == b.seed(k), tag
            assert a.counters(k) == b.counters(k), tag
            assert np.array_equal(a.green(k), b.green(k)), tag
            assert a.obs_equal(k) == b.obs_equal(k), tag
            assert a.drift(k) == b.drift(k), tag
            if toggle:
                assert np.array_equal(a.obs_tau(k), b.obs_tau(k)), tag
                assert [x.tolist() for x in a.dyn_drift(k)] == [x.tolist() for x in b.dyn_drift(k)], tag
        print(f"{precision} L={L} sweep {sw} (toggle {toggle}): bit-identical", flush=True)


def test_dyn_reset_refuses_the_wrong_place():
    """Dyn%reset belongs at the start of a right pass; after pre (or a right pass)
    the second propagator would start from the full product B(beta, 0)"""
    import z2_gauge_zfy_amd as z
    e = z.Engine(4, 4, 20, 5, [z.ChainSpec(Beta=2.0)])
    e.enable_tau()
    e.init_fields()
    e.pre()
    with pytest.raises(z.Z2Error, match="right pass"):
        e.dyn_reset()
    with pytest.raises(z.Z2Error, match="right pass"):
        e.dyn_measure()
