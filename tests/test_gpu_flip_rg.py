"""The flip kernel with G in registers (mixed mode, N = 256, k_flip_rg) against
the oracle and against the delayed-update kernel (k_flip) on the same inputs.

Bar: the decisions (field, seed, counters) are the oracle's serial ones; G stays
within FP32 rounding of the oracle's FP64 G over a few slices (2e-5 of max|G|),
and within 1e-4 of the FP64 engine over a sweep with replayed decisions, as the
delayed-update kernel does."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_chain(L, T, beta, w, spec):
    p = O.make_params(RT=spec.RT, J=spec.J, h=spec.h, mu=spec.mu, Q=spec.Q, Nlx=L, Nly=L,
                      Ltrot=T, Beta=beta, Nwrap=w, absolute=spec.absolute)
    return O.Chain(O.Model(p), spec.seed)


@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_register_flip_slices_match_oracle(mu):
    """three left-pass and three right-pass slices (flips, wraps, mu factors) from a
    random G that FP32 represents exactly"""
    import z2_gauge_zfy_amd as z
    L, T, beta, w = 16, 20, 2.0, 10
    spec = z.ChainSpec(h=0.5, mu=mu, Beta=beta, seed=2024)
    e = z.Engine(L, L, T, w, [spec], precision=z.Z2_PREC_MIXED)
    e.set_flip_variant(z.Z2_FLIP_REGISTER)
    e.init_fields()
    c = _oracle_chain(L, T, beta, w, spec)
    rng = np.random.default_rng(5)
    N = L * L
    G = rng.standard_normal((N, N)) * 0.02 + 0.5 * np.eye(N)
    G = np.asfortranarray(G.astype(np.float32).astype(np.float64))
    e.set_green(0, G)
    c.set_green(G)
    c.reset_min_gap()
    acc0 = e.counters(0)[0]
    for nt, left in ((7, True), (6, True), (5, True), (5, False), (6, False), (7, False)):
        (e.slice_L if left else e.slice_R)(nt)
        (c.slice_L if left else c.slice_R)(nt)
        tag = f"{'L' if left else 'R'} slice {nt}"
        if not np.array_equal(e.sigma(0), c.sigma()):
            assert c.min_gap < 1e-5, f"decisions differ away from a near tie ({tag})"
            return
        scale = max(1.0, np.abs(c.green()).max())
        err = np.abs(e.green(0) - c.green()).max() / scale
        print(f"{tag}: max|G_rg - G_oracle| = {err:.2e}, accepts {e.counters(0)[0] - acc0}", flush=True)
        assert err < 2e-5, (tag, err)
    assert e.seed(0) == c.seed
    assert e.counters(0)[0] > acc0 + 20  # the slices accepted updates (pairs and a lone last one)


def test_register_and_delayed_kernels_shadow_fp64():
    """one sweep at L = 16 (Ltrot = 20, beta = 2) with recomputes, each mixed engine
    replaying its own FP64 engine's decisions: both stay within 1e-4 of FP64 and of
    each other; chains 1-3 have mu != 0"""
    import z2_gauge_zfy_amd as z
    L, T, beta, w = 16, 20, 2.0, 10
    specs = [z.ChainSpec(h=0.4 + 0.2 * k, mu=0.2 * k, Beta=beta, seed=31337 + 2 * 7919 * k) for k in range(4)]
    refs = [z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_FP64) for _ in range(2)]
    mixes = [z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_MIXED) for _ in range(2)]
    mixes[0].set_flip_variant(z.Z2_FLIP_REGISTER)
    mixes[1].set_flip_variant(z.Z2_FLIP_DELAYED)
    for e in refs + mixes:
        e.init_fields()
        e.pre()
    d = [mixes[i].sweep_shadow(refs[i], True) for i in range(2)]
    print("max|dG| register", d[0], "delayed", d[1], flush=True)
    assert np.all(d[0] < 1e-4) and np.all(d[1] < 1e-4), d
    for k in range(len(specs)):
        assert np.array_equal(mixes[0].sigma(k), refs[0].sigma(k))
        assert np.array_equal(mixes[0].sigma(k), mixes[1].sigma(k))
        assert mixes[0].counters(k) == mixes[1].counters(k)
        dg = np.abs(mixes[0].green(k) - mixes[1].green(k)).max()
        assert dg < 1e-4, (k, dg)


def test_register_kernel_own_decisions_follow_oracle():
    """mixed engine on its own decisions at L = 16 from pre(): one sweep's field
    history equals the FP64 oracle's unless a decision falls within FP32 rounding
    of its random number"""
    import z2_gauge_zfy_amd as z
    L, T, beta, w = 16, 20, 2.0, 10
    spec = z.ChainSpec(h=0.6, mu=-0.2, Beta=beta, seed=4711)
    e = z.Engine(L, L, T, w, [spec], precision=z.Z2_PREC_MIXED)
    e.set_flip_variant(z.Z2_FLIP_REGISTER)
    e.init_fields()
    c = _oracle_chain(L, T, beta, w, spec)
    assert np.array_equal(e.sigma(0), c.sigma())
    c.pre()
    e.pre()
    c.reset_min_gap()
    e.sweep(True)
    c.sweep()
    if not np.array_equal(e.sigma(0), c.sigma()):
        assert c.min_gap < 1e-5
        return
    assert e.seed(0) == c.seed
    assert e.counters(0)[0] == c.counters()["local_acc"]
    scale = max(1.0, np.abs(c.green()).max())
    assert np.abs(e.green(0) - c.green()).max() / scale < 1e-4


def test_register_kernel_wolff_trials_follow_oracle():
    """Wolff global updates through the register kernel's forced mode (the trial
    flips one bond at the cluster's slices and accumulates log |det|^2), then a
    local sweep in each chain's direction, at L = 16: directions, fields, seeds
    and global counters are the FP64 oracle's, unless a decision falls within
    FP32 rounding of its random number (the oracle's smallest gap)"""
    import z2_gauge_zfy_amd as z
    L, T, beta, w, ng = 16, 20, 1.0, 10, 2
    specs = [z.ChainSpec(RT=1.0, J=1.0, h=0.5 + 0.2 * k, mu=0.1 * k, Beta=beta, seed=2468 + 2 * 7919 * k)
             for k in range(2)]
    e = z.Engine(L, L, T, w, specs, precision=z.Z2_PREC_MIXED)
    e.set_flip_variant(z.Z2_FLIP_REGISTER)
    e.init_fields()
    chains = []
    for s in specs:
        p = O.make_params(RT=s.RT, J=s.J, h=s.h, mu=s.mu, Q=s.Q, Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=w,
                          absolute=s.absolute, is_global=1, Nglobal=ng)
        chains.append(O.Chain(O.Model(p), s.seed))
    e.pre()
    for c in chains:
        c.pre()
        c.is_beta = True
    for nbc in range(2):
        e.global_sweep(ng)
        for c in chains:
            c.global_sweep()
        for k, c in enumerate(chains):
            if e.is_beta(k) != c.is_beta or not np.array_equal(e.sigma(k), c.sigma()):
                assert c.global_min_gap < 1e-4, f"chain {k} bin {nbc}: global decision away from a near tie"
                return
            (ga, gp, cs), _ = e.global_counters(k)
            kk = c.counters()
            assert (ga, gp, cs) == (kk["glob_acc"], kk["glob_prop"], kk["cluster_sum"]), (k, nbc)
        e.sweep(None)
        for c in chains:
            c.reset_min_gap()
            c.sweep()
        for k, c in enumerate(chains):
            if not np.array_equal(e.sigma(k), c.sigma()):
                assert c.min_gap < 1e-5, f"chain {k} bin {nbc}: local decision away from a near tie"
                return
            assert e.seed(k) == c.seed
            scale = max(1.0, np.abs(c.green()).max())
            assert np.abs(e.green(k) - c.green()).max() / scale < 1e-4
        print(f"bin {nbc}: directions {[e.is_beta(k) for k in range(2)]}, global {[e.global_counters(k)[0] for k in range(2)]}",
              flush=True)
