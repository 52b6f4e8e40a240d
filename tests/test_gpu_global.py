"""Wolff global update and boson warm-up on the GPU against the reference's own
end-to-end fixture (test_run/) and against the CPU oracle.

Bar: field history, RNG state, per-chain pass direction and all counters
identical; G within 1e-10 relative; a global decision may differ only where
|min(log ratio, 0) - log r| < 1e-12 (oracle global_min_gap)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "test_run")


def _lines(name):
    with open(os.path.join(GOLD, name)) as f:
        return [float(x) for x in f.read().split()]


def _close_G(Gg, Go, tol=1e-10):
    scale = max(1.0, np.abs(Go).max())
    err = np.abs(Gg - Go).max() / scale
    assert err < tol, err


def test_test_run_replay_on_gpu():
    """The reference's test_run/: L=4, beta=2, Ltrot=40, t=0.5, J=1, h=0.52,
    mu=0.3, Nwrap=5, absolute=5, seed 12439, 100 bins of (Global_sweep with
    Nglobal=4, then Local_sweep in the chain's direction).  Final seed and field
    are the reference's confout.txt, the counters its info.txt, the flux series
    its per-bin files."""
    import z2_gauge_zfy_amd as z
    spec = z.ChainSpec(RT=0.5, J=1.0, h=0.52, mu=0.3, Beta=2.0, seed=12439, absolute=5)
    e = z.Engine(4, 4, 40, 5, [spec])
    e.init_fields()
    e.pre()
    obs = []
    for nbc in range(100):
        e.global_sweep(4)
        e.sweep(None)
        obs.append(e.obs_equal(0))
        if nbc % 20 == 19:
            print(f"bin {nbc + 1}: global {e.global_counters(0)}, local {e.counters(0)}", flush=True)
    conf = _lines("confout.txt")
    assert e.seed(0) == int(conf[0]) == 1854583831
    assert np.array_equal(e.sigma(0).reshape(-1), np.array(conf[1:1281]))
    assert e.counters(0) == (2839, 256000)
    (gacc, gprop, csum), size_cluster = e.global_counters(0)
    assert (gacc, gprop, csum) == (29, 400, 13536)
    assert size_cluster / 100 == pytest.approx(33.840000000000003, rel=1e-15)
    for name in ("flux_avg", "chi_flux"):
        ref = np.array(_lines(name + ".last100"))
        assert np.array_equal(np.array([o[name] for o in obs]), ref), name
    for name in ("density", "cdw_pi", "pair_sc_q0", "cdw_ratio"):
        ref = np.array(_lines(name + ".last100"))
        got = np.array([o[name] for o in obs])
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12, err_msg=name)
    xmax, _ = e.drift(0)
    assert xmax == pytest.approx(2.3676807588932092e-3, rel=1e-6)


def _oracle(L, T, beta, w, spec, nglobal):
    p = O.make_params(RT=spec.RT, J=spec.J, h=spec.h, mu=spec.mu, Q=spec.Q, Nlx=L, Nly=L, Ltrot=T,
                      Beta=beta, Nwrap=w, absolute=spec.absolute, is_global=1, Nglobal=nglobal)
    m = O.Model(p)
    return m, O.Chain(m, spec.seed)


@pytest.mark.parametrize("precision", ["fp64"])
def test_wolff_batched_chains_follow_own_direction(precision):
    """Six chains whose Wolff trials are accepted at different bins, so their
    pass directions differ and the engine regroups them; every chain follows
    its oracle chain bin by bin."""
    import z2_gauge_zfy_amd as z
    L, T, beta, w, ng = 4, 20, 1.0, 5, 3
    specs = [z.ChainSpec(RT=0.5, J=1.0, h=0.4 + 0.05 * k, mu=(0.2 if k % 2 else 0.0), Beta=beta,
                         seed=12439 + 2 * 7919 * k, absolute=5) for k in range(6)]
    e = z.Engine(L, L, T, w, specs)
    e.init_fields()
    chains = [_oracle(L, T, beta, w, s, ng) for s in specs]
    e.pre()
    for _, c in chains:
        c.pre()
        c.is_beta = True
    flips = 0
    for nbc in range(6):
        e.global_sweep(ng)
        for _, c in chains:
            c.global_sweep()
        dirs = [e.is_beta(k) for k in range(len(specs))]
        for k, (_, c) in enumerate(chains):
            gap = c.global_min_gap
            if e.is_beta(k) != c.is_beta or not np.array_equal(e.sigma(k), c.sigma()):
                assert gap < 1e-12, f"chain {k} bin {nbc}: global decision diverged (gap {gap:.3e})"
                return
            assert e.seed(k) == c.seed, (k, nbc)
            (ga, gp, cs), sz = e.global_counters(k)
            kk = c.counters()
            assert (ga, gp, cs) == (kk["glob_acc"], kk["glob_prop"], kk["cluster_sum"]), (k, nbc)
            assert sz == c.acc()["size_cluster"]
            _close_G(e.green(k), c.green())
        e.sweep(None)
        for _, c in chains:
            c.sweep()
        for k, (_, c) in enumerate(chains):
            if not np.array_equal(e.sigma(k), c.sigma()):
                assert c.min_gap < 1e-12, f"chain {k} bin {nbc}: local history diverged"
                return
            assert e.seed(k) == c.seed
            _close_G(e.green(k), c.green())
            og, oo = e.obs_equal(k), c.obs_equal()
            assert og["flux_avg"] == oo["flux_avg"]
            assert abs(og["density"] - oo["density"]) < 1e-10
        flips += len(set(dirs)) > 1
        print(f"bin {nbc}: directions {dirs}", flush=True)
    assert flips > 0, "no bin had chains in both directions"


def test_therm_matches_oracle():
    """LocalK_therm sweeps: field, seed and counters identical to the oracle,
    including T = 2 (the two temporal neighbours are the same slice)."""
    import z2_gauge_zfy_amd as z
    for L, T in ((4, 12), (6, 2)):
        specs = [z.ChainSpec(RT=1.0, J=1.0, h=0.3 + 0.3 * k, Beta=2.0, seed=555 + 2 * k) for k in range(3)]
        e = z.Engine(L, L, T, 1 if T == 2 else 4, specs)
        e.init_fields()
        chains = []
        for s in specs:
            p = O.make_params(RT=s.RT, J=s.J, h=s.h, mu=s.mu, Nlx=L, Nly=L, Ltrot=T, Beta=2.0,
                              Nwrap=1 if T == 2 else 4)
            m = O.Model(p)
            chains.append((m, O.Chain(m, s.seed)))
        for _ in range(3):
            e.therm()
            for _, c in chains:
                c.therm()
        for k, (_, c) in enumerate(chains):
            assert np.array_equal(e.sigma(k), c.sigma()), (L, T, k)
            assert e.seed(k) == c.seed
            kk = c.counters()
            assert e.therm_counters(k) == (kk["therm_acc"], kk["therm_prop"])


def test_mixed_precision_wolff_runs():
    """mixed engine with global updates: runs and stays physical (its own decisions)"""
    import z2_gauge_zfy_amd as z
    specs = [z.ChainSpec(RT=0.5, J=1.0, h=0.52, Beta=2.0, seed=12439 + 2 * k, absolute=5) for k in range(4)]
    e = z.Engine(4, 4, 40, 5, specs, precision=z.Z2_PREC_MIXED)
    e.init_fields()
    e.pre()
    for _ in range(4):
        e.global_sweep(2)
        e.sweep(None)
    for k in range(4):
        (ga, gp, cs), _ = e.global_counters(k)
        assert gp == 8 and cs >= 8
        assert abs(e.obs_equal(k)["density"] - 1.0) < 1e-3
