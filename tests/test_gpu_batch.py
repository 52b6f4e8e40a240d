"""Batches of chains of uneven lattice size, Ltrot, Nwrap, beta, h and mu
(z2_batch_*): every chain follows its oracle chain (the reference runs one chain
per MPI rank); the eager and the graph-replayed sweeps give the same results bit
for bit, replay after replay; the autotuner's choices persist and reload."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle(bc):
    s = bc.spec
    p = O.make_params(RT=s.RT, J=s.J, h=s.h, mu=s.mu, Q=s.Q, Nlx=bc.L, Nly=bc.L, Ltrot=bc.Ltrot, Beta=s.Beta,
                      Nwrap=bc.Nwrap, absolute=s.absolute)
    m = O.Model(p)
    return m, O.Chain(m, s.seed)


def _uneven(z, n_per=2):
    out = []
    shapes = [(4, 20, 5, 2.0), (6, 24, 6, 3.0), (8, 20, 10, 2.5)]
    k = 0
    for L, T, w, beta in shapes:
        for j in range(n_per):
            out.append(z.BatchChain(L, T, w, z.ChainSpec(h=0.3 + 0.25 * j, mu=0.2 * (k % 2), J=1.0 + 0.1 * j,
                                                         Beta=beta, seed=4321 + 2 * 7919 * k)))
            k += 1
    return out


def _same(a, b, k, tag):
    assert np.array_equal(a.sigma(k), b.sigma(k)), tag
    assert a.seed(k) == b.seed(k), tag
    assert np.array_equal(a.green(k), b.green(k)), tag
    assert a.counters(k) == b.counters(k), tag
    assert a.obs_equal(k) == b.obs_equal(k), tag


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(precision):
    """one batch sweeps eagerly, an identical one through the captured graph (the
    first sweep eager, then replays); after every sweep both agree bit for bit;
    the groups are split (max_group) so the graph has several parallel branches"""
    import z2_gauge_zfy_amd as z
    prec = z.Z2_PREC_FP64 if precision == "fp64" else z.Z2_PREC_MIXED
    chains = _uneven(z, n_per=3)
    a = z.Batch(chains, precision=prec, max_group=2)
    b = z.Batch(chains, precision=prec, max_group=2)
    assert a.info() == (9, 6)
    b.set_exec(z.Z2_EXEC_GRAPH)
    for x in (a, b):
        x.init_fields()
        x.pre()
    for sw in range(4):
        a.sweep(True)
        b.sweep(True)
        a.sync()
        b.sync()
        for k in range(len(chains)):
            _same(a, b, k, f"{precision} sweep {sw} chain {k}")
        print(f"{precision} sweep {sw}: graph == eager", flush=True)


def test_uneven_batch_follows_the_oracle():
    """FP64: every chain of an uneven batch (three lattice sizes, Ltrot, Nwrap,
    beta; couplings and mu per chain) against its oracle chain over two sweeps,
    graph mode"""
    import z2_gauge_zfy_amd as z
    chains = _uneven(z, n_per=2)
    bt = z.Batch(chains)
    bt.set_exec(z.Z2_EXEC_GRAPH)
    bt.init_fields()
    orc = [_oracle(c) for c in chains]
    for k, (m, c) in enumerate(orc):
        assert np.array_equal(bt.sigma(k), c.sigma()) and bt.seed(k) == c.seed
        c.pre()
    bt.pre()
    for sw in range(2):
        bt.sweep(True)
        bt.sync()
        for k, (m, c) in enumerate(orc):
            c.sweep()
            assert np.array_equal(bt.sigma(k), c.sigma()) or c.min_gap < 1e-12, (sw, k)
            assert bt.seed(k) == c.seed
            G, Go = bt.green(k), c.green()
            assert np.abs(G - Go).max() < 1e-10 * max(1.0, np.abs(Go).max()), (sw, k)
            assert bt.counters(k)[0] == c.counters()["local_acc"]


def test_c3_shaped_batch_one_sweep():
    """the uneven-batch workload: 256 chains, L in {8, 10, 12} (86/85/85), h on a
    grid of 32 values over [0.1, 1.5], beta = 10, Ltrot = 100, Nwrap = 10, FP64,
    graph mode; two chains of every size checked against the oracle for one sweep"""
    import z2_gauge_zfy_amd as z
    hs = np.linspace(0.1, 1.5, 32)
    counts = {8: 86, 10: 85, 12: 85}
    chains, k = [], 0
    for L, n in counts.items():
        for j in range(n):
            chains.append(z.BatchChain(L, 100, 10, z.ChainSpec(h=float(hs[k % 32]), Beta=10.0,
                                                               seed=12345 + 2 * 7919 * k)))
            k += 1
    bt = z.Batch(chains)
    assert bt.info() == (256, 3)
    bt.set_exec(z.Z2_EXEC_GRAPH)
    bt.init_fields()
    picks = [0, 85, 86, 170, 171, 255]
    orc = {k: _oracle(chains[k]) for k in picks}

    def run(c):
        c.pre()
        c.sweep()

    with ThreadPoolExecutor(len(picks)) as pool:
        futs = [pool.submit(run, c) for m, c in orc.values()]
        bt.pre()
        bt.sweep(True)
        bt.sync()
        for f in futs:
            f.result()
    for k, (m, c) in orc.items():
        tag = f"chain {k} (L={chains[k].L}, h={chains[k].spec.h:.3f})"
        assert np.array_equal(bt.sigma(k), c.sigma()) or c.min_gap < 1e-12, tag
        assert bt.seed(k) == c.seed, tag
        G, Go = bt.green(k), c.green()
        assert np.abs(G - Go).max() < 1e-10 * max(1.0, np.abs(Go).max()), tag
        print(tag, "identical", flush=True)


def test_autotune_persists_and_reloads(tmp_path):
    """the autotuner leaves the batch's own chains untouched (same fields as an
    untuned batch after init), writes its choices, and a fresh batch loading the
    file gets the same settings and runs the same sweep bit for bit"""
    import z2_gauge_zfy_amd as z
    chains = [z.BatchChain(L, 20, 5, z.ChainSpec(h=0.5, Beta=2.0, seed=999 + 2 * k))
              for k, L in enumerate((4, 4, 6, 6))]
    path = str(tmp_path / "tune.txt")
    a = z.Batch(chains)
    a.init_fields()
    ref = z.Batch(chains)
    ref.init_fields()
    a.autotune(warm=1, sweeps=1, drift_tol=5e-3, path=path)
    text = open(path).read()
    print(text)
    lines = text.splitlines()
    assert sum(x.startswith("group ") for x in lines) == 2 and any(x.startswith("exec ") for x in lines)
    for k in range(len(chains)):
        assert np.array_equal(a.sigma(k), ref.sigma(k)) and a.seed(k) == ref.seed(k)
    b = z.Batch(chains)
    b.init_fields()
    b.load_tuning(path)
    assert b.exec_mode() == a.exec_mode()
    for x in (a, b):
        x.pre()
        x.sweep(True)
        x.sweep(True)
        x.sync()
    for k in range(len(chains)):
        _same(a, b, k, f"chain {k}")
    with pytest.raises(z.Z2Error, match="before z2_batch_pre"):
        a.autotune()


def test_profile_spans_lay_groups_on_one_time_line():
    """z2_profile_spans: every timed flip launch of two stream groups, measured from
    the first group's z2_profile_kernel; per group one launch per slice step, in
    order, and their durations sum to z2_profile_read's total"""
    import z2_gauge_zfy_amd as z
    T = 20
    chains = [z.BatchChain(4, T, 5, z.ChainSpec(h=0.4 + 0.2 * k, Beta=2.0, seed=777 + 2 * 7919 * k)) for k in range(4)]
    b = z.Batch(chains, precision=z.Z2_PREC_MIXED, max_group=2)
    assert b.info() == (4, 2)
    b.init_fields()
    b.pre()
    b.sync()
    for g in range(2):
        b.group(g).profile_kernel("flip")
    b.sweep(True)
    b.sync()
    ref = b.group(0)
    spans = [b.group(g).profile_spans(ref) for g in range(2)]  # before any profile_kernel resets ref
    for g in range(2):
        e = b.group(g)
        st, en = spans[g]
        total, n = e.profile_read()
        assert n == 2 * T and len(st) == n, (g, n, len(st))
        assert np.all(st >= -1e-3) and np.all(en >= st), g
        assert np.all(np.diff(st) >= 0) and np.all(st[1:] >= en[:-1] - 1e-3), g  # one stream: in order
        assert abs(float(np.sum(en - st)) - total) <= 1e-2 * total + 1e-3, (g, float(np.sum(en - st)), total)
    for g in range(2):
        b.group(g).profile_kernel("")
