"""Pins of four oracle functions that the reference's own fixtures do not reach,
each against what the mathematics fixes:

* conf_set with absolute = 1 (src/fields.f90:322-333) and lambda_random_fill +
  lambda_enforce_sector (:559-603): the field and lambda from the closed-form
  ranf stream computed here with integer arithmetic;
* LocalK_therm (src/localK.f90:400-420) driven by Local_sweep_therm
  (src/local_sweep.f90:73-85): the seed after a warm-up sweep in closed form, and
  the distribution the boson-only Metropolis chain samples against the exact
  Boltzmann weights of all 2^16 fields of an L = 2, Ltrot = 2 lattice;
* Obs_tau_calc (src/obser_tau.f90:56-133) at the first slice after Dyn%reset,
  where G(0,0) = G(t,t) = G(t,0) = G and G(0,t) = -(I - G)
  (src/process_matrix.f90:88-96): sums over displacements against Obs_equal_calc's
  own sums on the same G (src/obser_equal.f90:73-149, pinned by the reference's
  test_run series) and against trace identities evaluated by matrix algebra;
* conf_transfer (src/fields.f90:429-507): the run-size field as the periodic
  tiling of the thermalisation-size field, built here by index arithmetic.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

A_LCG, M_LCG = 48828125, 2 ** 31


def lcg_draws(seed, n):
    """ran_imada.f:55-68 in closed form: iq_k = seed * a^k mod 2^31, ranf = iq_k / 2^31"""
    out = np.empty(n)
    iq = seed % M_LCG
    for k in range(n):
        iq = iq * A_LCG % M_LCG
        out[k] = iq / M_LCG
    return out, iq


# ------------------------------------------------------------ conf_set (a3)
@pytest.mark.parametrize("L, T, seed, Q", [(4, 8, 12345, 1.0), (6, 4, 987654321, -1.0), (2, 2, 7, 1.0)])
def test_conf_set_absolute_1_closed_form(L, T, seed, Q):
    N = L * L
    p = O.make_params(Nlx=L, Nly=L, Ltrot=T, Beta=1.0, Nwrap=T, Q=Q, absolute=1)
    c = O.Chain(O.Model(p), seed)
    r, iq = lcg_draws(seed, 2 * N * T + N)
    # tau outer, bond inner: draw (nt - 1) * 2N + ii; +1 iff ranf >= 0.5
    sigma = np.where(r[:2 * N * T] >= 0.5, 1.0, -1.0).reshape(T, 2 * N)
    lam = np.where(r[2 * N * T:] >= 0.5, 1.0, -1.0)
    if np.prod(lam) * (1.0 if Q >= 0 else -1.0) < 0:  # sector fix flips the last site
        lam[-1] = -lam[-1]
    assert np.array_equal(c.sigma(), sigma)
    assert np.array_equal(c.lam(), lam)
    assert np.prod(c.lam()) == (1.0 if Q >= 0 else -1.0)
    assert c.seed == iq


# ------------------------------------------------------ LocalK_therm (f4)
def _plaquettes(L):
    """bond numbers (1-based) of the plaquette at every site: (s,+x), (s+x,+y),
    (s+y,+x), (s,+y); site (x,y) -> (y-1) L + x, bond 2 s - 1 = +x, 2 s = +y"""
    site = lambda x, y: ((y - 1) % L) * L + ((x - 1) % L) + 1
    out = []
    for y in range(1, L + 1):
        for x in range(1, L + 1):
            s, sx, sy = site(x, y), site(x + 1, y), site(x, y + 1)
            out.append((2 * s - 1, 2 * sx, 2 * sy - 1, 2 * s))
    return out


def _exact_means(L, T, dtau, J, h):
    """<sum of plaquette products> and <sum of temporal neighbour products> under
    W = exp(dtau J sum_tau sum_plaq prod sigma + K sum_tau sum_b sigma_b(tau) sigma_b(tau+1)),
    K = -1/2 ln tanh(dtau h), tau periodic -- by enumerating every field"""
    N2 = 2 * L * L
    K = -0.5 * np.log(np.tanh(dtau * h))
    plaq = np.array(_plaquettes(L)) - 1
    cfg = np.array(list(itertools.product((-1.0, 1.0), repeat=N2 * T))).reshape(-1, T, N2)
    P = np.prod(cfg[:, :, plaq], axis=3).sum(axis=(1, 2))
    Tn = (cfg * np.roll(cfg, -1, axis=1)).sum(axis=(1, 2))
    logw = dtau * J * P + K * Tn
    w = np.exp(logw - logw.max())
    w /= w.sum()
    return (w * P).sum(), (w * Tn).sum(), P, Tn, w


def test_therm_seed_closed_form():
    L, T = 4, 6
    p = O.make_params(Nlx=L, Nly=L, Ltrot=T, Beta=1.5, Nwrap=T, h=0.7)
    c = O.Chain(O.Model(p), 24680)
    s0 = c.seed
    for k in range(3):
        c.therm()  # one ranf per (nt, ii)
        assert c.seed == s0 * pow(A_LCG, 2 * L * L * T * (k + 1), M_LCG) % M_LCG


def test_therm_samples_the_boson_weight():
    """L = 2 (8 bonds), Ltrot = 2: 2^16 fields.  The boson-only chain's averages of
    the plaquette sum and the temporal-neighbour sum agree with the exact ones
    within 3 standard errors (batch means); flipping the sign of either coupling
    in the exact weights moves them far outside."""
    L, T, beta, J, h = 2, 2, 1.0, 1.0, 0.5
    dtau = beta / T
    ep, et, P, Tn, w = _exact_means(L, T, dtau, J, h)
    p = O.make_params(Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=T, J=J, h=h)
    c = O.Chain(O.Model(p), 13579)
    plaq = np.array(_plaquettes(L)) - 1
    nsw, nb = 60000, 60
    for _ in range(200):
        c.therm()
    ps, ts = np.empty(nsw), np.empty(nsw)
    for k in range(nsw):
        c.therm()
        s = c.sigma()
        ps[k] = np.prod(s[:, plaq], axis=2).sum()
        ts[k] = (s * np.roll(s, -1, axis=0)).sum()
    for name, x, exact in (("plaquette", ps, ep), ("temporal", ts, et)):
        bm = x.reshape(nb, -1).mean(axis=1)
        err = bm.std(ddof=1) / np.sqrt(nb)
        print(f"{name}: MC {x.mean():.4f} +- {err:.4f}, exact {exact:.4f}")
        assert abs(x.mean() - exact) < 3 * err, name
    # the test discriminates: the exact means with a wrong sign are far away
    ep_wrong, et_wrong, *_ = _exact_means(L, T, dtau, -J, h)
    assert abs(ps.mean() - ep_wrong) > 20 * ps.reshape(nb, -1).mean(axis=1).std(ddof=1) / np.sqrt(nb)


# ------------------------------------------------------ Obs_tau_calc (a18)
def test_obs_tau_first_slice_identities():
    """after a left pass (the place of Dyn%reset), the tau = 0 column of every
    correlator, summed over displacements, from the oracle's Obs_tau_calc and
    from independent evaluations on the same G"""
    L, T, beta, w, RT = 4, 20, 2.0, 5, 1.0
    p = O.make_params(RT=RT, J=1.0, h=0.5, mu=0.2, Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=w, is_tau=1)
    m = O.Model(p)
    c = O.Chain(m, 2468)
    c.pre()
    c.sweep()
    c.pass_L()
    G = c.green()
    terms = c.obs_equal_terms(1)
    c.dyn_measure()
    o = c.obs_tau()  # name -> [N, Ltrot], displacement imj - 1
    N = L * L
    I = np.eye(N)
    n = 1.0 - np.diag(G)
    # Obs_equal_calc's own sums on the same G
    assert o["pair_s"][:, 0].sum() == pytest.approx(8.0 * terms["pair_sum"], rel=1e-12)
    assert o["pair_d"][:, 0].sum() == pytest.approx(8.0 * terms["pair_sum"], rel=1e-12)
    zero = m.tables()["imj"][0, 0] - 1  # imj(i, i): the zero displacement
    assert o["single"][zero, 0] == pytest.approx(4.0 * N - 2.0 * terms["density_slice"], rel=1e-12)
    # trace identities (matrix algebra)
    assert o["pair_s"][:, 0].sum() == pytest.approx(8.0 * np.sum((I - G) ** 2), rel=1e-12)
    assert o["single"][:, 0].sum() == pytest.approx(4.0 * G.sum(), rel=1e-12, abs=1e-12)
    exch = np.trace(G) - np.trace(G @ G)  # sum_ij (delta_ij - G_ji) G_ij
    assert o["charge_s"][:, 0].sum() == pytest.approx(16.0 * n.sum() ** 2 + 4.0 * exch, rel=1e-12)
    assert o["charge_d"][:, 0].sum() == pytest.approx(4.0 * exch, rel=1e-10, abs=1e-10)
    s1 = c.sigma()[0, :N]
    assert o["spin"][:, 0].sum() == pytest.approx(s1.sum() ** 2, abs=1e-12)
    # curxx: with S the +x neighbour permutation (S e_i = e_{i+x}), X = G(0,t), Y = G(t,0),
    # sum_ij term_ij = -tr(XSYS) - tr(S^T X S^T Y) + tr(X S^T Y S) + tr(S^T X S Y)
    Lb = m.tables()["L_bonds"]
    S = np.zeros((N, N))
    for i in range(N):
        S[Lb[i, 1] - 1, i] = 1.0
    X, Y = -(I - G), G
    tsum = (-np.trace(X @ S @ Y @ S) - np.trace(S.T @ X @ S.T @ Y) + np.trace(X @ S.T @ Y @ S)
            + np.trace(S.T @ X @ S @ Y))
    assert o["curxx"][:, 0].sum() == pytest.approx(-4.0 * RT * RT * tsum, rel=1e-10, abs=1e-10)


# --------------------------------------------------- conf_transfer (a3)
@pytest.mark.parametrize("LxT, LyT, TT, L, T, Q", [(2, 2, 10, 4, 20, 1.0), (4, 2, 5, 4, 20, -1.0),
                                                  (2, 4, 20, 4, 20, 1.0), (3, 3, 4, 6, 12, -1.0)])
def test_conf_transfer_is_the_periodic_tiling(LxT, LyT, TT, L, T, Q):
    """src/fields.f90:429-507: every bond of the run-size lattice takes the value of
    the bond of the same direction at the site (x mod LxT, y mod LyT) of the
    thermalisation-size field, slice nt takes slice (nt - 1) mod TT + 1, lambda is
    tiled the same way and then put into the sector of Q (last site flipped when the
    product has the wrong sign); the seed is the one the thermalisation-size conf_set
    leaves (conf_set's draws are pinned above)"""
    seed = 97531
    pT = O.make_params(Nlx=LxT, Nly=LyT, Ltrot=TT, Beta=1.0, Nwrap=TT, Q=Q, absolute=1)
    cT = O.Chain(O.Model(pT), seed)
    sT, lT = cT.sigma(), cT.lam()
    p = O.make_params(Nlx=L, Nly=L, Ltrot=T, Beta=1.0, Nwrap=T, Q=Q, absolute=1,
                      NlxTherm=LxT, NlyTherm=LyT, LtrotTherm=TT)
    c = O.Chain(O.Model(p), seed)
    N = L * L
    sig = np.empty((T, 2 * N))
    lam = np.empty(N)
    for y in range(L):
        for x in range(L):
            s, st = y * L + x, (y % LyT) * LxT + (x % LxT)
            for d in range(2):  # 0: +x bond 2s-1, 1: +y bond 2s (1-based)
                for nt in range(T):
                    sig[nt, 2 * s + d] = sT[nt % TT, 2 * st + d]
            lam[s] = lT[st]
    if np.prod(lam) * (1.0 if Q >= 0 else -1.0) < 0:
        lam[-1] = -lam[-1]
    assert np.array_equal(c.sigma(), sig)
    assert np.array_equal(c.lam(), lam)
    assert c.seed == cT.seed
