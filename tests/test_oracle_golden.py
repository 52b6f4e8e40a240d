"""Pins of the CPU oracle against the reference's own data and against closed
forms.  The oracle is the checker for the GPU library; these tests make sure a
plausible restatement mistake (dropped term, wrong sign/index, transposed
operand) fails somewhere."""
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "test_run")
A_LCG, M_LCG = 48828125, 2 ** 31


def _lines(name):
    with open(os.path.join(GOLD, name)) as f:
        return [float(x) for x in f.read().split()]


# ---------------------------------------------------------------- RNG (a1)
def test_ranf_known_answers():
    """ran_imada.f:55-68 closed form iq <- iq*5^11 mod 2^31."""
    r, s = O.ranf_stream(12439, 4)
    assert [int(x * 2 ** 31) for x in r] == [1782658139, 670706831, 1569671539, 1516352071]
    assert r[0] == 0.8301148838363588
    r, s = O.ranf_stream(12345, 4)
    assert s == 2086348681


def test_ranf_jump_ahead_and_negative_seed():
    for seed in (1, 12439, 2**31 - 1, -5, -2**31 + 7):
        r, s = O.ranf_stream(seed, 1000)
        assert s == (seed % M_LCG) * pow(A_LCG, 1000, M_LCG) % M_LCG
        assert np.all((r >= 0) & (r < 1))


# ------------------------------------------------- end-to-end reference fixture
@pytest.fixture(scope="module")
def test_run_chain():
    """The reference's test_run/ directory: L=4, beta=2, Ltrot=40, t=0.5, J=1,
    h=0.52, mu=0.3, Nwrap=5, 100 bins, Wolff Nglobal=4, absolute=5, seed 12439."""
    p = O.make_params(RT=0.5, J=1.0, h=0.52, mu=0.3, Q=1.0, Nlx=4, Nly=4, Ltrot=40, Beta=2.0,
                      Nwrap=5, Nbin=100, Nsweep=1, is_global=1, Nglobal=4, absolute=5)
    m = O.Model(p)
    c = O.Chain(m, 12439)
    obs = c.run_program()
    return m, c, obs


def test_test_run_final_configuration(test_run_chain):
    m, c, obs = test_run_chain
    conf = _lines("confout.txt")
    assert c.seed == int(conf[0]) == 1854583831
    assert np.array_equal(c.sigma().reshape(-1), np.array(conf[1:1281]))
    assert np.array_equal(c.lam(), np.array(conf[1281:1297]))


def test_test_run_counters(test_run_chain):
    """info.txt:36-42 exact integer pins and the drift metric."""
    m, c, obs = test_run_chain
    k = c.counters()
    assert (k["local_acc"], k["local_prop"]) == (2839, 256000)
    assert (k["glob_acc"], k["glob_prop"]) == (29, 400)
    assert k["cluster_sum"] == 13536
    acc = c.acc()
    assert acc["Kl"] / 100 == pytest.approx(1.1089843750000000e-2, rel=1e-15)
    assert acc["Kg"] / 100 == pytest.approx(7.2499999999999995e-2, rel=1e-15)
    assert acc["size_cluster"] / 100 == pytest.approx(33.840000000000003, rel=1e-15)
    xmax, xmean = c.drift()
    assert xmax == pytest.approx(2.3676807588932092e-3, rel=1e-10)
    assert xmean / 1600 == pytest.approx(3.8208843973362001e-4, rel=1e-10)


def test_test_run_observable_series(test_run_chain):
    m, c, obs = test_run_chain
    for name in ("flux_avg", "chi_flux"):  # pure functions of sigma: exact
        ref = np.array(_lines(name + ".last100"))
        assert np.array_equal(np.array([o[name] for o in obs]), ref), name
    for name in ("density", "cdw_pi", "pair_sc_q0", "cdw_ratio"):
        ref = np.array(_lines(name + ".last100"))
        got = np.array([o[name] for o in obs])
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14, err_msg=name)


# ------------------------------------------------------------- lattice (a2)
@pytest.mark.parametrize("L", [2, 4, 6])
def test_lattice_groups_are_perfect_matchings(L):
    m = O.Model(O.make_params(Nlx=L, Nly=L, Ltrot=4, Nwrap=2))
    t = m.tables()
    N = L * L
    allb = np.concatenate(t["groups"])
    assert sorted(allb.tolist()) == list(range(1, 2 * N + 1))
    for g in t["groups"]:
        sites = t["bond_list"][g - 1].reshape(-1)
        assert len(set(sites.tolist())) == N  # every site exactly once


def test_plaquettes_from_bond_bonds_match_geometry():
    L = 4
    m = O.Model(O.make_params(Nlx=L, Nly=L, Ltrot=4, Nwrap=2))
    t = m.tables()
    bl, bb = t["bond_list"], t["bond_bonds"]

    def site(x, y):
        return ((y - 1) % L) * L + ((x - 1) % L) + 1

    def bond(x, y, d):  # d=0: +x, d=1: +y from site (x,y)
        return 2 * (site(x, y) - 1) + 1 + d

    plaqs = set()
    for x in range(1, L + 1):
        for y in range(1, L + 1):
            plaqs.add(frozenset([bond(x, y, 0), bond(x + 1, y, 1), bond(x, y + 1, 0), bond(x, y, 1)]))
    for b in range(1, 2 * L * L + 1):
        p1 = frozenset([b, *bb[b - 1, 1:4]])
        p2 = frozenset([b, *bb[b - 1, 4:7]])
        assert p1 in plaqs and p2 in plaqs and p1 != p2


# --------------------------------------------------- kinetic operator (a5-a7)
def _dense_B(m, sigma, nt, alpha):
    t = m.tables()
    N = m.N
    E = []
    for g in t["groups"]:
        K = np.zeros((N, N))
        for b in g:
            i, j = t["bond_list"][b - 1] - 1
            K[i, j] = K[j, i] = sigma[nt - 1, b - 1]
        E.append(expm(alpha * K))
    return E[0] @ E[1] @ E[2] @ E[3] @ E[3] @ E[2] @ E[1] @ E[0]


def test_checkerboard_matches_dense_exponentials():
    p = O.make_params(RT=1.0, Nlx=4, Nly=4, Ltrot=10, Beta=1.0, Nwrap=5)
    m = O.Model(p)
    rng = np.random.default_rng(1)
    sigma = rng.choice([-1.0, 1.0], size=(10, 32))
    nt = 3
    alpha = 0.5 * 0.1 * 1.0
    B = _dense_B(m, sigma, nt, alpha)
    X = rng.standard_normal((16, 16))
    np.testing.assert_allclose(m.kin_R(X, sigma, nt, 1), B @ X, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.kin_L(X, sigma, nt, 1), X @ B, rtol=1e-13, atol=1e-13)
    Bi = np.linalg.inv(B)
    np.testing.assert_allclose(m.kin_R(X, sigma, nt, -1), Bi @ X, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(m.kin_L(X, sigma, nt, -1), X @ Bi, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(B, B.T, atol=1e-15)
    # particle-hole: S B S = B^{-1} with S the sublattice sign
    x, y = np.meshgrid(np.arange(4), np.arange(4))
    S = np.diag(np.where((x + y).reshape(-1) % 2 == 0, 1.0, -1.0))
    np.testing.assert_allclose(S @ B @ S, Bi, atol=1e-12)


def test_delta_closed_form():
    p = O.make_params(RT=1.0, Ltrot=10, Beta=1.0, Nwrap=5)
    m = O.Model(p)
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    a2 = 0.1
    for so in (1.0, -1.0):
        sn = -so
        want = expm(a2 * sn * X) @ expm(-a2 * so * X) - np.eye(2)
        np.testing.assert_allclose(m.delta(so, sn), want, atol=1e-15)


def test_flip_update_is_exact_sherman_morrison():
    """G' = G (I + D~(I-G))^{-1}, det ratio = det[I + Delta (I-G)_PP]."""
    p = O.make_params(RT=1.0, Nlx=4, Nly=4, Ltrot=10, Beta=1.0, Nwrap=5)
    m = O.Model(p)
    rng = np.random.default_rng(7)
    G = rng.standard_normal((16, 16)) * 0.2 + 0.5 * np.eye(16)
    bl = m.tables()["bond_list"]
    for ii, so in ((5, 1.0), (12, -1.0)):
        Gn, det = m.flip_update(G, ii, so)
        i, j = bl[ii - 1] - 1
        Dt = np.zeros((16, 16))
        Dt[np.ix_([i, j], [i, j])] = m.delta(so, -so)
        M = np.eye(16) + Dt @ (np.eye(16) - G)
        np.testing.assert_allclose(Gn, G @ np.linalg.inv(M), rtol=1e-11, atol=1e-12)
        assert det == pytest.approx(np.linalg.det(M), rel=1e-12)


# ------------------------------------------------------------ boson ratio (a4)
def _action(sigma, L, dtau, J, h):
    T = sigma.shape[0]
    gamma = -0.5 * np.log(np.tanh(dtau * h))
    S = 0.0
    for t in range(T):
        s = sigma[t]
        for x in range(L):
            for y in range(L):
                b = lambda xx, yy, d: 2 * (((yy % L) * L) + (xx % L)) + d
                S += dtau * J * s[b(x, y, 0)] * s[b(x + 1, y, 1)] * s[b(x, y + 1, 0)] * s[b(x, y, 1)]
        S += gamma * np.sum(sigma[t] * sigma[(t + 1) % T])
    return S


def test_bosonratio_is_exp_of_action_difference():
    L, T = 4, 6
    p = O.make_params(Nlx=L, Nly=L, Ltrot=T, Beta=0.6, Nwrap=3, h=0.7, J=1.3)
    m = O.Model(p)
    rng = np.random.default_rng(3)
    sigma = rng.choice([-1.0, 1.0], size=(T, 2 * L * L))
    for ii, nt in ((1, 1), (7, 3), (32, 6), (18, 4)):
        sn = sigma.copy()
        sn[nt - 1, ii - 1] *= -1
        want = np.exp(_action(sn, L, 0.1, 1.3, 0.7) - _action(sigma, L, 0.1, 1.3, 0.7))
        assert m.bosonratio(sigma, sn, ii, nt) == pytest.approx(want, rel=1e-12)


# --------------------------------------------------------- stabilization (a12-14)
def test_qdrp_factorization():
    rng = np.random.default_rng(11)
    n = 12
    A = rng.standard_normal((n, n)) @ np.diag(np.logspace(0, 8, n))[:, rng.permutation(n)]
    M, D, ipvt, tau = O.qdrp(A)
    R = np.triu(M) * D[:, None]
    # rebuild Q from the reflectors
    Q = np.eye(n)
    for i in range(n):
        v = np.zeros(n)
        v[i] = 1.0
        v[i + 1:] = M[i + 1:, i]
        Q = Q @ (np.eye(n) - tau[i] * np.outer(v, v))
    np.testing.assert_allclose(Q @ R, A[:, ipvt - 1], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert np.all(np.diff(D) <= 1e-12 * D[:-1])  # pivoting orders |R_ii|
    np.testing.assert_allclose(np.abs(np.diag(M)), 1.0)


@pytest.mark.parametrize("nt", [3, 8])
def test_stab_green_equals_dense_inverse(nt):
    """both branches (nt <= Ltrot/2 and nt > Ltrot/2) of stab_green."""
    rng = np.random.default_rng(nt)
    n = 10

    def orth():
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        return q

    UR, UL = orth(), orth()
    VR = np.triu(rng.standard_normal((n, n))) * 0.3 + np.eye(n)
    VL = np.triu(rng.standard_normal((n, n))) * 0.3 + np.eye(n)
    DR, DL = np.logspace(3, -3, n), np.logspace(2, -2, n)
    G = O.stab_green(10, nt, UR, DR, VR, UL, DL, VL)
    X = np.eye(n) + UR @ np.diag(DR) @ VR @ VL @ np.diag(DL) @ UL
    np.testing.assert_allclose(G, np.linalg.inv(X), rtol=1e-9, atol=1e-11)


def test_pre_green_equals_dense_inverse_small_beta():
    L, T = 4, 8
    p = O.make_params(RT=1.0, Nlx=L, Nly=L, Ltrot=T, Beta=0.8, Nwrap=4, mu=0.2, h=0.5)
    m = O.Model(p)
    c = O.Chain(m, 4321)
    sigma = c.sigma()
    c.pre()
    ep, _ = m.mu_factors()
    B = np.eye(L * L)
    for nt in range(1, T + 1):
        B = ep * (_dense_B(m, sigma, nt, 0.05) @ B)
    np.testing.assert_allclose(c.green(), np.linalg.inv(np.eye(L * L) + B), rtol=1e-10,
                               atol=1e-12)


def test_half_filling_density_at_stabilization():
    """mu=0, even L: a freshly recomputed G has G_ii = 1/2 (particle-hole)."""
    p = O.make_params(RT=1.0, Nlx=4, Nly=4, Ltrot=20, Beta=2.0, Nwrap=5, mu=0.0, h=0.5)
    m = O.Model(p)
    c = O.Chain(m, 999)
    c.pre()
    np.testing.assert_allclose(np.diag(c.green()), 0.5, atol=1e-12)
    c.sweep()
    np.testing.assert_allclose(np.diag(c.green()), 0.5, atol=1e-12)


# ------------------------------------------------------ time-displaced (a17-a18)
def test_time_displaced_measurement_dense():
    """Dyn%sweep_R: G(tau,0) = B(tau,0) G(0), G(0,tau) = -(I-G(0)) B(tau,0)^{-1},
    G(tau,tau) = B(tau,0) G(0) B(tau,0)^{-1} with recomputes every Nwrap slices;
    the seven tau observables against dense formulas (small beta)."""
    L, T, beta, w, mu, RT = 4, 8, 0.8, 4, 0.2, 1.0
    p = O.make_params(RT=RT, J=1.0, h=0.5, mu=mu, Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=w,
                      is_tau=1)
    m = O.Model(p)
    c = O.Chain(m, 2468)
    c.pre()
    c.pass_L()
    sigma = c.sigma()
    G0 = c.green()
    c.dyn_measure()
    got = c.obs_tau()
    N = L * L
    ep, em = m.mu_factors()
    Bs = [ep * _dense_B(m, sigma, t, 0.5 * (beta / T) * RT) for t in range(1, T + 1)]
    Bbeta = np.eye(N)
    for B in Bs:
        Bbeta = B @ Bbeta
    np.testing.assert_allclose(G0, np.linalg.inv(np.eye(N) + Bbeta), rtol=1e-10, atol=1e-12)
    t = m.tables()
    imj, Lb = t["imj"], t["L_bonds"]
    want = {k: np.zeros((N, T)) for k in ("spin", "pair_s", "pair_d", "charge_s", "charge_d",
                                           "single", "curxx")}
    Bt0 = np.eye(N)
    I = np.eye(N)
    for nt in range(1, T + 1):
        Gt0 = Bt0 @ G0
        G0t = -(I - G0) @ np.linalg.inv(Bt0)
        Gtt = Bt0 @ G0 @ np.linalg.inv(Bt0)
        tt = 2 * (1 - np.diag(Gtt))
        t0 = 2 * (1 - np.diag(G0))
        for j in range(N):
            for i in range(N):
                q = imj[i, j] - 1
                want["spin"][q, nt - 1] += sigma[nt - 1, i] * sigma[0, j]
                want["pair_d"][q, nt - 1] += 8 * G0t[j, i] ** 2
                want["pair_s"][q, nt - 1] += 8 * G0t[j, i] ** 2
                a = G0t[j, i] * Gt0[i, j]
                want["charge_d"][q, nt - 1] += -4 * a
                want["charge_s"][q, nt - 1] += -4 * a + 4 * tt[i] * t0[j]
                want["single"][q, nt - 1] += 4 * Gt0[i, j]
                ip, jp = Lb[i, 1] - 1, Lb[j, 1] - 1
                term = (-G0t[j, ip] * Gt0[i, jp] - G0t[jp, i] * Gt0[ip, j] + G0t[j, i] * Gt0[ip, jp]
                        + G0t[jp, ip] * Gt0[i, j])
                want["curxx"][q, nt - 1] += -4 * RT * RT * term
        Bt0 = Bs[nt - 1] @ Bt0
    for k in want:
        np.testing.assert_allclose(got[k], want[k], rtol=1e-8, atol=1e-10, err_msg=k)
    xmax, _ = c.dyn_drift()
    assert np.all(xmax < 1e-9)  # no flips during the measurement: recomputes agree
