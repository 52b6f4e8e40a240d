"""The Fortran driver (reference module names over the C ABI) runs a small
job on the GPU and writes the reference's per-bin files and confout.txt; the
results match the CPU oracle run on the same seeds."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortran", "build", "z2_gauge_gpu.out")


def test_fortran_driver_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("fortran driver not built")
    seeds = [12345, 12345 + 2 * 7919]
    (tmp_path / "paramC_sets.txt").write_text(
        "1.0 1.0 0.5 0.0 1\n4 4 40 4.0\n4 4 40\n10 3 1\n.false. 0\n.false. 0\n.false. 0\n1\n")
    (tmp_path / "seeds.txt").write_text("\n".join(str(s) for s in seeds) + "\n")
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    flux = [float(x) for x in (tmp_path / "flux_avg").read_text().split()]
    dens = [float(x) for x in (tmp_path / "density").read_text().split()]
    assert len(flux) == 3 and len(dens) == 3
    p = O.make_params(RT=1.0, J=1.0, h=0.5, mu=0.0, Nlx=4, Nly=4, Ltrot=40, Beta=4.0, Nwrap=10)
    m = O.Model(p)
    chains = [O.Chain(m, s) for s in seeds]
    for c in chains:
        c.pre()
    for b in range(3):
        obs = []
        for c in chains:
            c.sweep()
            obs.append(c.obs_equal())
        f = 0.0
        d = 0.0
        for o in obs:
            f += o["flux_avg"]
            d += o["density"]
        assert flux[b] == pytest.approx(f / 2, rel=1e-15, abs=1e-15)
        assert dens[b] == pytest.approx(d / 2, rel=1e-10)
    conf = (tmp_path / "confout.txt").read_text().split()
    n = 1 + 2 * 16 * 40 + 16  # seed, sigma, lambda
    for k, c in enumerate(chains):
        blk = conf[k * n:(k + 1) * n]
        assert int(blk[0]) == c.seed
        assert np.array_equal(np.array([float(x) for x in blk[1:1281]]), c.sigma().reshape(-1))
        assert np.array_equal(np.array([float(x) for x in blk[1281:]]), c.lam())


GOLD = os.path.join(ROOT, "tests", "golden", "test_run")


def _gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return f.read()


def test_fortran_driver_replays_reference_test_run(tmp_path):
    """The reference's test_run/ job files (paramC_sets.txt with Wolff updates,
    seeds.txt, fresh start) run on one chain, as the reference's single MPI
    rank did: confout.txt equals the reference's file value for value, the
    per-bin flux files and the info.txt acceptance figures agree."""
    if not os.path.exists(EXE):
        pytest.skip("fortran driver not built")
    for name in ("paramC_sets.txt", "seeds.txt"):
        (tmp_path / name).write_text(_gold(name))
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([EXE, "fp64", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = [float(x) for x in (tmp_path / "confout.txt").read_text().split()]
    ref = [float(x) for x in _gold("confout.txt").split()]
    assert got == ref
    for name in ("flux_avg", "chi_flux", "density", "pair_sc_q0", "cdw_pi", "cdw_ratio"):
        g = np.array([float(x) for x in (tmp_path / name).read_text().split()])
        want = np.array([float(x) for x in _gold(name + ".last100").split()])
        tol = 1e-14 if name in ("flux_avg", "chi_flux") else 1e-9
        np.testing.assert_allclose(g, want, rtol=tol, atol=1e-14, err_msg=name)
    info = {}
    for line in (tmp_path / "info.txt").read_text().splitlines():
        if ":" in line:
            k, v = line.rsplit(":", 1)
            info[k.strip()] = v.split()
    assert float(info["Accept_Kglobal"][0]) == pytest.approx(7.2499999999999995e-2, rel=1e-14)
    assert float(info["Average size of Wolff cluster"][0]) == pytest.approx(33.840000000000003, rel=1e-14)
    assert float(info["Accept_Klocal_shift"][0]) == pytest.approx(1.1089843750000000e-2, rel=1e-14)


def _numbers(path):
    """All numbers of a list-directed Fortran output file, complex (re,im) split."""
    txt = open(path).read().replace("(", " ").replace(")", " ").replace(",", " ")
    return np.array([float(x) for x in txt.split()])


def test_fortran_driver_time_displaced_files(tmp_path):
    """is_tau: the driver's time-displaced files (m_write_obs_tau layout) against a
    numpy transform of the oracle's chain-averaged correlators, and the inversion
    identity sum_k g(k) = g(r = 0) of the naive Fourier transform."""
    if not os.path.exists(EXE):
        pytest.skip("fortran driver not built")
    L, T, beta, nbin = 4, 20, 2.0, 2
    seeds = [2468, 2468 + 2 * 7919]
    (tmp_path / "paramC_sets.txt").write_text(
        f"1.0 1.0 0.5 0.2 1\n{L} {L} {T} {beta}\n{L} {L} {T}\n5 {nbin} 1\n.true. 0\n.false. 0\n.false. 0\n1\n")
    (tmp_path / "seeds.txt").write_text("\n".join(str(s) for s in seeds) + "\n")
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    p = O.make_params(RT=1.0, J=1.0, h=0.5, mu=0.2, Nlx=L, Nly=L, Ltrot=T, Beta=beta, Nwrap=5, is_tau=1)
    m = O.Model(p)
    chains = [O.Chain(m, s) for s in seeds]
    for c in chains:
        c.pre()
    N, dtau = L * L, beta / T
    # momenta (ix-1, iy-1) 2 pi / L and displacements (ix, iy) of site (iy-1) L + ix
    ix, iy = np.meshgrid(np.arange(1, L + 1), np.arange(1, L + 1))
    ix, iy = ix.reshape(-1), iy.reshape(-1)
    k = np.stack([(ix - 1) * 2 * np.pi / L, (iy - 1) * 2 * np.pi / L], 1)
    rv = np.stack([ix, iy], 1).astype(float)
    F = np.exp(1j * (k @ rv.T)) / N  # F[nk, imj]
    ipi = (L // 2) * L + L // 2  # 0-based index of (pi, pi)
    want = {n: [] for n in ("sdw_susc", "cdw_d_pipi_tau", "g_r0_tau", "curxx_freq", "pair_s_zero_susc")}
    for b in range(nbin):
        acc = None
        for c in chains:
            c.sweep(toggle=True)
            o = c.obs_tau()
            acc = o if acc is None else {n: acc[n] + o[n] for n in acc}
        avg = {n: acc[n] / len(chains) for n in acc}
        spin_k = F @ avg["spin"]
        susc = dtau * spin_k.sum(1)
        want["sdw_susc"].append(np.concatenate([np.concatenate([k[q], [susc[q].real, susc[q].imag]]) for q in range(N)]))
        # identity of the transform: sum over k of g(k) is the zero-displacement element
        assert abs(susc.sum() - dtau * avg["spin"][N - 1].sum()) < 1e-9 * max(1.0, np.abs(susc).max())
        cdk = F @ avg["charge_d"]
        want["cdw_d_pipi_tau"].append(np.concatenate([k[ipi], np.stack([cdk[ipi].real, cdk[ipi].imag], 1).reshape(-1)]))
        g0 = avg["single"][N - 1] / N
        want["g_r0_tau"].append(np.concatenate([rv[N - 1], np.stack([g0.real, g0.imag], 1).reshape(-1)]))
        cur0 = (F @ avg["curxx"])[0]
        t = np.arange(T)
        gw = dtau * np.array([(np.exp(2j * np.pi * t * w / T) * cur0).sum() for w in range(T)])
        want["curxx_freq"].append(np.stack([2 * np.pi * np.arange(T) / beta, gw.real, gw.imag], 1).reshape(-1))
        ps = dtau * (F @ avg["pair_s"]).sum(1)
        want["pair_s_zero_susc"].append(np.array([ps[0].real, ps[0].imag]))
    for name, parts in want.items():
        got = _numbers(tmp_path / name)
        ref = np.concatenate(parts)
        assert got.shape == ref.shape, (name, got.shape, ref.shape)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() < 1e-8 * scale, (name, np.abs(got - ref).max())


def test_fortran_driver_restart_from_confin(tmp_path):
    """confin.txt restart (conf_therm_in layout: seed, sigma, lambda per chain):
    a run restarted from a previous run's confout.txt follows the oracle chain
    built from that configuration."""
    if not os.path.exists(EXE):
        pytest.skip("fortran driver not built")
    seeds = [777, 777 + 2 * 7919]
    par = "1.0 1.0 0.5 0.0 1\n4 4 20 2.0\n4 4 20\n5 2 1\n.false. 0\n.false. 0\n.false. 0\n1\n"
    (tmp_path / "paramC_sets.txt").write_text(par)
    (tmp_path / "seeds.txt").write_text("\n".join(str(s) for s in seeds) + "\n")
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    first = (tmp_path / "confout.txt").read_text()
    (tmp_path / "confin.txt").write_text(first)
    for f in ("flux_avg", "density"):
        (tmp_path / f).unlink()
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = 1 + 2 * 16 * 20 + 16
    vals = first.split()
    p = O.make_params(RT=1.0, J=1.0, h=0.5, mu=0.0, Nlx=4, Nly=4, Ltrot=20, Beta=2.0, Nwrap=5)
    m = O.Model(p)
    flux = [float(x) for x in (tmp_path / "flux_avg").read_text().split()]
    conf2 = (tmp_path / "confout.txt").read_text().split()
    fsum = [0.0, 0.0]
    for kc in range(2):
        blk = vals[kc * n:(kc + 1) * n]
        sig = np.array([float(x) for x in blk[1:1 + 640]]).reshape(20, 32)
        lam = np.array([float(x) for x in blk[1 + 640:]])
        c = O.Chain(m, int(blk[0]), sigma_therm=sig, lambda_therm=lam)
        c.pre()
        for b in range(2):
            c.sweep()
            fsum[b] += c.obs_equal()["flux_avg"]
        blk2 = conf2[kc * n:(kc + 1) * n]
        assert int(blk2[0]) == c.seed
        assert np.array_equal(np.array([float(x) for x in blk2[1:1 + 640]]), c.sigma().reshape(-1))
        assert np.array_equal(np.array([float(x) for x in blk2[1 + 640:]]), c.lam())
    assert flux == pytest.approx([f / 2 for f in fsum], rel=1e-15, abs=1e-15)


CHECK = os.path.join(ROOT, "fortran", "build", "z2_finegrained_check.out")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fortran_fine_grained_sequence_is_bit_identical(tmp_path, precision):
    """The reference's procedure sequence through the Fortran modules
    (Local_sweep_pre/_L/_R of src/local_sweep.f90:87-150 calling Multiply_mod,
    LocalK_mod, Wrap_pre/L/R, Obs_equal%calc/ave and, with is_tau, Dyn%reset,
    Dyn%sweep_R, Obs_tau%calc, propgrK_R, propgrT_R, Wrap_tau, Obs_tau%ave: one
    engine call each) against z2_pre + z2_sweep_tau on a second context of the
    same three chains: fields, seeds, G and observables bit-identical after every
    bin (mu = 0.3; the first bin without, the next two with the time-displaced
    measurement)."""
    if not os.path.exists(CHECK):
        pytest.skip("fortran check program not built")
    (tmp_path / "paramC_sets.txt").write_text(
        "1.0 1.0 0.5 0.3 1\n6 6 24 3.0\n6 6 24\n6 3 1\n.true. 1\n.false. 0\n.false. 0\n1\n")
    (tmp_path / "seeds.txt").write_text("\n".join(str(12345 + 2 * 7919 * k) for k in range(3)) + "\n")
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([CHECK, precision], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "bit-identical over 3 bins" in r.stdout


def test_fortran_driver_thermalisation_size(tmp_path):
    """paramC_sets.txt with a thermalisation size (2 x 2 sites, 10 slices) below the
    run size (4 x 4, 20): the driver starts every chain by conf_set at that size
    and conf_transfer, and its confout.txt follows the oracle chain made the same way"""
    if not os.path.exists(EXE):
        pytest.skip("fortran driver not built")
    seeds = [777, 777 + 2 * 7919]
    (tmp_path / "paramC_sets.txt").write_text(
        "1.0 1.0 0.5 0.0 1\n4 4 20 2.0\n2 2 10\n5 2 1\n.false. 0\n.false. 0\n.false. 0\n1\n")
    (tmp_path / "seeds.txt").write_text("\n".join(str(s) for s in seeds) + "\n")
    (tmp_path / "confin.txt").write_text("0\n")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    p = O.make_params(RT=1.0, J=1.0, h=0.5, mu=0.0, Nlx=4, Nly=4, Ltrot=20, Beta=2.0, Nwrap=5,
                      NlxTherm=2, NlyTherm=2, LtrotTherm=10)
    m = O.Model(p)
    conf = (tmp_path / "confout.txt").read_text().split()
    n = 1 + 2 * 16 * 20 + 16
    for k, s in enumerate(seeds):
        c = O.Chain(m, s)
        c.pre()
        for _ in range(2):
            c.sweep()
        blk = conf[k * n:(k + 1) * n]
        assert int(blk[0]) == c.seed
        assert np.array_equal(np.array([float(x) for x in blk[1:1 + 640]]), c.sigma().reshape(-1))
        assert np.array_equal(np.array([float(x) for x in blk[1 + 640:]]), c.lam())
