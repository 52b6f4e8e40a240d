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
