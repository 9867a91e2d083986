"""The oracle's droplet start (hetddft_init_droplet restated), pinned by what
does not come from the same code:

* the vectorised Philox4x64-10 equals the scalar one bit for bit (the scalar
  one equals numpy's Philox, test_oracle_sorteo.py), so the per-point and
  per-particle streams are the generator's;
* the Woods-Saxon density with ripple and cavity equals the workload
  generator's, which evaluates the harmonics with scipy's lpmv: the
  oracle's Rodrigues-form P_l^m (Condon-Shortley phase) against an
  independent routine; a flipped phase or a swapped cos/sin fails it;
* the density integrates to N, and psi's phase is x.lamda where there is no
  noise (DFT4He3d.f90:587-596);
* the Box-Muller normals have the moments and quantiles of N(0, 1); the
  particle cloud has mean centre and width sigma; velocities are v0 Ktops.
"""
import math

import numpy as np
import pytest

from conftest import oracle_params, rel_max


def _orc():
    from oracle import hetddft_oracle as orc
    return orc


def test_vectorised_philox_equals_the_scalar_one():
    orc = _orc()
    ctr = np.array([0, 1, 2, 7, 123456789, (1 << 40) + 3, (1 << 64) - 1], dtype=np.uint64)
    for key in ((0, 0), (5, 1), (0xDEADBEEFCAFEBABE, 2), ((1 << 64) - 1, (1 << 64) - 1)):
        w = orc.philox4x64_many(ctr, key)
        for i, c in enumerate(ctr):
            ref = orc.philox4x64((int(c), 0, 0, 0), key)
            assert [int(w[k][i]) for k in range(4)] == ref, (key, int(c))


def test_mulhilo_against_python_integers():
    orc = _orc()
    rng = np.random.default_rng(3)
    a = rng.integers(0, 1 << 63, size=1000, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    for m in (0xD2E7470EE14C6C93, 0xCA5A826395121157, (1 << 64) - 1, 1):
        with np.errstate(over="ignore"):
            hi, lo = orc._mulhilo64(a, m)
        for x, h, l in zip(a[:200], hi[:200], lo[:200]):
            p = int(x) * m
            assert int(h) == p >> 64 and int(l) == p & ((1 << 64) - 1)


@pytest.mark.parametrize("ripple,cavity", [(0.0, 0.0), (0.0, 4.2), (0.3, 0.0), (0.3, 4.2)])
def test_density_equals_the_workload_generator(ripple, cavity):
    import inputs
    orc = _orc()
    wl = inputs.Workload("r", 40, 0.5, 150, 64, 0.5, ripple=ripple, seed=5, cavity=cavity)
    prm = oracle_params(wl)
    ref = inputs.woods_saxon(wl)
    spec = inputs.droplet_spec(wl)
    got = orc.droplet_density(prm, spec["n_he"], cavity=spec["cavity"], cavity_w=spec["cavity_w"],
                              ripple=spec["ripple"], clm=spec["clm"])
    assert rel_max(got, ref) < 1e-12
    assert abs(got.sum() * prm.dxyz - 150.0) < 1e-10


def test_planted_legendre_mistakes_fail():
    """The comparison above would catch a wrong Condon-Shortley sign or a
    cos/sin swap: the perturbed densities differ from the generator's."""
    import inputs
    orc = _orc()
    wl = inputs.Workload("r", 32, 0.5, 150, 64, 0.5, ripple=0.3, seed=5)
    prm = oracle_params(wl)
    ref = inputs.woods_saxon(wl)
    spec = inputs.droplet_spec(wl)
    good = orc.assoc_legendre
    try:
        orc.assoc_legendre = lambda l, m, x: abs(good(l, m, x)) if m % 2 else good(l, m, x)
        bad = orc.droplet_density(prm, 150.0, ripple=0.3, clm=spec["clm"])
    finally:
        orc.assoc_legendre = good
    assert rel_max(bad, ref) > 1e-6
    clm = list(spec["clm"])
    clm[0], clm[2] = clm[2], clm[0]          # l = 1: sin(phi) <-> cos(phi)
    assert rel_max(orc.droplet_density(prm, 150.0, ripple=0.3, clm=clm), ref) > 1e-6


def test_psi_phase_and_norm():
    orc = _orc()
    prm = orc.Params(nx=32, ny=32, nz=32, xmax=8.0, ymax=8.0, zmax=8.0, n_par=16)
    lam = (0.3, -0.2, 0.5)
    psi, rimp, vimp = orc.droplet_state(prm, 120.0, seed=9, lamda=lam, v0=(1.0, 0.0, -2.0))
    assert rimp is None
    assert abs((np.abs(psi) ** 2).sum() * prm.dxyz - 120.0) < 1e-10
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    th = x[None, None, :] * lam[0] + y[None, :, None] * lam[1] + z[:, None, None] * lam[2]
    d = np.angle(psi * np.exp(-1j * th))
    assert np.abs(d).max() < 1e-12
    np.testing.assert_array_equal(vimp[0], 7.63822291)
    np.testing.assert_array_equal(vimp[2], -2.0 * 7.63822291)


def test_box_muller_normals():
    orc = _orc()
    prm = orc.Params(nx=48, ny=48, nz=48, xmax=12.0, ymax=12.0, zmax=12.0, n_par=20000)
    sig = 0.05
    psi, rimp, _ = orc.droplet_state(prm, 100.0, seed=11, phase_noise=sig, sigma=0.8, centre=(0.5, -1.0, 2.0))
    g = np.angle(psi).ravel() / sig            # |phase| << pi: the normals themselves
    n = g.size
    assert abs(g.mean()) < 5.0 / math.sqrt(n)
    assert abs(g.var() - 1.0) < 5.0 * math.sqrt(2.0 / n)
    for q, zq in ((0.1587, -1.0), (0.5, 0.0), (0.8413, 1.0), (0.0228, -2.0)):
        assert abs(np.quantile(g, q) - zq) < 0.02
    for c, (m0, s) in enumerate(((0.5, 0.8), (-1.0, 0.8), (2.0, 0.8))):
        assert abs(rimp[c].mean() - m0) < 5.0 * s / math.sqrt(prm.n_par)
        assert abs(rimp[c].std() - s) < 0.02 * s
    # g0 and g1 of a particle come from one Box-Muller pair: uncorrelated
    assert abs(np.corrcoef(rimp[0], rimp[1])[0, 1]) < 0.03
