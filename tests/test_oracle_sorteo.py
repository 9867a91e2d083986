"""The oracle's particle sampler (sorteo.f90 restated with a Philox stream):
the generator against numpy's Philox4x64-10, trilinear interpolation exact on
affine densities, and the sample's distribution against the density."""
import numpy as np
import pytest


def test_philox_matches_numpy():
    """numpy's Philox (4x64, 10 rounds) increments its 256-bit counter before
    each block: its outputs for counter c0 are our blocks c0+1, c0+2, ..."""
    from oracle import hetddft_oracle as orc
    for key, c0 in (((1, 2), 0), ((0xDEADBEEF12345678, 7), 41), ((2**64 - 1, 2**64 - 1), 2**64 - 3)):
        bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array([c0, 0, 0, 0], dtype=np.uint64))
        raw = [int(v) for v in bg.random_raw(12)]
        ours = []
        for b in range(1, 4):
            c = c0 + b
            ours += orc.philox4x64((c & (2**64 - 1), c >> 64, 0, 0), key)
        assert ours == raw


def test_interpolation_exact_on_affine_density():
    from oracle import hetddft_oracle as orc
    prm = orc.Params(nx=8, ny=6, nz=10, xmax=2.0, ymax=1.5, zmax=2.5, n_par=1)
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    den = 3.0 + 0.5 * X - 0.25 * Y + 0.125 * Z
    rng = np.random.default_rng(1)
    for _ in range(50):
        p = [rng.uniform(x[0], x[-1] - 1e-9), rng.uniform(y[0], y[-1] - 1e-9), rng.uniform(z[0], z[-1] - 1e-9)]
        assert abs(orc.interpola3denx(prm, den, p) - (3.0 + 0.5 * p[0] - 0.25 * p[1] + 0.125 * p[2])) < 1e-12
    assert orc.interpola3denx(prm, den, [x[-1] + 0.01, 0.0, 0.0]) == 0.0   # outside [1, n-1]


def test_sample_follows_density():
    """A Gaussian cloud (sigma 0.6 A) sampled in the whole box: sample mean and
    variance within 5 standard errors of the density's."""
    from oracle import hetddft_oracle as orc
    prm = orc.Params(nx=24, ny=24, nz=24, xmax=4.0, ymax=4.0, zmax=4.0, n_par=1)
    x = prm.axis_grid("x")
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s0, mu = 0.6, np.array([0.3, -0.2, 0.1])
    den = np.exp(-((X - mu[0]) ** 2 + (Y - mu[1]) ** 2 + (Z - mu[2]) ** 2) / (2 * s0 * s0))
    n = 400
    r, tries = orc.sorteo(prm, den, n, [-4.0] * 3, [4.0 - prm.hx] * 3, seed=5)
    assert np.all(tries >= 1)
    for c in range(3):
        assert abs(r[c].mean() - mu[c]) < 5 * s0 / np.sqrt(n)
        assert abs(r[c].var() - s0 * s0) < 5 * s0 * s0 * np.sqrt(2.0 / n) + 0.02
