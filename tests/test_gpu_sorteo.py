"""Device particle sampler (hetddft_sample_particles) against the oracle's
sorteo with the same Philox stream: identical positions; the distribution of a
large sample; the bounded-attempts error path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gauss(prm, mu=(0.3, -0.2, 0.1), s0=0.6):
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.exp(-((X - mu[0]) ** 2 + (Y - mu[1]) ** 2 + (Z - mu[2]) ** 2) / (2 * s0 * s0))


def _instance(hip, n, h, n_par):
    p = hip.default_params(nx=n, ny=n, nz=n, xmax=n * h / 2, ymax=n * h / 2, zmax=n * h / 2, n_par=n_par)
    return hip.Instance(p)


def test_sampler_matches_oracle(hip):
    from oracle import hetddft_oracle as orc
    n, h, npar = 24, 1.0 / 3.0, 300
    prm = orc.Params(nx=n, ny=n, nz=n, xmax=n * h / 2, ymax=n * h / 2, zmax=n * h / 2, n_par=npar)
    den = _gauss(prm)
    lo, hi = [-2.5, -3.0, -2.0], [3.0, 2.5, 3.5]
    print("sorteo device vs oracle", flush=True)
    with _instance(hip, n, h, npar) as g:
        got = g.sample_particles(den, lo, hi, seed=123456789)
    ref, _ = orc.sorteo(prm, den, npar, lo, hi, seed=123456789)
    assert np.array_equal(got, ref)


def test_sampler_distribution(hip):
    n, h, npar = 32, 0.25, 20000
    from oracle import hetddft_oracle as orc
    prm = orc.Params(nx=n, ny=n, nz=n, xmax=n * h / 2, ymax=n * h / 2, zmax=n * h / 2, n_par=npar)
    mu, s0 = np.array([0.3, -0.2, 0.1]), 0.6
    den = _gauss(prm, mu, s0)
    print("sorteo distribution", flush=True)
    with _instance(hip, n, h, npar) as g:
        r = g.sample_particles(den, [-4.0] * 3, [3.75] * 3, seed=7)
    for c in range(3):
        assert abs(r[c].mean() - mu[c]) < 5 * s0 / np.sqrt(npar)
        assert abs(r[c].var() - s0 * s0) < 5 * s0 * s0 * np.sqrt(2.0 / npar) + 0.01


def test_sampler_gives_up_after_max_attempts(hip):
    """Density only outside the sampling box: every attempt is rejected, each
    thread stops after max_attempts and the call reports the failure."""
    n, h = 16, 0.5
    den = np.zeros((n, n, n))
    den[2, 2, 2] = 1.0                       # near the corner (-3.0, -3.0, -3.0)
    with _instance(hip, n, h, 50) as g:
        with pytest.raises(hip.HetddftError):
            g.sample_particles(den, [1.0] * 3, [3.0] * 3, seed=1, max_attempts=2000)
