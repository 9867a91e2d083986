"""Pins the oracle's FFT layer, integrators and particle/grid operators by mathematics."""
import math

import numpy as np
import pytest

from oracle import hetddft_oracle as orc


# ---------------------------------------------------------------- FFT layer
def _dft3(f):
    nz, ny, nx = f.shape
    out = np.zeros((nz, ny, nx // 2 + 1), complex)
    z = np.arange(nz)[:, None, None]
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kz * z / nz + ky * y / ny + kx * x / nx))
                out[kz, ky, kx] = np.sum(f * ph)
    return out


def test_r2c_is_the_dft():
    f = np.random.default_rng(1).standard_normal((5, 6, 8))
    np.testing.assert_allclose(orc.r2c(f), _dft3(f), rtol=0, atol=1e-12)


def test_c2r_roundtrip_with_renor():
    f = np.random.default_rng(2).standard_normal((6, 10, 12))
    back = orc.c2r(orc.r2c(f), f.shape, 1.0 / f.size)
    assert np.abs(back - f).max() < 1e-14


# ------------------------------------------------------ RK-Gill + Hamming PMC
def _linear_oracle(dt_steps, lam_v, w2, lambdah=0.0):
    """Oracle on a 2x2x2 grid: the stencil is empty (n < 13), so
    Hpsi = timec*(-V psi); particles feel a = -w2 r."""
    prm = orc.Params(nx=2, ny=2, nz=2, xmax=1.0, ymax=1.0, zmax=1.0, n_par=3, lambdah=lambdah)
    H = (2, 2, 2)
    k = {n: np.zeros(H) for n in ("fvlj4", "wcgk", "kalfs", "vq")}
    o = orc.Oracle(prm, k)
    o.psi = np.full(prm.shape, 0.6 + 0.8j)
    o.pot4 = np.full(prm.shape, lam_v)
    o.rimp = np.array([[1.0, 0.5, -0.3], [0.0, 0.2, 0.1], [0.4, -1.0, 0.0]])
    o.vimp = np.array([[0.0, 0.1, 0.3], [1.0, 0.0, -0.2], [0.0, 0.5, 0.0]])

    def veval():
        o.aimp = -w2 * o.rimp
    o.veval = veval
    o.den = o.dens(o.psi)
    veval()
    return o


def _run(o, dt, nsteps):
    for it in range(nsteps):
        o.iterate(it, dt)


@pytest.mark.parametrize("lambdah", [0.0, 0.4])
def test_first_gill_step_is_taylor4(lambdah):
    V, dt = 3.0, 0.05
    o = _linear_oracle(1, V, 1.0, lambdah)
    y0 = o.psi[0, 0, 0]
    o.steprk(dt)
    tc = o.timec_z[0]
    z = tc * (-V) * dt
    taylor = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert abs(o.psi[0, 0, 0] - taylor * y0) < 1e-14


@pytest.mark.parametrize("lambdah", [0.0, 0.4])
def test_rk_start_plus_hamming_is_fourth_order(lambdah):
    """Global error of 3 Gill steps + Hamming PMC on y' = -i V y and r'' = -w^2 r.

    With the modifier and the final 9/121 correction (steppc.f90:52,106) the
    scheme is fifth order on a fixed interval, so the error falls about 32x
    per halving of dt; a wrong history slot or rotation breaks the order."""
    V, w2, T = 2.0, 1.7, 1.2
    w = math.sqrt(w2)
    errs_psi, errs_r, errs_v = [], [], []
    for nsteps in (96, 192, 384):
        dt = T / nsteps
        o = _linear_oracle(nsteps, V, w2, lambdah)
        r0, v0 = o.rimp.copy(), o.vimp.copy()
        y0 = o.psi.copy()
        _run(o, dt, nsteps)
        lam = o.timec_z[:, None, None] * (-V)
        errs_psi.append(np.abs(o.psi - y0 * np.exp(lam * T)).max())
        errs_r.append(np.abs(o.rimp - (r0 * math.cos(w * T) + v0 / w * math.sin(w * T))).max())
        errs_v.append(np.abs(o.vimp - (-r0 * w * math.sin(w * T) + v0 * math.cos(w * T))).max())
    for errs in (errs_psi, errs_r, errs_v):
        r1 = errs[0] / errs[1]
        r2 = errs[1] / errs[2]
        assert 24.0 < r1 < 40.0 and 24.0 < r2 < 40.0, errs


def test_pc_error_estimates_scale():
    """errHe ~ (9/121)|p - c| shrinks like dt^5 per step (Milne-Hamming estimate)."""
    V, w2 = 2.0, 1.7
    est = []
    for dt in (0.02, 0.01):
        o = _linear_oracle(8, V, w2)
        for it in range(3):
            o.iterate(it, dt)
        est.append(o.steppc(dt)[0])
    ratio = est[0] / est[1]
    assert 24.0 < ratio < 40.0, est


# ------------------------------------------------------ grid/particle operators
def test_ironing_preserves_constants_and_mass():
    f = np.full((8, 9, 10), 2.5)
    g = orc.ironing(f)
    assert np.abs(g - 2.5).max() < 1e-15
    h = np.zeros((12, 12, 12))
    h[4:8, 5:7, 3:9] = np.random.default_rng(0).random((4, 2, 6))
    assert abs(orc.ironing(h).sum() - h.sum()) < 1e-14 * h.sum()


def test_ngp_histogram_counts_and_drops():
    prm = orc.Params(nx=16, ny=16, nz=16, xmax=4.0, ymax=4.0, zmax=4.0, n_par=5)
    # on a grid point, half a cell off, outside the box
    r = np.array([[0.0, 0.24, -4.0, 3.74, 9.0],
                  [0.0, 0.0, -4.0, 3.5, 0.0],
                  [0.0, 0.0, -4.0, 3.5, 0.0]])
    c = orc.ngp_counts(prm, r)
    assert c.sum() == 4
    # x = 0 is grid index 8 (x(i) = -4 + 0.5 (i-1))
    assert c[8, 8, 8] == 2 and c[0, 0, 0] == 1 and c[15, 15, 15] == 1


def test_trilinear_exact_on_affine_fields():
    prm = orc.Params(nx=12, ny=10, nz=14, xmax=3.0, ymax=2.5, zmax=3.5, n_par=50)
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    f = 0.3 + 1.1 * X - 0.7 * Y + 0.25 * Z
    rng = np.random.default_rng(4)
    r = np.stack([rng.uniform(-2.9, 2.4, 50), rng.uniform(-2.4, 1.9, 50), rng.uniform(-3.4, 2.9, 50)])
    a = orc.interpola3(prm, r, f, f, f)
    want = 0.3 + 1.1 * r[0] - 0.7 * r[1] + 0.25 * r[2]
    for c in range(3):
        np.testing.assert_allclose(a[c], want, rtol=0, atol=1e-13)
    # outside the interpolation cube -> zero (potenimp.f90:104-105)
    out = orc.interpola3(prm, np.array([[2.9], [0.0], [0.0]]), f, f, f)
    assert np.all(out == 0.0)


def test_kinetic_energy_of_gaussian():
    """E_kin = h2o2m4 int |grad psi|^2 = h2o2m4 (3/2) pi^(3/2) s for psi = exp(-r^2/2s^2)."""
    s = 1.3
    prm = orc.Params(nx=64, ny=64, nz=64, xmax=8.0, ymax=8.0, zmax=8.0, n_par=1)
    x = prm.axis_grid("x")
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    psi = np.exp(-(X * X + Y * Y + Z * Z) / (2 * s * s)).astype(complex)
    g = [orc.deriv(psi, a, 1, prm.hx) for a in "xyz"]
    ekin = prm.h2o2m4 * sum(np.sum(np.abs(d) ** 2) for d in g) * prm.dxyz
    want = prm.h2o2m4 * 1.5 * math.pi ** 1.5 * s
    assert abs(ekin - want) < 2e-9 * want   # 13-point truncation error ~ (h/s)^12


def test_threaded_oracle_forms_the_same_values():
    """oracle/threaded.py (the multi-core oracle of the size-scale tests and
    the CPU baseline) gives the single-threaded oracle's stencils bit for bit
    and its transforms to rounding."""
    from oracle.threaded import ThreadedOracle
    g = np.random.default_rng(3)
    f = g.standard_normal((16, 20, 24))
    c = f + 1j * g.standard_normal(f.shape)
    plain = [orc.deriv(f, a, k, 0.3) for a in "xyz" for k in (1, 2)] + [orc.deriv(c, "y", 2, 0.4)]
    F = orc.r2c(f)
    back = orc.c2r(F, f.shape, 1.0 / f.size)
    with ThreadedOracle(orc, 3):
        threaded = [orc.deriv(f, a, k, 0.3) for a in "xyz" for k in (1, 2)] + [orc.deriv(c, "y", 2, 0.4)]
        F2 = orc.r2c(f)
        back2 = orc.c2r(F2, f.shape, 1.0 / f.size)
    for x, y in zip(plain, threaded):
        assert np.array_equal(x, y)
    assert np.abs(F2 - F).max() <= 1e-13 * np.abs(F).max()
    assert np.abs(back2 - back).max() <= 1e-14
    assert orc.deriv is not None and orc.r2c.__module__ == orc.__name__
