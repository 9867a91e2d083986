"""The oracle's diagnostics against closed forms for a Gaussian wave packet
psi = A exp(-|r - r0|^2 / (2 s^2)) exp(i k x): norm N = A^2 pi^(3/2) s^3,
centre-of-mass moments r0 N, integral of j_x = 2 h2o2m4 k N (transverse
components zero), chemical potential h2o2m4 (k^2 + 3/(2 s^2)) + V0 for a
constant potential V0, and the mean-square radius of a Gaussian denx.

The packet (s = 1.6 A on h = 0.3 A, box +-9.6 A) is resolved well enough that
the 13-point stencils' truncation error stays below 1e-8 (mu) and 1e-7 (j);
the tolerances sit just above that."""
import math

import numpy as np
import pytest


def _packet(orc, n=64, h=0.3, s=1.6, k=0.6, r0=(0.45, -0.3, 0.2)):
    prm = orc.Params(nx=n, ny=n, nz=n, xmax=n * h / 2, ymax=n * h / 2, zmax=n * h / 2, n_par=4)
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    A = 0.13
    psi = A * np.exp(-((X - r0[0]) ** 2 + (Y - r0[1]) ** 2 + (Z - r0[2]) ** 2) / (2 * s * s)) * np.exp(1j * k * X)
    return prm, psi, A, s, k, np.array(r0), (X, Y, Z)


V0 = -7.25
SD = 0.9


def _check_closed_forms(orc):
    """Returns the list of closed forms the oracle misses (empty when all hold)."""
    prm, psi, A, s, k, r0, (X, Y, Z) = _packet(orc)
    denx = np.exp(-(X * X + Y * Y + Z * Z) / (2 * SD * SD))
    rng = np.random.default_rng(2)
    vimp, aimp = rng.standard_normal((3, 4)), rng.standard_normal((3, 4))
    out = orc.diagnostics(prm, psi, np.full(psi.shape, V0), denx, vimp, aimp)
    N = A * A * math.pi ** 1.5 * s ** 3
    bad = []
    if not abs(out[0] - N) < 1e-13 * N:
        bad.append("norm")
    if not np.allclose(out[1:4], r0 * N, rtol=0, atol=1e-13 * N):
        bad.append("cm")
    mu = prm.h2o2m4 * (k * k + 1.5 / (s * s)) + V0
    if not abs(out[4] - mu) < 1e-8 * abs(mu):
        bad.append("mu")
    if not abs(out[5] - 2 * prm.h2o2m4 * k * N) < 1e-7 * N:
        bad.append("jtot")
    if not np.abs(out[6:8]).max() < 1e-15 * N:
        bad.append("jtot transverse")
    msx = math.sqrt(SD * SD * (2 * math.pi * SD * SD) ** 1.5)
    if not np.allclose(out[8:11], msx, rtol=1e-12):
        bad.append("msx")
    if not (np.array_equal(out[11:17:2], vimp.min(axis=1)) and np.array_equal(out[12:17:2], vimp.max(axis=1))
            and np.array_equal(out[17:23:2], aimp.min(axis=1)) and np.array_equal(out[18:23:2], aimp.max(axis=1))):
        bad.append("minmax")
    jx, jy, jz = orc.current(prm, psi)
    exact = 2 * prm.h2o2m4 * k * np.abs(psi) ** 2
    inner = (slice(6, -6),) * 3
    if not (np.abs(jx - exact)[inner].max() < 3e-7 * exact.max() and np.abs(jy).max() < 1e-15 * exact.max()
            and np.abs(jz).max() < 1e-15 * exact.max()):
        bad.append("current")
    return bad


def test_diagnostics_closed_forms():
    from oracle import hetddft_oracle as orc
    assert _check_closed_forms(orc) == []


@pytest.mark.parametrize("plant", ["drop_lap_z", "swap_axis", "flip_d1_sign"])
def test_planted_mistake_is_caught(plant, monkeypatch):
    """Plausible mistakes planted in the oracle's derivative calls fail the
    closed forms: a Laplacian without its z term, a current built from the
    wrong axis, a first derivative of the wrong sign."""
    from oracle import hetddft_oracle as orc
    real = orc.deriv

    def planted(f, axis, order, h):
        if plant == "drop_lap_z" and order == 2 and axis == "z":
            return np.zeros_like(f)
        if plant == "swap_axis" and order == 1:
            axis = {"x": "y", "y": "x", "z": "z"}[axis]
        if plant == "flip_d1_sign" and order == 1:
            return -real(f, axis, order, h)
        return real(f, axis, order, h)

    monkeypatch.setattr(orc, "deriv", planted)
    assert _check_closed_forms(orc) != []
