"""k-space kernels: oracle pinned by quadrature/closed forms; library generator vs oracle (CPU)."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import hetddft_oracle as orc

PRM = orc.Params(nx=24, ny=24, nz=24, xmax=7.2, ymax=7.2, zmax=7.2, n_par=8)


def _radial_ft(f, p, r0, r1):
    """(4 pi / q) int_r0^r1 r f(r) sin(q r) dr, q = 2 pi p (3D Fourier transform of f(|r|))."""
    q = 2 * math.pi * p
    if np.isinf(r1):
        val, _ = quad(lambda r: r * f(r), r0, np.inf, weight="sin", wvar=q, limlst=200)
    else:
        val, _ = quad(lambda r: r * f(r) * math.sin(q * r), r0, r1, limit=400, epsabs=0, epsrel=1e-13)
    return 4 * math.pi / q * val


@pytest.mark.parametrize("p", [0.05, 0.3, 0.9, 1.7])
def test_wcgk_is_ft_of_normalised_ball(p):
    h = PRM.h4
    vol = 4.0 / 3.0 * math.pi * h ** 3
    want = _radial_ft(lambda r: 1.0 / vol, p, 0.0, h)
    x = 2 * math.pi * p * h
    got = 3 * (math.sin(x) - x * math.cos(x)) / x ** 3
    assert abs(got - want) < 1e-11


@pytest.mark.parametrize("p", [0.05, 0.3, 0.9])
def test_kalfs_is_ft_of_normalised_gaussian(p):
    l = PRM.l
    want = _radial_ft(lambda r: math.exp(-(r / l) ** 2) / (math.pi ** 1.5 * l ** 3), p, 0.0, 12.0)
    got = math.exp(-(math.pi * l * p) ** 2)
    assert abs(got - want) < 1e-11


def _lj_core(r):
    s = PRM.sigma4 / r
    return 4 * PRM.eps4 * (s ** 12 - s ** 6)


@pytest.mark.parametrize("p", [0.08, 0.4, 1.1])
def test_fvlj4_is_ft_of_cored_lj(p):
    """OT core: V = 0 inside h4 (fforma.f90:50-53)."""
    want = _radial_ft(_lj_core, p, PRM.h4, np.inf)
    q = 2 * math.pi * p
    got = 8 * PRM.eps4 * PRM.sigma4 ** 6 * (PRM.sigma4 ** 6 * orc._sinx11(q, PRM.h4)
                                            - orc._sinx5(q, PRM.h4)) / p
    assert abs(got - want) < 1e-7 * max(1.0, abs(want))


def test_fvlj4_at_zero_is_volume_integral():
    want, _ = quad(lambda r: 4 * math.pi * r * r * _lj_core(r), PRM.h4, np.inf, epsrel=1e-13)
    assert abs(orc._bforce_ot(PRM.eps4, PRM.sigma4, PRM.h4) - want) < 1e-9 * abs(want)


def _morse_ft_quad(pot, p):
    inner = _radial_ft(lambda r: pot.umax, p, 0.0, pot.rc)
    outer = _radial_ft(lambda r: float(pot.v_raw(np.array([r]))[0]), p, pot.rc, 60.0)
    return inner + outer


@pytest.fixture(scope="module")
def ftv():
    pot = orc.Morse()
    pmaxx = 1.0 / (2.0 * PRM.hx)
    qmax = math.sqrt(3) * pmaxx
    return pot, orc.FTVSpline(pot, qmax, qmax / 1499)


@pytest.mark.parametrize("p", [0.0, 0.02, 0.1, 0.35, 0.8, 1.4])
def test_spline_ft_matches_analytic_morse_transform(ftv, p):
    """FT_V_spline of the cored Morse potential vs direct quadrature.

    The reference procedure is accurate to about 2.5e-7 of V(0): its clamped
    spline takes the end second derivative from a 0.5 A finite difference
    (FT_V_spline.f90:88-101), which is coarse next to the steep core."""
    pot, ft = ftv
    if p == 0.0:
        want, _ = quad(lambda r: 4 * math.pi * r * r * float(pot(np.array([r]))[0]), 0.0, 60.0,
                       points=[pot.rc], limit=400, epsrel=1e-13)
    else:
        want = _morse_ft_quad(pot, p)
    scale = abs(ft(0.0))
    assert abs(ft(p) - want) < 4e-7 * scale, (p, ft(p), want)


def test_library_generator_matches_oracle(hip):
    """hetddft_make_kernels (host C++) against the oracle's restatement."""
    p = hip.default_params(nx=PRM.nx, ny=PRM.ny, nz=PRM.nz, xmax=PRM.xmax, ymax=PRM.ymax,
                           zmax=PRM.zmax, n_par=PRM.n_par)
    k = hip.make_kernels(p)
    np.testing.assert_allclose(k["wcgk"], orc.kernel_wcgk(PRM), rtol=0, atol=1e-15)
    np.testing.assert_allclose(k["kalfs"], orc.kernel_kalfs(PRM), rtol=0, atol=1e-15)
    f = orc.kernel_fvlj4(PRM)
    # sinx11 cancels ~8 digits at the largest |p|; both sides carry it
    np.testing.assert_allclose(k["fvlj4"], f, rtol=0, atol=1e-8 * np.abs(f).max())
    vq = orc.kernel_vq(PRM)
    diff = np.abs(k["vq"] - vq)
    ulp = np.spacing(np.abs(vq).astype(np.float32)).astype(np.float64)
    assert np.all(diff <= ulp)
