"""Pins the oracle's 13-point derivative (derivnD, icon=13) against the reference table and mathematics."""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import hetddft_oracle as orc

GOLD = json.loads((Path(__file__).parent / "golden" / "stencil_m13.json").read_text())


def _f(tok):
    return float(tok.replace("D", "e"))


def _ref_matrix(order):
    """13x13 weight matrix as newder(ICON=1) returns it (newder.f:99-117):
    the first 85 entries come from the table, the rest mirror them with
    sign (-1)^order."""
    block = [_f(t) for t in GOLD["K%d" % order]]
    M = 13
    dk = np.zeros(M * M)
    sig = 1.0 - 2.0 * (order % 2)
    for j in range(84):
        dk[j] = block[j]
        dk[M * M - 1 - j] = sig * block[j]
    dk[84] = block[84]
    return dk.reshape(M, M)


def _redef13(cd):
    """Boundary-row rewrite of redef.f:151-170 (icon=13) then icon=12 (:133-150)."""
    cd = cd.copy()
    K, KO2 = 13, 6
    for i in range(KO2, 0, -1):            # rows KO2..1: shift of row i+1 to the left
        for j in range(1, K):
            cd[i - 1, j - 1] = cd[i, j]
        cd[i - 1, K - 1] = 0.0
    for i in range(KO2 + 2, K + 1):        # rows KO2+2..K: shift of row i-1 to the right
        for j in range(K, 1, -1):
            cd[i - 1, j - 1] = cd[i - 2, j - 2]
        cd[i - 1, 0] = 0.0
    return cd


def _deriv1d_table(f, order, h):
    """Deriv1D_p loop structure (init_deriv_parallel.f90:141-159) with the rewritten matrix."""
    cd = _redef13(_ref_matrix(order)) * (1.0 / (h if order == 1 else h * h))
    n = len(f)
    out = np.zeros(n)
    for i in range(6):
        out[i] = sum(cd[i, j] * f[j] for j in range(13))
    for i in range(6, n - 6):
        out[i] = sum(cd[6, j] * f[i - 6 + j] for j in range(13))
    for i in range(n - 6, n):
        out[i] = sum(cd[i - n + 13, j] * f[n - 13 + j] for j in range(13))
    return out


@pytest.mark.parametrize("order", [1, 2])
def test_centre_row_equals_reference_table(order):
    table = _ref_matrix(order)[6]
    exact = [float(c) for c in orc.stencil_rationals(order)]
    assert list(table) == exact


@pytest.mark.parametrize("order", [1, 2])
def test_weights_are_fornberg(order):
    """Centre weights reproduce d^order/dx^order of x^k exactly for k <= 12."""
    w = orc.stencil_rationals(order)
    for k in range(13):
        s = sum(wj * Fraction(m) ** k for wj, m in zip(w, range(-6, 7)))
        want = Fraction(0)
        if k == order:
            want = Fraction(1 if order == 1 else 2)
        assert s == want, (order, k)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_boundary_rows_match_reference_rewrite(order, axis):
    rng = np.random.default_rng(order * 10 + ord(axis))
    f = rng.standard_normal((17, 19, 21))
    h = 0.37
    got = orc.deriv(f, axis, order, h)
    ax = orc._AX[axis]
    want = np.apply_along_axis(_deriv1d_table, ax, f, order, h)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_polynomial_exactness_interior():
    h = 0.25
    x = h * np.arange(40) - 5.0
    f = np.broadcast_to(x ** 7 - 3 * x ** 4 + x, (13, 13, 40)).copy()
    d1 = orc.deriv(f, "x", 1, h)[0, 0, 6:-6]
    d2 = orc.deriv(f, "x", 2, h)[0, 0, 6:-6]
    xi = x[6:-6]
    np.testing.assert_allclose(d1, 7 * xi ** 6 - 12 * xi ** 3 + 1, rtol=1e-8, atol=1e-6)
    np.testing.assert_allclose(d2, 42 * xi ** 5 - 36 * xi ** 2, rtol=1e-8, atol=1e-6)


def test_short_axis_gives_zero():
    """n < npd returns df = 0 (init_deriv_parallel.f90:408-412)."""
    f = np.ones((4, 20, 20))
    assert np.all(orc.deriv(f, "z", 2, 0.5) == 0.0)
