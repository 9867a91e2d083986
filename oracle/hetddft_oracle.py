"""CPU FP64 oracle of the He-4 Orsay-Trento TDDFT + Bohmian-impurity time step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library: it
may be imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, never by the product path.

It restates, in numpy, the per-step algorithm of the reference Fortran solver
(bcntls2016/4hetddft-isotropic-bohmian).  Every function cites the reference
file:line it follows.  Arrays use numpy C order with shape ``(nz, ny, nx)``,
which is the Fortran ``(nx, ny, nz)`` column-major layout of the reference
byte for byte; half spectra have shape ``(nz, ny, nx//2+1)``; particle arrays
have shape ``(3, N_par)``, the Fortran ``rimp(N_par,3)`` layout.

How each part is pinned (tests/test_oracle_*.py):
  * stencil weights: exact Fornberg rationals == the reference table
    ``mates.f:827-832,874-879`` (golden file tests/golden/stencil_m13.json);
    exactness on polynomials of degree <= 12; the zero-padding rule.
  * FFT layer: numpy.fft (pocketfft) against a direct O(n^2) DFT and the
    r2c/c2r round trip with the 1/(nx ny nz) factor of ``fft.f90:125-207``.
  * ironing / NGP histogram / trilinear gather: conservation and exactness
    invariants (constants preserved, affine fields reproduced, counts exact).
  * RK-Gill and Hamming PMC bookkeeping: the linear test ODE y' = lam*y in
    exact rational arithmetic (fractions), for psi and both particle slots.
  * energy: closed form for a Gaussian and conservation checks.
  * k-space kernels: closed forms (wcgk, kalfs), quadrature of the cored
    Lennard-Jones Fourier integral (fvlj4), analytic Fourier transform of
    the cored Morse potential (Vq).
  * droplet start (droplet_state): the vectorised Philox against the scalar
    one (itself equal to numpy's Philox bit for bit); the Woods-Saxon
    density with ripple against the workload generator's scipy-based one;
    the Box-Muller normals' moments; sum(|psi|^2) dxyz = N.
  * composite trajectory (Veval ordering inside the PC step): parity
    unpinned against a reference run -- the reference needs FFTW/MKL, which
    this image lacks, so it cannot be built here; it is pinned only through
    the per-component checks above.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

# --------------------------------------------------------------------------
# Constants (modules.f90:122-141 he4, :201-213 lenard4, :314 classicimp)
# --------------------------------------------------------------------------
PI = 4.0 * math.atan(1.0)                   # DFT4He3d.f90:170
TWOPI = 2.0 * PI
MP_U = 0.0206149591055                      # modules.f90:314
KTOPS = 7.63822291                          # DFT4He3d.f90:110 (ps <-> K^-1)
PSTOK = 1.0 / 7.63822291


@dataclass
class Params:
    """Grid, functional and impurity parameters of one instance.

    Defaults are the Orsay-Trento values of modules.f90:122-141 and
    :201-213, with the production settings npd=13, icon=13, core4='OTC'.
    """
    nx: int = 32
    ny: int = 32
    nz: int = 32
    xmax: float = 8.0
    ymax: float = 8.0
    zmax: float = 8.0
    n_par: int = 256
    mimpur: float = 40.0
    ironx: int = 3
    core: str = "OTC"
    lambdah: float = 0.0
    tzmean: float = 0.0
    tzsurf: float = 1.0
    cp4: float = -2.41186e4
    cpp4: float = 1.85850e6
    alphas: float = 54.31
    l: float = 1.0
    den0s: float = 0.04
    h2o2m4: float = 6.05969638298
    C: float = 3.1577504e4
    beta: float = 40.0
    den_m: float = 0.37
    h4: float = 2.190323                    # h4ot, DFT4He3d.f90:474-481
    eps4: float = 10.22
    sigma4: float = 2.556

    # derived quantities (DFT4He3d.f90:271-284, 204, 429)
    @property
    def hx(self):
        return 2.0 * abs(self.xmax) / self.nx

    @property
    def hy(self):
        return 2.0 * abs(self.ymax) / self.ny

    @property
    def hz(self):
        return 2.0 * abs(self.zmax) / self.nz

    @property
    def dxyz(self):
        return self.hx * self.hy * self.hz

    @property
    def nxyz(self):
        return self.nx * self.ny * self.nz

    @property
    def renor(self):
        return 1.0 / (self.nx * self.ny * self.nz)   # fft.f90:29

    @property
    def m_imp(self):
        return self.mimpur * MP_U                    # mAg_u, DFT4He3d.f90:204

    @property
    def h2o2mx(self):
        return self.h2o2m4 * (4.002602 / self.mimpur)  # DFT4He3d.f90:429

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def hshape(self):
        return (self.nz, self.ny, self.nx // 2 + 1)

    def axis_grid(self, which):
        """x(i) = -xmax + hx*(i-1)  (DFT4He3d.f90:302-310)."""
        n, xm, h = {"x": (self.nx, self.xmax, self.hx),
                    "y": (self.ny, self.ymax, self.hy),
                    "z": (self.nz, self.zmax, self.hz)}[which]
        return -xm + h * np.arange(n, dtype=np.float64)


# --------------------------------------------------------------------------
# Finite-difference weights (newder.f:99-129, redef.f:151-170, mates.f)
# --------------------------------------------------------------------------
# Centre rows of the 13-point Bickley/Fornberg table for offsets -6..+6.
_D1 = [Fraction(6, 7), Fraction(-15, 56), Fraction(5, 63), Fraction(-1, 56),
       Fraction(1, 385), Fraction(-1, 5544)]
_D2 = [Fraction(12, 7), Fraction(-15, 56), Fraction(10, 189), Fraction(-1, 112),
       Fraction(2, 1925), Fraction(-1, 16632)]
_D2C = Fraction(-5369, 1800)


def stencil_rationals(order):
    """Exact centre-row weights w[m+6], m=-6..6, of the 13-point derivative."""
    if order == 1:
        return [-c for c in reversed(_D1)] + [Fraction(0)] + list(_D1)
    if order == 2:
        return list(reversed(_D2)) + [_D2C] + list(_D2)
    raise ValueError(order)


def stencil_weights(order, h):
    """Weights scaled by 1/h**order as Deriv3D_p does (init_deriv_parallel.f90:415-428)."""
    sto = 1.0 / (h if order == 1 else h * h)
    return [float(c) * sto for c in stencil_rationals(order)]


_AX = {"x": 2, "y": 1, "z": 0}


def deriv(f, axis, order, h):
    """derivnD(order, nn, h, iv, f, df, icon=13) for 3D real or complex arrays.

    icon=13 makes every boundary row the centre stencil with f == 0 outside
    the box (redef.f:121-170 shifts the centre row; the Deriv3D loops at
    init_deriv_parallel.f90:428-535 then apply rows 1..6 and n-5..n).
    The sum runs over the 13 points in index order, as ``Sum(Cd*f)`` does.
    """
    w = stencil_weights(order, h)
    ax = _AX[axis]
    n = f.shape[ax]
    if n < 13:                       # init_deriv_parallel.f90:408-412
        return np.zeros_like(f)
    pad = [(0, 0)] * 3
    pad[ax] = (6, 6)
    fp = np.pad(f, pad)
    out = np.zeros_like(f)
    sl = [slice(None)] * 3
    for j in range(13):
        sl[ax] = slice(j, j + n)
        out = out + w[j] * fp[tuple(sl)]
    return out


# --------------------------------------------------------------------------
# FFT layer (fft.f90:5-207).  FFTW r2c is unnormalised; every fftbk_*
# multiplies the c2r output by renor = 1/(nx ny nz).
# --------------------------------------------------------------------------
def r2c(f):
    return np.fft.rfftn(f)


def c2r(F, shape, renor):
    return np.fft.irfftn(F, s=shape, axes=(0, 1, 2), norm="forward") * renor


# --------------------------------------------------------------------------
# k-space grid and kernels (DFT4He3d.f90:312-354, 517-570; fforma.f90;
# initcg.f90; FT_V_spline.f90)
# --------------------------------------------------------------------------
def pgrid_1d(n, h):
    """FFTW-ordered frequencies with +pmax at n/2 (DFT4He3d.f90:321-339)."""
    hp = 1.0 / (n * h)
    pmax = 1.0 / (2.0 * h)
    p = np.empty(n)
    for i in range(n):
        if i <= n // 2:
            p[i] = hp * i
        else:
            p[i] = -pmax + hp * (i - n // 2)
    return p


def pmod(prm: Params):
    """|p| on the half grid (DFT4He3d.f90:343-351)."""
    px = pgrid_1d(prm.nx, prm.hx)[: prm.nx // 2 + 1]
    py = pgrid_1d(prm.ny, prm.hy)
    pz = pgrid_1d(prm.nz, prm.hz)
    px2 = (px * px)[None, None, :]
    py2 = (py * py)[None, :, None]
    pz2 = (pz * pz)[:, None, None]
    return np.sqrt(px2 + py2 + pz2)


def kernel_wcgk(prm: Params):
    """Coarse-graining kernel 3(sin x - x cos x)/x^3, x = 2 pi p h4 (initcg.f90:31-46)."""
    p = pmod(prm)
    out = np.ones_like(p)
    m = p > 0.0
    arg = TWOPI * p[m] * prm.h4
    aux = np.sin(arg) - arg * np.cos(arg)
    val = 3.0 * aux / (arg * arg * arg)
    val = np.where(np.abs(aux) < 1e-10, 1.0, val)
    out[m] = val
    return out


def kernel_kalfs(prm: Params):
    """Gaussian alpha_s kernel exp(-(pi l p)^2) (DFT4He3d.f90:566-570)."""
    p = pmod(prm)
    t = PI * prm.l * p
    return np.exp(-(t * t))


def _si(x):
    from scipy.special import sici
    return sici(x)[0]


def _sinint(a, x):
    """pi/2 - Si(a x) for a > 0 (fforma.f90:158-205)."""
    return PI * 0.5 - _si(a * x)


def _sinx11(a, x):
    """int_x^inf sin(a r)/r^11 dr in closed form (fforma.f90:73-100)."""
    y = a * x
    y2 = y * y
    dot_b = 40320.0 - 720.0 * y2 + 24.0 * y2 ** 2 - 2.0 * y2 ** 3 + 1.0 * y2 ** 4
    dot_c = 362880.0 - 5040.0 * y2 + 120.0 * y2 ** 2 - 6.0 * y2 ** 3 + 1.0 * y2 ** 4
    res1 = (a * dot_b * np.cos(y) + dot_c * np.sin(y) / x) / (3628800 * x ** 9)
    res2 = -(a ** 10 * _sinint(a, x)) / 3628800.0
    return res1 + res2


def _sinx5(a, x):
    """int_x^inf sin(a r)/r^5 dr in closed form (fforma.f90:106-133)."""
    y = a * x
    y2 = y * y
    dot_b = -2.0 + 1.0 * y2
    dot_c = -6.0 + 1.0 * y2
    res1 = -(a * dot_b * np.cos(y) + dot_c * np.sin(y) / x) / (24 * x ** 3)
    res2 = (a ** 4 * _sinint(a, x)) / 24.0
    return res1 + res2


def _bforce_ot(eps, sigma, hc):
    """p = 0 value of the OT-cored LJ transform (fforma.f90:256-280)."""
    pi4 = 4.0 * PI
    eps4 = 4.0 * eps
    s6 = sigma ** 6
    s12 = s6 * s6
    bb2 = pi4 * eps4 * s12 / (9.0 * hc ** 9)
    bb3 = -pi4 * eps4 * s6 / (3.0 * hc ** 3)
    return bb2 + bb3


def kernel_fvlj4(prm: Params):
    """Fourier transform of the OT-cored LJ potential (fforma.f90:11-61).

    ``core4='OTC'`` is passed into a CHARACTER(len=2) dummy, so the 'OT'
    branch applies (fforma.f90:14,50-53).
    """
    p = pmod(prm)
    out = np.empty_like(p)
    small = np.abs(p) < 1e-10
    out[small] = _bforce_ot(prm.eps4, prm.sigma4, prm.h4)
    pp = p[~small]
    aux1 = 8.0 * prm.eps4
    sigma6 = prm.sigma4 ** 6
    aux2 = TWOPI * pp
    t1 = _sinx11(aux2, prm.h4)
    t2 = _sinx5(aux2, prm.h4)
    out[~small] = (aux1 * sigma6 * (sigma6 * t1 - t2)) / pp
    return out


# ---- cored Morse He-impurity pair potential and FT_V_spline -------------
@dataclass
class Morse:
    """V(r) = D[e^{-2a(r-re)} - 2 e^{-a(r-re)}], V = umax for r < rc.

    Same functional form as the reference's surface Morse (V_impur.f90:2455,
    morse.f90:15); the Select_Pot core convention (V_impur.f90:179-189) sets
    V = umax below r_cutoff.
    """
    D: float = 20.0
    re: float = 5.0
    alpha: float = 0.8
    rc: float = 2.0
    umax: float | None = None

    def __post_init__(self):
        if self.umax is None:
            self.umax = float(self.v_raw(np.array([self.rc]))[0])

    def v_raw(self, r):
        e = np.exp(-self.alpha * (r - self.re))
        return self.D * (e * e - 2.0 * e)

    def __call__(self, r):
        r = np.asarray(r, dtype=np.float64)
        return np.where(r < self.rc, self.umax, self.v_raw(r))


def spline_coeffs(x, y, s1, sn):
    """Maron cubic spline in global-polynomial form (mates.f:36-116)."""
    n = len(x)
    a = np.zeros(n)
    b = np.zeros(n)
    c = np.zeros(n)
    d = np.zeros(n)
    n1 = n - 1
    n2 = n - 2
    a[:n1] = x[1:] - x[:-1]
    d[:n1] = (y[1:] - y[:-1]) / a[:n1]
    # loop i=1..n2 with k = n1-i+1 (1-based): d(k) = 6(d(k)-d(k-1)), descending k
    for i in range(1, n2 + 1):
        b[i - 1] = 2.0 * (a[i - 1] + a[i])
        k = n1 - i + 1
        d[k - 1] = 6.0 * (d[k - 1] - d[k - 2])
    d[1] = d[1] - a[0] * s1
    d[n1 - 1] = d[n1 - 1] - a[n1 - 1] * sn
    for i in range(2, n2 + 1):
        r = a[i - 1] / b[i - 2]
        b[i - 1] = b[i - 1] - r * a[i - 1]
        d[i] = d[i] - r * d[i - 1]
    d[n1 - 1] = d[n1 - 1] / b[n2 - 1]
    for i in range(2, n2 + 1):
        k = n1 - i + 1
        d[k - 1] = (d[k - 1] - a[k - 1] * d[k]) / b[k - 2]
    d[n - 1] = sn
    A = np.zeros(n)
    B = np.zeros(n)
    Cc = np.zeros(n)
    Dd = np.zeros(n)
    si_next = s1
    for i in range(n1):
        si = si_next
        si_next = d[i + 1]
        h = a[i]
        hi = 1.0 / h
        xi, xi1 = x[i], x[i + 1]
        A[i] = ((hi / 6.0) * (si * xi1 ** 3 - si_next * xi ** 3)
                + hi * (y[i] * xi1 - y[i + 1] * xi)
                + (h / 6.0) * (si_next * xi - si * xi1))
        B[i] = ((hi / 2.0) * (si_next * xi ** 2 - si * xi1 ** 2)
                + hi * (y[i + 1] - y[i]) + (h / 6.0) * (si - si_next))
        Cc[i] = (hi / 2.0) * (si * xi1 - si_next * xi)
        Dd[i] = (hi / 6.0) * (si_next - si)
    return A, B, Cc, Dd


def spline_coeffs_fast(x, y, s1, sn):
    """Vectorised equivalent of spline_coeffs for long node lists."""
    n = len(x)
    n1 = n - 1
    n2 = n - 2
    a = np.zeros(n)
    d = np.zeros(n)
    a[:n1] = x[1:] - x[:-1]
    d[:n1] = (y[1:] - y[:-1]) / a[:n1]
    b = np.zeros(n)
    b[:n2] = 2.0 * (a[:n2] + a[1:n1])
    dd = d.copy()
    # d(k) = 6(d(k) - d(k-1)) for k = n1..2 (1-based) uses the old d(k-1)
    dd[1:n1] = 6.0 * (d[1:n1] - d[0:n1 - 1])
    d = dd
    d[1] = d[1] - a[0] * s1
    d[n1 - 1] = d[n1 - 1] - a[n1 - 1] * sn
    for i in range(2, n2 + 1):
        r = a[i - 1] / b[i - 2]
        b[i - 1] = b[i - 1] - r * a[i - 1]
        d[i] = d[i] - r * d[i - 1]
    d[n1 - 1] = d[n1 - 1] / b[n2 - 1]
    for i in range(2, n2 + 1):
        k = n1 - i + 1
        d[k - 1] = (d[k - 1] - a[k - 1] * d[k]) / b[k - 2]
    d[n - 1] = sn
    si = np.empty(n1)
    si[0] = s1
    si[1:] = d[1:n1]
    si1 = d[1:n]
    h = a[:n1]
    hi = 1.0 / h
    xi = x[:n1]
    xi1 = x[1:n]
    yi = y[:n1]
    yi1 = y[1:n]
    A = np.zeros(n)
    B = np.zeros(n)
    Cc = np.zeros(n)
    Dd = np.zeros(n)
    A[:n1] = ((hi / 6.0) * (si * xi1 ** 3 - si1 * xi ** 3)
              + hi * (yi * xi1 - yi1 * xi)
              + (h / 6.0) * (si1 * xi - si * xi1))
    B[:n1] = ((hi / 2.0) * (si1 * xi ** 2 - si * xi1 ** 2)
              + hi * (yi1 - yi) + (h / 6.0) * (si - si1))
    Cc[:n1] = (hi / 2.0) * (si * xi1 - si1 * xi)
    Dd[:n1] = (hi / 6.0) * (si1 - si)
    return A, B, Cc, Dd


def findi(xs, xc):
    """Bisection interval search (mates.f:4-34), returns a 0-based index."""
    n = len(xs)
    if xc > xs[n - 1]:
        return n - 2
    if xc < xs[0]:
        return 0
    i, i1 = 1, n
    while True:
        it = (i + i1) // 2
        if xc > xs[it - 1]:
            i = it
        if xc <= xs[it - 1]:
            i1 = it
        if i1 - i <= 1:
            break
    return i - 1


class FTVSpline:
    """Spline-based Fourier transform of a radial potential (FT_V_spline.f90:14-276)."""

    N = 100001
    R_MAX = 2000.0
    EPS = 1e-8
    SMALL_Q = 1e-2
    DR = 0.5

    def __init__(self, pot, pmax, hp):
        self.pot = pot
        rc, fc = pot.rc, pot.umax
        N = self.N
        X1 = rc + self.EPS
        XN = self.R_MAX
        dr = self.DR
        r0 = X1
        r1 = r0 + dr
        r2 = r1 + dr
        f0, f1, f2 = (float(pot(np.array([v]))[0]) for v in (r0, r1, r2))
        S1 = (f0 - 2.0 * f1 + f2) / dr ** 2
        fm1, f0n, fp1 = (float(pot(np.array([v]))[0]) for v in (XN - dr, XN, XN + dr))
        Sn = (fm1 - 2.0 * f0n + fp1) / dr ** 2
        h = (XN - X1) / (N - 1)
        x = np.arange(N, dtype=np.float64) * h + X1
        y = pot(x)
        A, B, C, D = spline_coeffs_fast(x, y, S1, Sn)
        Qpi = 4.0 * PI
        n_p = int(pmax / hp + 1.5)
        P = np.arange(n_p, dtype=np.float64) * hp
        Vpp = np.empty(n_p)
        xj, xj1 = x[:-1], x[1:]
        Aj, Bj, Cj, Dj = A[:-1], B[:-1], C[:-1], D[:-1]
        for i in range(n_p):
            p = P[i]
            q = 2.0 * PI * p
            if p == 0.0:
                vp = rc ** 3 * fc / 3.0
                terms = (Aj / 3.0 * (xj1 ** 3 - xj ** 3) + Bj / 4.0 * (xj1 ** 4 - xj ** 4)
                         + Cj / 5.0 * (xj1 ** 5 - xj ** 5) + Dj / 6.0 * (xj1 ** 6 - xj ** 6))
                Fi = np.cumsum(terms)[-1]
                Vpp[i] = (vp + Fi) * Qpi
                continue
            c1 = 1.0 / q
            c2 = c1 * c1
            c3 = c2 * c1
            c4 = c2 * c2
            c5 = c4 * c1
            vp = fc * c1 * (math.sin(q * rc) * c1 - rc * math.cos(q * rc))
            sj, cj = np.sin(q * xj), np.cos(q * xj)
            sj1, cj1 = np.sin(q * xj1), np.cos(q * xj1)
            rj, rj1 = xj, xj1
            terms = ((Aj + (Bj + (Cj + Dj * rj) * rj) * rj) * rj * c1 * cj
                     - (Aj + (Bj + (Cj + Dj * rj1) * rj1) * rj1) * rj1 * c1 * cj1
                     + (Aj + (2.0 * Bj + (3.0 * Cj + 4.0 * Dj * rj1) * rj1) * rj1) * c2 * sj1
                     - (Aj + (2.0 * Bj + (3.0 * Cj + 4.0 * Dj * rj) * rj) * rj) * c2 * sj
                     + (2.0 * Bj + (6.0 * Cj + 12.0 * Dj * rj1) * rj1) * c3 * cj1
                     - (2.0 * Bj + (6.0 * Cj + 12.0 * Dj * rj) * rj) * c3 * cj
                     + (6.0 * Cj + 24.0 * Dj * rj) * c4 * sj
                     - (6.0 * Cj + 24.0 * Dj * rj1) * c4 * sj1
                     + 24.0 * Dj * c5 * (cj - cj1))
            Fi = np.cumsum(terms)[-1]
            Vpp[i] = (vp + Fi) * c1 * Qpi
        # small-q redefinition (FT_V_spline.f90:230-264)
        xs = [0.0]
        ys = [Vpp[0]]
        for i in range(1, n_p):
            if P[i] > self.SMALL_Q:
                xs.append(P[i])
                ys.append(Vpp[i])
            if len(xs) == 4:
                break
        a4, b4, c4_, d4 = spline_coeffs(np.array(xs), np.array(ys), 0.0, 0.0)
        for i in range(1, n_p):
            p = P[i]
            if p <= self.SMALL_Q:
                Vpp[i] = a4[0] + p * (b4[0] + p * (c4_[0] + p * d4[0]))
            else:
                break
        self.P = P
        self.Vpp = Vpp
        self.A, self.B, self.C, self.D = spline_coeffs_fast(P, Vpp, 0.0, 0.0)

    def __call__(self, p):
        k = findi(self.P, p)
        return self.A[k] + p * (self.B[k] + p * (self.C[k] + p * self.D[k]))


def kernel_vq(prm: Params, pot=None, nq=1500, round_fp32=True):
    """Vq(p) = cmplx(FT_V_spline(|p|,...)) (DFT4He3d.f90:527-536).

    ``cmplx`` without KIND rounds to default REAL(4) (DFT4He3d.f90:533), so
    the stored value is the FP32 rounding of the spline transform.
    """
    pot = pot or Morse()
    pmaxx, pmaxy, pmaxz = (1.0 / (2.0 * h) for h in (prm.hx, prm.hy, prm.hz))
    qmax = math.sqrt(pmaxx ** 2 + pmaxy ** 2 + pmaxz ** 2)
    hq = qmax / (nq - 1)
    ft = FTVSpline(pot, qmax, hq)
    pm = pmod(prm)
    flat = pm.ravel()
    uniq, inv = np.unique(flat, return_inverse=True)
    vals = np.array([ft(float(p)) for p in uniq])
    out = vals[inv].reshape(pm.shape)
    if round_fp32:
        out = out.astype(np.float32).astype(np.float64)
    return out


def make_kernels(prm: Params, pot=None, nq=1500):
    return dict(fvlj4=kernel_fvlj4(prm), wcgk=kernel_wcgk(prm),
                kalfs=kernel_kalfs(prm), vq=kernel_vq(prm, pot, nq))


# --------------------------------------------------------------------------
# Per-step physics (potenimp.f90, poten.f90, term_alfa.f90, denpart.f90,
# ironing.f90, derden.f90, energy.f90, steprk.f90, steppc.f90)
# --------------------------------------------------------------------------
IRON_A0 = 8.333333333333333e-2    # ironing.f90:12
IRON_A1 = 0.5                     # ironing.f90:13


def ironing(f):
    """One 7-point smoothing pass on the interior points (ironing.f90:16-28)."""
    aux = f.copy()
    out = f.copy()
    c = aux[1:-1, 1:-1, 1:-1]
    s = (aux[1:-1, 1:-1, :-2] + aux[1:-1, 1:-1, 2:]) + aux[1:-1, :-2, 1:-1]
    s = s + aux[1:-1, 2:, 1:-1]
    s = s + aux[:-2, 1:-1, 1:-1]
    s = s + aux[2:, 1:-1, 1:-1]
    out[1:-1, 1:-1, 1:-1] = IRON_A0 * s + IRON_A1 * c
    return out


def ngp_counts(prm: Params, rimp):
    """Nearest-grid-point counts (denpart.f90:24-32): ix = int(1.5+(r-x1)/h)."""
    xmin, ymin, zmin = -prm.xmax, -prm.ymax, -prm.zmax
    ix = np.trunc(1.5 + (rimp[0] - xmin) / prm.hx).astype(np.int64)
    iy = np.trunc(1.5 + (rimp[1] - ymin) / prm.hy).astype(np.int64)
    iz = np.trunc(1.5 + (rimp[2] - zmin) / prm.hz).astype(np.int64)
    ok = ((iz >= 1) & (iz <= prm.nz) & (iy >= 1) & (iy <= prm.ny)
          & (ix >= 1) & (ix <= prm.nx))
    cnt = np.zeros(prm.shape, dtype=np.float64)
    np.add.at(cnt, (iz[ok] - 1, iy[ok] - 1, ix[ok] - 1), 1.0)
    return cnt


def blend_103(r, s, t, x000, x001, x010, x011, x100, x101, x110, x111):
    """Trilinear blend in the reference's polynomial form (potenimp.f90:278-286)."""
    return (1.0 * (+x000)
            + r * (-x000 + x100)
            + s * (-x000 + x010)
            + t * (-x000 + x001)
            + r * s * (+x000 - x100 - x010 + x110)
            + r * t * (+x000 - x100 - x001 + x101)
            + s * t * (+x000 - x010 - x001 + x011)
            + r * s * t * (-x000 + x100 + x010 + x001 - x011 - x101 - x110 + x111))


def interpola3(prm: Params, rimp, fx, fy, fz):
    """Trilinear gather at the particles; zero outside [1, n-1] (potenimp.f90:76-116)."""
    x0, y0, z0 = -prm.xmax, -prm.ymax, -prm.zmax
    ax = (rimp[0] - x0) / prm.hx + 1
    ay = (rimp[1] - y0) / prm.hy + 1
    az = (rimp[2] - z0) / prm.hz + 1
    ix = np.trunc(ax).astype(np.int64)
    iy = np.trunc(ay).astype(np.int64)
    iz = np.trunc(az).astype(np.int64)
    r = ax - ix
    s = ay - iy
    t = az - iz
    bad = ((ix >= prm.nx) | (iy >= prm.ny) | (iz >= prm.nz)
           | (ix <= 0) | (iy <= 0) | (iz <= 0))
    out = np.zeros((3, rimp.shape[1]))
    ok = ~bad
    i0, j0, k0 = ix[ok] - 1, iy[ok] - 1, iz[ok] - 1
    rr, ss, tt = r[ok], s[ok], t[ok]
    for c, f in enumerate((fx, fy, fz)):
        g = lambda dz, dy, dx: f[k0 + dz, j0 + dy, i0 + dx]
        out[c, ok] = blend_103(rr, ss, tt, g(0, 0, 0), g(1, 0, 0), g(0, 1, 0), g(1, 1, 0),
                               g(0, 0, 1), g(1, 0, 1), g(0, 1, 1), g(1, 1, 1))
    return out


# --------------------------------------------------------------------------
# Particle sampling (sorteo.f90:1-37 with interpola3denx :40-74)
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1
_PHILOX_M = (0xD2E7470EE14C6C93, 0xCA5A826395121157)
_PHILOX_W = (0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B)


def philox4x64(ctr, key, rounds=10):
    """Philox4x64-10 block (Salmon, Moraes, Dror, Shaw, SC11): four 64-bit
    words from a 256-bit counter and a 128-bit key."""
    c = [int(v) & _M64 for v in ctr]
    k = [int(v) & _M64 for v in key]
    for r in range(rounds):
        if r > 0:
            k = [(k[0] + _PHILOX_W[0]) & _M64, (k[1] + _PHILOX_W[1]) & _M64]
        p0 = _PHILOX_M[0] * c[0]
        p1 = _PHILOX_M[1] * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k[0], p1 & _M64, (p0 >> 64) ^ c[3] ^ k[1], p0 & _M64]
    return c


def _u53(v):
    return (v >> 11) * (1.0 / 9007199254740992.0)


def interpola3denx(prm: Params, denx, xyz):
    """Trilinear interpolation of denx at a point (sorteo.f90:40-74): index
    int((r - x1)/h + 1), zero outside [1, n-1], blend_103 weights."""
    xg, yg, zg = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    # x0 = x(1), Dx = x(2) - x(1) from the grid arrays (sorteo.f90:10-12)
    ax = (xyz[0] - xg[0]) / (xg[1] - xg[0]) + 1.0
    ay = (xyz[1] - yg[0]) / (yg[1] - yg[0]) + 1.0
    az = (xyz[2] - zg[0]) / (zg[1] - zg[0]) + 1.0
    ix, iy, iz = int(ax), int(ay), int(az)
    r, s, t = ax - ix, ay - iy, az - iz
    if ix >= prm.nx or iy >= prm.ny or iz >= prm.nz or ix <= 0 or iy <= 0 or iz <= 0:
        return 0.0
    d = lambda i, j, k: float(denx[k - 1, j - 1, i - 1])
    return blend_103(r, s, t, d(ix, iy, iz), d(ix, iy, iz + 1), d(ix, iy + 1, iz), d(ix, iy + 1, iz + 1),
                     d(ix + 1, iy, iz), d(ix + 1, iy, iz + 1), d(ix + 1, iy + 1, iz), d(ix + 1, iy + 1, iz + 1))


def sorteo(prm: Params, denx, n_par, lo, hi, seed, max_attempts=1 << 20):
    """Rejection sampling of sorteo.f90:19-30: a uniform point in the box
    [lo, hi] and a uniform dum in [0, max(denx)); accepted when dum < f(point).
    The random numbers come from Philox4x64-10 keyed by (seed, 0): attempt a of
    particle i uses the block of counter (i, a, 0, 0), words 0-2 for the point
    and word 3 for dum, each as a 53-bit uniform.  Returns rimp (3, n_par) and
    the attempts used per particle."""
    denmax = float(np.max(denx))
    out = np.zeros((3, n_par))
    tries = np.zeros(n_par, dtype=np.int64)
    for i in range(n_par):
        for a in range(max_attempts):
            w = philox4x64((i, a, 0, 0), (seed, 0))
            xyz = [_u53(w[c]) * (hi[c] - lo[c]) + lo[c] for c in range(3)]
            dum = _u53(w[3]) * denmax
            if dum < interpola3denx(prm, denx, xyz):
                out[:, i] = xyz
                tries[i] = a + 1
                break
        else:
            raise RuntimeError(f"particle {i} not accepted in {max_attempts} attempts")
    return out, tries


# --------------------------------------------------------------------------
# Fresh start of a droplet (mode 0, DFT4He3d.f90:585-605) from a closed-form
# density: the layout and random streams of hetddft_init_droplet
# --------------------------------------------------------------------------
_U32 = np.uint64(0xFFFFFFFF)


def _mulhilo64(a, m):
    """High and low 64-bit words of a * m for a uint64 array a and a 64-bit
    constant m (schoolbook product of 32-bit halves)."""
    a = a.astype(np.uint64)
    m = np.uint64(m)
    al, ah = a & _U32, a >> np.uint64(32)
    ml, mh = m & _U32, m >> np.uint64(32)
    ll, hl, lh, hh = al * ml, ah * ml, al * mh, ah * mh
    cross = (ll >> np.uint64(32)) + (hl & _U32) + lh
    hi = hh + (hl >> np.uint64(32)) + (cross >> np.uint64(32))
    return hi, a * m


def philox4x64_many(c0, key, rounds=10):
    """philox4x64 of counters (c0[i], 0, 0, 0) under one key, vectorised over
    i; returns the four words as uint64 arrays."""
    with np.errstate(over="ignore"):
        c = [np.asarray(c0, dtype=np.uint64), np.zeros(len(c0), np.uint64), np.zeros(len(c0), np.uint64),
             np.zeros(len(c0), np.uint64)]
        k0, k1 = np.uint64(int(key[0]) & _M64), np.uint64(int(key[1]) & _M64)
        for r in range(rounds):
            if r > 0:
                k0 = np.uint64((int(k0) + _PHILOX_W[0]) & _M64)
                k1 = np.uint64((int(k1) + _PHILOX_W[1]) & _M64)
            h0, l0 = _mulhilo64(c[0], _PHILOX_M[0])
            h1, l1 = _mulhilo64(c[2], _PHILOX_M[1])
            c = [h1 ^ c[1] ^ k0, l1, h0 ^ c[3] ^ k1, l0]
    return c


def _u53_open(w):
    """((w >> 11) + 1) / 2^53 in (0, 1]."""
    return ((w >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _u53_vec(w):
    return (w >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def assoc_legendre(l, m, x):
    """P_l^m(x) with the Condon-Shortley phase, 0 <= m <= l <= 4, from the
    Rodrigues form (-1)^m (1-x^2)^(m/2) d^m/dx^m P_l(x) with the Legendre
    polynomials written out (independent of scipy's recurrence)."""
    legendre = {0: [1.0], 1: [0.0, 1.0], 2: [-0.5, 0.0, 1.5], 3: [0.0, -1.5, 0.0, 2.5],
                4: [0.375, 0.0, -3.75, 0.0, 4.375]}
    coef = np.polynomial.polynomial.polyder(np.array(legendre[l]), m) if m else np.array(legendre[l])
    d = np.polynomial.polynomial.polyval(x, coef)
    s = np.sqrt(np.maximum(1.0 - x * x, 0.0))
    return (-1.0) ** m * s ** m * d


def droplet_density(prm: Params, n_he, rho0=0.0218, r0=2.22, a=0.7, cavity=0.0, cavity_w=0.3, ripple=0.0,
                    clm=None, centre=(0.0, 0.0, 0.0)):
    """Woods-Saxon density of hetddft_init_droplet on the grid x(ix) = -xmax
    + hx (ix-1): rho0/(1+exp((r-R)/a)), R = r0 N^(1/3) + ripple s/max|s|,
    s = sum_{l<=4, |m|<=l} c_lm P_l^|m|(cos theta) (cos m phi | sin |m| phi),
    divided by 1+exp(-(r-cavity)/cavity_w), rescaled to sum*dxyz = N."""
    x = prm.axis_grid("x") - centre[0]
    y = prm.axis_grid("y") - centre[1]
    z = prm.axis_grid("z") - centre[2]
    z3, y3, x3 = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt(x3 * x3 + y3 * y3 + z3 * z3)
    R = r0 * n_he ** (1.0 / 3.0)
    if ripple > 0.0:
        ct = np.divide(z3, r, out=np.zeros_like(r), where=r > 0)
        phi = np.arctan2(y3, x3)
        surf = np.zeros_like(r)
        k = 0
        for l in range(1, 5):
            for m in range(-l, l + 1):
                t = np.cos(m * phi) if m >= 0 else np.sin(-m * phi)
                surf = surf + clm[k] * (assoc_legendre(l, abs(m), ct) * t)
                k += 1
        surf = surf / max(np.abs(surf).max(), 1e-300)
        R = R + ripple * surf
    rho = rho0 / (1.0 + np.exp((r - R) / a))
    if cavity > 0.0:
        rho = rho / (1.0 + np.exp(-(r - cavity) / cavity_w))
    return rho * (n_he / (rho.sum() * prm.dxyz))


def droplet_state(prm: Params, n_he, seed, lamda=(0.0, 0.0, 0.0), phase_noise=0.0, sigma=0.0,
                  v0=(0.0, 0.0, 0.0), centre=(0.0, 0.0, 0.0), **dens):
    """psi, rimp, vimp of hetddft_init_droplet: psi = sqrt(rho) exp(i th),
    th = x lamdax + y lamday + z lamdaz (DFT4He3d.f90:587-596) plus
    phase_noise * g(point); particle p at centre + sigma (g0, g1, g2) with
    velocity v0 * Ktops (DFT4He3d.f90:600-605).  g: Box-Muller normals of
    Philox4x64-10 words, key (seed, 1) and counter (i, 0, 0, 0) for point i,
    key (seed, 2) and counter (p, 0, 0, 0) for particle p.  rimp is None when
    sigma <= 0 (positions left to sorteo)."""
    rho = droplet_density(prm, n_he, centre=centre, **dens)
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    th = (x[None, None, :] * lamda[0] + y[None, :, None] * lamda[1]) + z[:, None, None] * lamda[2]
    th = np.broadcast_to(th, rho.shape).copy()
    if phase_noise > 0.0:
        w = philox4x64_many(np.arange(rho.size, dtype=np.uint64), (seed, 1))
        g = np.sqrt(-2.0 * np.log(_u53_open(w[0]))) * np.cos(6.283185307179586 * _u53_vec(w[1]))
        th = th + phase_noise * g.reshape(rho.shape)
    amp = np.sqrt(rho)
    psi = amp * np.cos(th) + 1j * (amp * np.sin(th))
    n = prm.n_par
    rimp = None
    if sigma > 0.0:
        w = philox4x64_many(np.arange(n, dtype=np.uint64), (seed, 2))
        r01 = np.sqrt(-2.0 * np.log(_u53_open(w[0])))
        t01 = 6.283185307179586 * _u53_vec(w[1])
        r23 = np.sqrt(-2.0 * np.log(_u53_open(w[2])))
        t23 = 6.283185307179586 * _u53_vec(w[3])
        rimp = np.stack([centre[0] + sigma * (r01 * np.cos(t01)), centre[1] + sigma * (r01 * np.sin(t01)),
                         centre[2] + sigma * (r23 * np.cos(t23))])
    ktops = 7.63822291
    vimp = np.stack([np.full(n, v0[c] * ktops) for c in range(3)])
    return psi, rimp, vimp


def diagnostics(prm: Params, psi, pot4, denx, vimp, aimp):
    """Per-step diagnostics: the He count sum(den) dxyz (DFT4He3d.f90:884),
    the centre-of-mass moments sum(den x_c) dxyz of r_cm (r_cm.f90:23-58; the
    reference divides them by n4), the impurity mean-square radii
    sqrt(sum(denx x_c^2) dxyz) (DFT4He3d.f90:918-928), min/max of vimp and
    aimp per component (:900-903); and the chemical potential
    <psi|H|psi>/<psi|psi>, H = -h2o2m4 lap + pot4, and the integral of the
    current j = 2 h2o2m4 Im(conj(psi) grad psi).  Layout of hetddft_diagnostics."""
    den = np.abs(psi) ** 2
    dv = prm.dxyz
    x, y, z = prm.axis_grid("x"), prm.axis_grid("y"), prm.axis_grid("z")
    xs = [x[None, None, :], y[None, :, None], z[:, None, None]]
    out = np.zeros(23)
    out[0] = den.sum() * dv
    for k in range(3):
        out[1 + k] = (den * xs[k]).sum() * dv
    lap = (deriv(psi, "x", 2, prm.hx) + deriv(psi, "y", 2, prm.hy)) + deriv(psi, "z", 2, prm.hz)
    hp = pot4 * psi - prm.h2o2m4 * lap
    out[4] = np.sum(np.real(np.conj(psi) * hp)) / den.sum()
    for k, j in enumerate(current(prm, psi)):
        out[5 + k] = j.sum() * dv
    for k in range(3):
        out[8 + k] = math.sqrt((denx * xs[k] ** 2).sum() * dv)
    if vimp.shape[1] > 0:
        for k in range(6):
            col = vimp[k] if k < 3 else aimp[k - 3]
            out[11 + 2 * k] = col.min()
            out[12 + 2 * k] = col.max()
    return out


def current(prm: Params, psi):
    """j = 2 h2o2m4 Im(conj(psi) d_c psi), c = x, y, z, with the 13-point stencil."""
    hs = {"x": prm.hx, "y": prm.hy, "z": prm.hz}
    return tuple(2.0 * prm.h2o2m4 * np.imag(np.conj(psi) * deriv(psi, a, 1, hs[a])) for a in ("x", "y", "z"))


class Oracle:
    """State and per-step operators of one droplet+impurity instance."""

    C112 = 112.0 / 121.0          # steppc.f90:21-25
    C9 = 9.0 / 121.0
    C4O3 = 4.0 / 3.0
    THIRD = 0.3333333333          # steppc.f90:133,147 literal

    def __init__(self, prm: Params, kernels: dict):
        self.p = prm
        self.fvlj4 = kernels["fvlj4"]
        self.wcgk = kernels["wcgk"]
        self.kalfs = kernels["kalfs"]
        self.vq = kernels["vq"]
        shp = prm.shape
        self.psi = np.zeros(shp, np.complex128)
        self.psiold = np.zeros((3,) + shp, np.complex128)
        self.hpsiold = np.zeros((2,) + shp, np.complex128)
        self.pc = np.zeros(shp, np.complex128)          # dimen.f90:133
        self.q = np.zeros(shp, np.complex128)
        n = prm.n_par
        self.rimp = np.zeros((3, n))
        self.vimp = np.zeros((3, n))
        self.aimp = np.zeros((3, n))
        self.rimpold = np.zeros((3, 3, n))
        self.vimpold = np.zeros((3, 3, n))
        self.aimpold = np.zeros((2, 3, n))
        self.qr = np.zeros((3, n))
        self.qv = np.zeros((3, n))
        self.pcr = np.zeros((3, n))
        self.pcv = np.zeros((3, n))
        self.ioldp = [0, 1, 2]
        self.ioldh = [0, 1]
        self.ioldr = [0, 1, 2]
        self.ioldv = [0, 1, 2]
        self.iolda = [0, 1]
        z = prm.axis_grid("z")
        # timec = cmplx(Lambdah*(1+tanh((z-tzmean)/tzsurf)), 1) rounded to
        # REAL(4) by the KIND-less cmplx (DFT4He3d.f90:773-782)
        re = prm.lambdah * (1.0 + np.tanh((z - prm.tzmean) / prm.tzsurf))
        re = re.astype(np.float32).astype(np.float64)
        self.timec_z = re + 1j * np.ones_like(re)

    # ---- state -----------------------------------------------------------
    def set_state(self, psi, rimp, vimp):
        self.psi = np.array(psi, dtype=np.complex128).reshape(self.p.shape)
        self.rimp = np.array(rimp, dtype=np.float64).reshape(3, self.p.n_par)
        self.vimp = np.array(vimp, dtype=np.float64).reshape(3, self.p.n_par)

    @staticmethod
    def dens(psi):
        return psi.real * psi.real + psi.imag * psi.imag

    def timec(self):
        return self.timec_z[:, None, None]

    # ---- potential evaluation -------------------------------------------
    def denpart(self):
        """NGP histogram, ironing, psix, hpsix (denpart.f90:5-49)."""
        p = self.p
        hxyz = p.hx * p.hy * p.hz
        cnt = ngp_counts(p, self.rimp)
        denx = cnt / p.n_par / hxyz
        for _ in range(p.ironx):
            denx = ironing(denx)
        self.set_denx(denx)

    def set_denx(self, denx):
        """psix = sqrt(denx), hpsix = -h2o2mx lap psix from a given impurity
        density (denpart.f90:40-46)."""
        p = self.p
        self.denx = denx
        self.psix = np.sqrt(denx)
        s1 = deriv(self.psix, "x", 2, p.hx)
        s2 = deriv(self.psix, "y", 2, p.hy)
        s3 = deriv(self.psix, "z", 2, p.hz)
        self.hpsix = -p.h2o2mx * ((s1 + s2) + s3)

    def potenimp(self):
        """uimp = V * denx by FFT (potenimp.f90:1-20)."""
        p = self.p
        self.denpart()
        wk1 = r2c(self.denx) * self.vq
        self.uimp = c2r(wk1, p.shape, p.renor)

    def poten(self):
        """OT mean field pot4 (poten.f90:5-129)."""
        p = self.p
        den = self.den
        fden = r2c(den)
        self.dencg = c2r(fden * self.wcgk, p.shape, p.renor)
        self.dxden = deriv(den, "x", 1, p.hx)
        self.dyden = deriv(den, "y", 1, p.hy)
        self.dzden = deriv(den, "z", 1, p.hz)
        self.delj4 = c2r(fden * self.fvlj4, p.shape, p.renor)
        a0 = p.cp4 / 2.0
        a1 = p.cpp4 / 3.0
        dcg = self.dencg
        sto1 = den * dcg * (p.cp4 + p.cpp4 * dcg)
        sto1 = c2r(r2c(sto1) * self.wcgk, p.shape, p.renor)
        t = np.tanh(p.beta * (den - p.den_m))
        penalty = p.C * (1.0 + t + p.beta * den * (1.0 - t * t))
        pot4 = self.delj4 + dcg * dcg * (a0 + a1 * dcg) + sto1
        if p.core == "OTC":
            self.denalf = c2r(fden * self.kalfs, p.shape, p.renor)
            self.term_alfa()
            pot4 = pot4 + self.ualphas
        self.pot4 = pot4 + self.uimp + penalty

    def term_alfa(self):
        """OT alpha_s field (term_alfa.f90:5-117)."""
        p = self.p
        falfs = 1.0 - self.denalf / p.den0s
        self.falfs = falfs
        ints = []
        for d in (self.dxden, self.dyden, self.dzden):
            ints.append(c2r(self.kalfs * r2c(falfs * d), p.shape, p.renor))
        self.intxalf, self.intyalf, self.intzalf = ints
        sto1 = ((self.dxden * self.intxalf) + (self.dyden * self.intyalf)) + (self.dzden * self.intzalf)
        ua = c2r(self.kalfs * r2c(sto1), p.shape, p.renor)
        ua = ua / p.den0s
        s4 = deriv(falfs * self.intxalf, "x", 1, p.hx)
        s5 = deriv(falfs * self.intyalf, "y", 1, p.hy)
        s6 = deriv(falfs * self.intzalf, "z", 1, p.hz)
        self.ualphas = (((ua + s4) + s5) + s6) * p.h2o2m4 * p.alphas

    def forceimp(self):
        """Force fields and trilinear gather onto the particles (potenimp.f90:24-73)."""
        p = self.p
        f = []
        for d in (self.dxden, self.dyden, self.dzden):
            f.append(c2r(-(r2c(d) * self.vq), p.shape, p.renor))
        s1 = deriv(self.psix, "x", 1, p.hx)
        s2 = deriv(self.psix, "y", 1, p.hy)
        s3 = deriv(self.psix, "z", 1, p.hz)
        s4 = deriv(self.hpsix, "x", 1, p.hx)
        s5 = deriv(self.hpsix, "y", 1, p.hy)
        s6 = deriv(self.hpsix, "z", 1, p.hz)
        sto7 = np.zeros_like(self.denx)
        pos = self.denx > 0.0
        sto7[pos] = 1.0 / self.denx[pos]
        m = p.m_imp
        self.fx = (f[0] + sto7 * (s1 * self.hpsix - self.psix * s4)) / m
        self.fy = (f[1] + sto7 * (s2 * self.hpsix - self.psix * s5)) / m
        self.fz = (f[2] + sto7 * (s3 * self.hpsix - self.psix * s6)) / m
        self.aimp = interpola3(p, self.rimp, self.fx, self.fy, self.fz)

    def veval(self):
        """potenimp; poten; forceimp (DFT4He3d.f90:861-863)."""
        self.potenimp()
        self.poten()
        self.forceimp()

    # ---- Hamiltonian ------------------------------------------------------
    def hpsi(self):
        """Sto4c = timec*((d2x+d2y+d2z)psi*h2o2m4 - pot4*psi) (steppc.f90:36-46)."""
        p = self.p
        s1 = deriv(self.psi, "x", 2, p.hx)
        s2 = deriv(self.psi, "y", 2, p.hy)
        s3 = deriv(self.psi, "z", 2, p.hz)
        lap = (s1 + s2) + s3
        return self.timec() * (lap * p.h2o2m4 - self.pot4 * self.psi)

    # ---- driver pieces ---------------------------------------------------
    def start(self):
        """den = conjg(psi)*psi, first Veval, zero particle histories (DFT4He3d.f90:630-634,841-845)."""
        self.den = self.dens(self.psi)
        self.veval()
        self.qr[:] = 0.0
        self.qv[:] = 0.0
        self.rimpold[:] = 0.0
        self.vimpold[:] = 0.0
        self.aimpold[:] = 0.0

    def steprk(self, dt):
        """Runge-Kutta-Gill start-up step (steprk.f90:1-128)."""
        s2 = math.sqrt(2.0)
        arun = [0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 1.0 / 6.0]
        brun = [2.0, 1.0, 1.0, 2.0]
        crun = [0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 0.5]
        for j in range(4):
            a, b, c = arun[j], brun[j], crun[j]
            sto4 = self.hpsi()
            sto1 = a * (sto4 - b * self.q)
            self.q = self.q + 3.0 * sto1 - c * sto4
            if j == 0:
                self.hpsiold[1] = self.hpsiold[0]
                self.hpsiold[0] = sto4
                self.psiold[2] = self.psiold[1]
                self.psiold[1] = self.psiold[0]
                self.psiold[0] = self.psi
            self.psi = self.psi + dt * sto1
            self.den = self.dens(self.psi)
            stor = a * (self.vimp - b * self.qr)
            self.qr = self.qr + 3.0 * stor - c * self.vimp
            if j == 0:
                self.rimpold[2] = self.rimpold[1]
                self.rimpold[1] = self.rimpold[0]
                self.rimpold[0] = self.rimp
            self.rimp = self.rimp + dt * stor
            stor = a * (self.aimp - b * self.qv)
            self.qv = self.qv + 3.0 * stor - c * self.aimp
            if j == 0:
                self.aimpold[1] = self.aimpold[0]
                self.aimpold[0] = self.aimp
                self.vimpold[2] = self.vimpold[1]
                self.vimpold[1] = self.vimpold[0]
                self.vimpold[0] = self.vimp
            self.vimp = self.vimp + dt * stor
            if j <= 2:
                self.veval()
        self.ioldp = [0, 1, 2]
        self.ioldr = [0, 1, 2]
        self.ioldv = [0, 1, 2]
        self.ioldh = [0, 1]
        self.iolda = [0, 1]

    def steppc(self, dt):
        """Hamming predictor-modifier-corrector step (steppc.f90:1-155)."""
        c112, c9, c4o3 = self.C112, self.C9, self.C4O3
        P, H = self.ioldp, self.ioldh
        sto4 = self.hpsi()
        sto1 = self.psiold[P[2]] + c4o3 * dt * (2.0 * sto4 - self.hpsiold[H[0]]
                                                 + 2.0 * self.hpsiold[H[1]])
        self.psiold[P[2]] = self.psi
        self.psi = sto1 - c112 * self.pc
        self.pc = sto1
        self.hpsiold[H[1]] = sto4
        self.den = self.dens(self.psi)
        H[0], H[1] = H[1], H[0]
        R, V, A = self.ioldr, self.ioldv, self.iolda
        stor = self.rimpold[R[2]] + c4o3 * dt * (2.0 * self.vimp - self.vimpold[V[0]]
                                                  + 2.0 * self.vimpold[V[1]])
        self.rimpold[R[2]] = self.rimp
        self.rimp = stor - c112 * self.pcr
        self.pcr = stor
        stor = self.vimpold[V[2]] + c4o3 * dt * (2.0 * self.aimp - self.aimpold[A[0]]
                                                  + 2.0 * self.aimpold[A[1]])
        self.vimpold[V[2]] = stor - c112 * self.pcv
        self.pcv = stor
        self.aimpold[A[1]] = self.aimp
        A[0], A[1] = A[1], A[0]

        self.veval()

        sto4 = self.hpsi()
        sto5 = 0.125 * (9.0 * self.psiold[P[2]] - self.psiold[P[1]]
                        + 3.0 * dt * (sto4 + 2.0 * self.hpsiold[H[0]] - self.hpsiold[H[1]]))
        self.pc = self.pc - sto5
        self.psi = sto5 + c9 * self.pc
        err_he = c9 * np.sum(np.abs(self.pc))
        self.den = self.dens(self.psi)
        err_he = err_he / self.p.nxyz
        P[0], P[1], P[2] = P[2], P[0], P[1]

        stor = 0.125 * (9.0 * self.rimpold[R[2]] - self.rimpold[R[1]]
                        + 3.0 * dt * (self.vimpold[V[2]] + 2.0 * self.vimp - self.vimpold[V[0]]))
        self.pcr = self.pcr - stor
        self.rimp = stor + c9 * self.pcr
        err_imp = np.sum(np.abs(c9 * self.pcr)) * self.THIRD / self.p.n_par
        R[0], R[1], R[2] = R[2], R[0], R[1]

        stor = 0.125 * (9.0 * self.vimp - self.vimpold[V[1]]
                        + 3.0 * dt * (self.aimp + 2.0 * self.aimpold[A[0]] - self.aimpold[A[1]]))
        self.pcv = self.pcv - stor
        self.vimpold[V[2]] = self.vimp
        self.vimp = stor + c9 * self.pcv
        err_vimp = np.sum(np.abs(c9 * self.pcv)) * self.THIRD / self.p.n_par
        V[0], V[1], V[2] = V[2], V[0], V[1]
        return err_he, err_imp, err_vimp

    def iterate(self, it_in_run, dt):
        """One driver iteration: RK for the first three, PC after, then Veval.

        Follows DFT4He3d.f90:849-863.  Returns the PC error triple or None.
        """
        if it_in_run < 3:
            self.steprk(dt)
            errs = None
        else:
            errs = self.steppc(dt)
        self.veval()
        return errs

    def particle_means(self):
        """Per-step means of rimp, vimp (A/ps), aimp (A/ps^2) (DFT4He3d.f90:864-878)."""
        n = self.p.n_par
        r = self.rimp.sum(axis=1) / n
        v = self.vimp.sum(axis=1) / n * PSTOK
        a = self.aimp.sum(axis=1) / n * PSTOK * PSTOK
        return np.concatenate([r, v, a])

    def energy(self):
        """Energy decomposition (energy.f90:12-129) plus the norm (DFT4He3d.f90:885)."""
        p = self.p
        s1 = deriv(self.psi, "x", 1, p.hx)
        s2 = deriv(self.psi, "y", 1, p.hy)
        s3 = deriv(self.psi, "z", 1, p.hz)
        aux1 = 0.5 * p.cp4
        aux2 = (1.0 / 3.0) * p.cpp4
        den, dcg = self.den, self.dencg
        ekin4 = np.sum(np.abs(s1) ** 2 + np.abs(s2) ** 2 + np.abs(s3) ** 2)
        elj4 = np.sum(self.delj4 * den)
        ecor4 = np.sum(den * dcg ** 2 * (aux1 + aux2 * dcg))
        ealphas = 0.0
        if p.core in ("OTE", "OTC"):
            ealphas = np.sum(self.falfs * (self.dxden * self.intxalf + self.dyden * self.intyalf
                                           + self.dzden * self.intzalf))
            ealphas = -p.h2o2m4 * 0.5 * p.alphas * ealphas * p.dxyz
        esolid = p.C * np.sum(den * (1.0 + np.tanh(p.beta * (den - p.den_m)))) * p.dxyz
        ekin4 = p.h2o2m4 * ekin4 * p.dxyz
        elj4 = 0.5 * elj4 * p.dxyz
        ecor4 = ecor4 * p.dxyz
        etot4 = ekin4 + elj4 + ecor4 + esolid + ealphas
        ekinx = 0.5 * p.m_imp * np.sum(self.vimp * self.vimp) / p.n_par
        ekinqx = np.sum(self.psix * self.hpsix) * p.dxyz
        ehex = np.sum(self.uimp * den) * p.dxyz
        eimpu = ekinx + ehex + ekinqx
        return dict(etot=etot4 + eimpu, etot4=etot4, ekin4=ekin4, elj4=elj4, ealphas=ealphas,
                    esolid=esolid, ecor4=ecor4, eimpu=eimpu, ekinx=ekinx, ekinqx=ekinqx,
                    ehex=ehex, norm=np.sum(den) * p.dxyz)
