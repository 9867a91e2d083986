"""Seeded synthetic initial states for the benchmark and test workloads.

The reference starts from a density file (readenc.f90, mode 0) and samples
the test particles by rejection (sorteo.f90).  The workloads here are built
instead from closed forms with a seeded Philox generator, so every run and
every machine sees the same inputs:

* He density: Woods-Saxon rho(r) = 0.0218/(1+exp((r-R)/0.7)), R = 2.22 N^(1/3),
  rescaled so that sum(rho)*dxyz = N; optional surface ripple of spherical
  harmonics l <= 4; optional cavity around the impurity, rho times
  1/(1+exp(-(r-r_b)/0.3)).  The benchmark workloads take r_b = 4.2 A, where
  the cored Morse potential (D = 20 K, r_e = 5 A, alpha = 0.8/A) changes sign,
  as a ground-state start has: with He inside the 2000 K core the density is
  expelled into a thin shell at the core edge and the explicit time step
  diverges after about 400 PC steps at 128^3;
* psi = sqrt(rho) exp(i phi), phi = 0 or a per-point N(0, sigma_phi^2) phase;
* particles: isotropic Gaussian cloud at the origin, common velocity v0 along
  a unit vector, converted from A/ps to A*K with 7.63822291
  (DFT4He3d.f90:600-605).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KTOPS = 7.63822291


@dataclass
class Workload:
    """One instance: grid, droplet, particles and time step."""
    name: str
    n: int                 # grid points per axis
    h: float               # grid step (A)
    n_he: int              # He atoms
    n_par: int             # test particles
    sigma: float           # particle cloud width (A)
    v0: float = 0.0        # impurity speed (A/ps); 100 m/s = 1 A/ps
    vdir: tuple = (0.0, 0.0, 1.0)
    phase_noise: float = 0.0
    ripple: float = 0.0    # surface ripple amplitude (A), harmonics l <= 4
    seed: int = 1
    dt_ps: float = 5e-4    # 0.5 fs
    n_pc: int = 100
    mimpur: float = 40.0
    ironx: int = 3
    cavity: float = 0.0    # radius of the He cavity around the impurity (A); 0: none

    @property
    def xmax(self):
        return self.n * self.h / 2.0

    @property
    def deltat(self):
        """Time step in K^-1 (DFT4He3d.f90:235)."""
        return self.dt_ps / KTOPS


WORKLOADS = {
    # reference's own CPU-runnable case: 64^3, dx = 0.5 A, N = 100
    "c1": Workload("c1", 64, 0.5, 100, 1000, 0.5, seed=1, n_pc=100, cavity=4.2),
    # the metric case: 128^3, dx = 0.4 A, N = 500, 1% phase noise, 10^4 particles
    "c2": Workload("c2", 128, 0.4, 500, 10000, 0.8, v0=1.0, phase_noise=0.01, seed=2, n_pc=1000, cavity=4.2),
    # large droplet: 256^3, dx = 0.3 A, N = 2000, rippled surface, 10^5 particles
    "c5": Workload("c5", 256, 0.3, 2000, 100000, 1.0, v0=2.0, ripple=0.3, seed=5, n_pc=2000, cavity=4.2),
    # small parity cases
    "t32": Workload("t32", 32, 0.5, 20, 256, 0.5, v0=1.0, phase_noise=0.01, seed=7, n_pc=5),
    "t24": Workload("t24", 24, 0.6, 12, 128, 0.5, v0=0.5, phase_noise=0.01, seed=8, n_pc=5),
}


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def ensemble_c4(count=32, seed=4, n_pc=500):
    """Ensemble of uneven instances: grids 48^3, 64^3, 80^3, 96^3 in turn at
    dx = 0.45 A; N uniform in [50, 300], clipped so that the droplet radius
    R = 2.22 N^(1/3) stays 5 A inside the box (a 48^3 box holds N <= 17 and
    takes that N); 2000 test particles (sigma 0.5 A) with a speed uniform in
    0-300 m/s along a random direction."""
    g = _rng(seed)
    out = []
    for i in range(count):
        n = (48, 64, 80, 96)[i % 4]
        h = 0.45
        half = n * h / 2.0
        nmax = int(((half - 5.0) / 2.22) ** 3)
        hi = min(300, nmax)
        lo = min(50, hi)
        n_he = int(g.integers(lo, hi + 1))
        speed = float(g.uniform(0.0, 3.0))          # A/ps; 100 m/s = 1 A/ps
        d = g.standard_normal(3)
        d /= np.linalg.norm(d)
        out.append(Workload(f"c4-{i}", n, h, n_he, 2000, 0.5, v0=speed, vdir=tuple(float(x) for x in d),
                            seed=400 + i, n_pc=n_pc, cavity=4.2))
    return out


def _ylm_real(l, m, ct, phi):
    """Real spherical harmonics (unnormalised) for the ripple, l <= 4."""
    from scipy.special import lpmv
    P = lpmv(abs(m), l, ct)
    return P * (np.cos(m * phi) if m >= 0 else np.sin(-m * phi))


def ripple_coeffs(wl: Workload):
    """The 24 ripple coefficients c_lm (l = 1..4, m = -l..l in that order)."""
    g = _rng(wl.seed + 1000)
    return [float(g.standard_normal()) for l in range(1, 5) for m in range(-l, l + 1)]


def droplet_spec(wl: Workload) -> dict:
    """The workload's droplet as keyword arguments of hetddft.droplet (the
    device-built start): the density of woods_saxon, phases and particles
    from the Philox streams of hetddft_init_droplet (same distributions as
    initial_state, other values)."""
    d = np.asarray(wl.vdir, dtype=np.float64)
    d = d / np.linalg.norm(d)
    return dict(n_he=float(wl.n_he), seed=wl.seed, cavity=wl.cavity, cavity_w=0.3, ripple=wl.ripple,
                clm=ripple_coeffs(wl) if wl.ripple > 0.0 else None, phase_noise=wl.phase_noise,
                sigma=wl.sigma, v0=tuple(float(wl.v0 * d[c]) for c in range(3)))


def woods_saxon(wl: Workload, shape=None):
    n = wl.n
    x = -wl.xmax + wl.h * np.arange(n)
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(x3 * x3 + y3 * y3 + z3 * z3)
    R = 2.22 * wl.n_he ** (1.0 / 3.0)
    if wl.ripple > 0.0:
        coeffs = iter(ripple_coeffs(wl))
        ct = np.divide(z3, r, out=np.zeros_like(r), where=r > 0)
        phi = np.arctan2(y3, x3)
        surf = np.zeros_like(r)
        for l in range(1, 5):
            for m in range(-l, l + 1):
                surf += next(coeffs) * _ylm_real(l, m, ct, phi)
        surf /= max(np.abs(surf).max(), 1e-300)
        R = R + wl.ripple * surf
    rho = 0.0218 / (1.0 + np.exp((r - R) / 0.7))
    if wl.cavity > 0.0:
        rho = rho / (1.0 + np.exp(-(r - wl.cavity) / 0.3))
    dxyz = wl.h ** 3
    rho *= wl.n_he / (rho.sum() * dxyz)
    return rho


def initial_state(wl: Workload):
    """psi (nz,ny,nx) complex, rimp (3,N_par), vimp (3,N_par) in A*K."""
    rho = woods_saxon(wl)
    g = _rng(wl.seed)
    if wl.phase_noise > 0.0:
        phi = wl.phase_noise * g.standard_normal(rho.shape)
        psi = np.sqrt(rho) * np.exp(1j * phi)
    else:
        psi = np.sqrt(rho).astype(np.complex128)
    rimp = wl.sigma * g.standard_normal((3, wl.n_par))
    d = np.asarray(wl.vdir, dtype=np.float64)
    d = d / np.linalg.norm(d)
    vimp = np.empty((3, wl.n_par))
    for c in range(3):
        vimp[c, :] = wl.v0 * d[c] * KTOPS
    return psi, rimp, vimp


def params_for(wl: Workload, precision=0, **kw):
    """hetddft_params for a workload (loads the library)."""
    from hetddft import default_params  # noqa: E402  (package-local import)
    kw.setdefault("ironx", wl.ironx)
    return default_params(nx=wl.n, ny=wl.n, nz=wl.n, xmax=wl.xmax, ymax=wl.xmax, zmax=wl.xmax,
                          n_par=wl.n_par, mimpur=wl.mimpur, precision=precision, **kw)
