"""Pins the oracle's potential evaluation and energy by mathematics and physics.

The GPU is compared with the oracle; these tests check the oracle itself,
each against something its formulas do not restate:

* pot4 = dE/drho: the mean field that ``poten`` builds (poten.f90:62-124,
  term_alfa.f90:28-113) is the functional derivative of the energy that
  ``energy`` sums (energy.f90:62-124).  A central difference of E along
  random density directions must equal sum(pot4 * drho) dV.  The discrete
  identity is exact up to the difference quotient's truncation: the
  zero-padded first-derivative stencil is an antisymmetric matrix and the
  k-space kernels are real and even, so the convolutions are symmetric.
* the bulk Orsay-Trento equation of state: at the saturation density
  rho0 = 0.021836 A^-3 the OT functional gives e/N = -7.15 K and zero
  pressure, so the uniform-liquid chemical potential equals e/N
  (Dalfovo et al., Phys. Rev. B 52, 1193 (1995), Table I and Eq. 4).
* the alpha_s energy of a Gaussian density with f = 1, in closed form
  (a k-space integral of k^2 |rho(k)|^2 exp(-k^2 l^2/4)).
* ``forceimp`` (potenimp.f90:42-69): the quantum part equals -grad Q for a
  Gaussian impurity density, Q = -h2o2mx lap(psi_X)/psi_X, which is linear in
  r; the convolution part obeys Newton's third law exactly on the grid
  (force of the He on the impurity = minus the force of the impurity on the
  He).
* ``denpart`` (denpart.f90:24-46): sum(denx) h^3 equals the fraction of the
  particles inside the box; the quantum kinetic energy of a Gaussian denx is
  h2o2mx * 3/(4 sigma^2).

Each check also runs on oracles with a planted mistake, which it must catch.
"""
import math

import numpy as np
import pytest

from oracle import hetddft_oracle as orc


# ------------------------------------------------------------------ fixtures
@pytest.fixture(scope="module")
def t24():
    import inputs
    wl = inputs.WORKLOADS["t24"]
    prm = orc.Params(nx=wl.n, ny=wl.n, nz=wl.n, xmax=wl.xmax, ymax=wl.xmax, zmax=wl.xmax,
                     n_par=wl.n_par, mimpur=wl.mimpur, ironx=wl.ironx)
    k = orc.make_kernels(prm, nq=200)
    psi, r, v = inputs.initial_state(wl)
    return prm, k, psi, r, v


def _oracle(prm, k, psi, r, v, cls=orc.Oracle):
    o = cls(prm, k)
    o.set_state(psi, r, v)
    o.den = o.dens(o.psi)
    o.potenimp()
    return o


# ---------------------------------------------------------- planted oracles
class FlippedAlpha(orc.Oracle):
    """pot4 with the alpha_s field of the wrong sign."""

    def poten(self):
        super().poten()
        self.pot4 = self.pot4 - 2.0 * self.ualphas


class NoDen0s(orc.Oracle):
    """alpha_s field without the 1/den0s of its convolution term."""

    def poten(self):
        super().poten()
        p = self.p
        sto1 = (self.dxden * self.intxalf + self.dyden * self.intyalf) + self.dzden * self.intzalf
        ua = orc.c2r(self.kalfs * orc.r2c(sto1), p.shape, p.renor)
        self.pot4 = self.pot4 + p.h2o2m4 * p.alphas * ua * (1.0 - 1.0 / p.den0s)


class BarePenalty(orc.Oracle):
    """solid penalty C(1 + t) without its beta rho (1 - t^2) term."""

    def poten(self):
        super().poten()
        p = self.p
        t = np.tanh(p.beta * (self.den - p.den_m))
        self.pot4 = self.pot4 - p.C * p.beta * self.den * (1.0 - t * t)


class UnsmoothedCorrelation(orc.Oracle):
    """correlation term rho rhobar (c2 + c3 rhobar) without its W* convolution."""

    def poten(self):
        super().poten()
        p = self.p
        g = self.den * self.dencg * (p.cp4 + p.cpp4 * self.dencg)
        self.pot4 = self.pot4 - orc.c2r(orc.r2c(g) * self.wcgk, p.shape, p.renor) + g


class DoubledLJ(orc.Oracle):
    """energy with E_LJ = sum(delj4 rho) dV (the 1/2 dropped)."""

    def energy(self):
        e = super().energy()
        e["etot"] = e["etot"] + e["elj4"]
        e["elj4"] = 2.0 * e["elj4"]
        return e


# ------------------------------------------------------------ pot4 = dE/drho
def _energy_at(o, rho):
    o.den = rho
    o.poten()
    return o.energy()["etot"]


def _directional(o, rho, drho, eps):
    """dE/d(eps) along drho: Richardson-extrapolated central differences."""
    def d(e):
        return (_energy_at(o, rho + e * drho) - _energy_at(o, rho - e * drho)) / (2.0 * e)
    return (4.0 * d(0.5 * eps) - d(eps)) / 3.0


def _directions(rho, seed):
    g = np.random.default_rng(seed)
    x = np.arange(rho.shape[2])
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    c = rho.shape[2] / 2.0
    bump = np.exp(-((x3 - c - 2) ** 2 + (y3 - c) ** 2 + (z3 - c + 3) ** 2) / 8.0) * rho.max()
    return [rho * g.standard_normal(rho.shape), rho * (1.0 + 0.5 * g.standard_normal(rho.shape)), bump]


def _variational_residuals(o, prm):
    rho = o.dens(o.psi)
    o.den = rho
    o.poten()
    pot4 = o.pot4.copy()
    out = []
    for drho in _directions(rho, 11):
        lhs = _directional(o, rho, drho, 1e-3)
        rhs = float(np.sum(pot4 * drho)) * prm.dxyz
        scale = float(np.sum(np.abs(pot4 * drho))) * prm.dxyz
        out.append(abs(lhs - rhs) / scale)
    return out


# "dense" scales the droplet to a peak density of 0.33 A^-3, near the solid
# penalty's onset rho_m = 0.37 A^-3, where the penalty term is not negligible
@pytest.mark.parametrize("core,dense", [("OTC", False), ("OT", False), ("OTC", True)])
def test_pot4_is_the_functional_derivative_of_the_energy(t24, core, dense):
    prm, k, psi, r, v = t24
    prm = orc.Params(**{**prm.__dict__, "core": core})
    if dense:
        psi = psi * math.sqrt(0.33 / float(np.abs(psi).max() ** 2))
    o = _oracle(prm, k, psi, r, v)
    res = _variational_residuals(o, prm)
    assert max(res) < 1e-8, res


@pytest.mark.parametrize("plant,core,dense", [(FlippedAlpha, "OTC", False), (NoDen0s, "OTC", False),
                                              (BarePenalty, "OT", True), (UnsmoothedCorrelation, "OT", False),
                                              (DoubledLJ, "OT", False)])
def test_variational_identity_catches_planted_mistakes(t24, plant, core, dense):
    prm, k, psi, r, v = t24
    prm = orc.Params(**{**prm.__dict__, "core": core})
    if dense:
        psi = psi * math.sqrt(0.33 / float(np.abs(psi).max() ** 2))
    o = _oracle(prm, k, psi, r, v, cls=plant)
    res = _variational_residuals(o, prm)
    assert max(res) > 1e-5, (plant.__name__, res)


# ------------------------------------------------- bulk OT equation of state
RHO0 = 0.021836     # saturation density of the OT functional (A^-3)
E0 = -7.15          # energy per atom at saturation (K)


def _uniform(core="OT"):
    prm = orc.Params(nx=24, ny=24, nz=24, xmax=7.2, ymax=7.2, zmax=7.2, n_par=1, core=core)
    k = dict(fvlj4=orc.kernel_fvlj4(prm), wcgk=orc.kernel_wcgk(prm), kalfs=orc.kernel_kalfs(prm),
             vq=np.zeros(prm.hshape))
    o = orc.Oracle(prm, k)
    o.set_state(np.full(prm.shape, math.sqrt(RHO0), dtype=complex), np.zeros((3, 1)), np.zeros((3, 1)))
    o.den = o.dens(o.psi)
    o.potenimp()
    o.poten()
    return prm, o


def test_bulk_liquid_energy_per_atom_and_zero_pressure():
    prm, o = _uniform()
    e = o.energy()
    n_atoms = RHO0 * prm.nxyz * prm.dxyz
    e_per_atom = (e["elj4"] + e["ecor4"] + e["esolid"]) / n_atoms
    assert abs(e_per_atom - E0) < 5e-3, e_per_atom
    # zero pressure at saturation: mu = d(rho e)/drho = e; pot4 is mu in the bulk
    c = prm.nx // 2
    mu = o.pot4[c, c, c]
    assert abs(mu - E0) < 5e-3, mu
    assert np.allclose(o.pot4, mu, rtol=0, atol=1e-9)
    # planted: the 1/2 of E_LJ dropped, or c3/3 read as c3/2
    assert abs((e["elj4"] * 2 + e["ecor4"] + e["esolid"]) / n_atoms - E0) > 1.0
    cor_wrong = e["ecor4"] + RHO0 ** 4 * prm.cpp4 * (0.5 - 1.0 / 3.0) * prm.nxyz * prm.dxyz
    assert abs((e["elj4"] + cor_wrong + e["esolid"]) / n_atoms - E0) > 0.5


# -------------------------------------------- alpha_s energy of a Gaussian
def _alpha_closed_form(prm, amp, s):
    a = s * s + prm.l * prm.l / 4.0
    integral = amp * amp * (2.0 * math.pi) ** 3 * s ** 6 * 3.0 / (16.0 * math.pi ** 1.5) * a ** -2.5
    return -0.5 * prm.h2o2m4 * prm.alphas * integral


def test_alpha_energy_of_a_gaussian_density():
    prm = orc.Params(nx=48, ny=48, nz=48, xmax=7.2, ymax=7.2, zmax=7.2, n_par=1, den0s=1e30)
    k = dict(fvlj4=np.zeros(prm.hshape), wcgk=orc.kernel_wcgk(prm), kalfs=orc.kernel_kalfs(prm),
             vq=np.zeros(prm.hshape))
    amp, s = 0.01, 1.2
    x = prm.axis_grid("x")
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    rho = amp * np.exp(-(x3 ** 2 + y3 ** 2 + z3 ** 2) / (2.0 * s * s))
    o = orc.Oracle(prm, k)
    o.set_state(np.sqrt(rho).astype(complex), np.zeros((3, 1)), np.zeros((3, 1)))
    o.den = rho
    o.potenimp()
    o.poten()
    assert np.abs(o.falfs - 1.0).max() == 0.0
    got = o.energy()["ealphas"]
    want = _alpha_closed_form(prm, amp, s)
    assert abs(got - want) < 1e-6 * abs(want), (got, want)
    # the closed form tells a sign slip, a dropped 1/2 and a kernel of the
    # wrong width apart
    for wrong in (-want, 2.0 * want, _alpha_closed_form(orc.Params(l=2.0 * prm.l), amp, s)):
        assert abs(got - wrong) > 1e-2 * abs(want)


# ---------------------------------------------- forceimp and denpart pins
def _gaussian_impurity(prm, sig):
    x = prm.axis_grid("x")
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    denx = (2.0 * math.pi * sig * sig) ** -1.5 * np.exp(-(x3 ** 2 + y3 ** 2 + z3 ** 2) / (2.0 * sig * sig))
    return denx, (x3, y3, z3)


def _bohm_oracle(denx, prm, plant=None):
    k = dict(fvlj4=np.zeros(prm.hshape), wcgk=np.zeros(prm.hshape), kalfs=np.zeros(prm.hshape),
             vq=np.zeros(prm.hshape))
    o = orc.Oracle(prm, k)
    o.rimp = np.zeros((3, 1))
    o.set_denx(denx)
    if plant == "denx_not_psix2":
        o.denx = 2.0 * denx
    o.dxden = o.dyden = o.dzden = np.zeros(prm.shape)
    o.forceimp()
    q = [o.fx * prm.m_imp, o.fy * prm.m_imp, o.fz * prm.m_imp]
    if plant == "sign":
        q = [-c for c in q]
    return o, q


@pytest.mark.parametrize("plant", [None, "denx_not_psix2", "sign", "lap_without_z"])
def test_quantum_force_is_minus_grad_q_for_a_gaussian(plant, monkeypatch):
    # psi_X = sqrt(denx) falls to 6e-8 of its peak at the box faces, so the
    # zero padding of the stencils does not reach the quantum kinetic energy
    prm = orc.Params(nx=48, ny=48, nz=48, xmax=4.5, ymax=4.5, zmax=4.5, n_par=1)
    sig = 0.55
    denx, xyz = _gaussian_impurity(prm, sig)
    if plant == "lap_without_z":
        real = orc.deriv
        monkeypatch.setattr(orc, "deriv", lambda f, a, o_, h: np.zeros_like(f) if (a == "z" and o_ == 2)
                            else real(f, a, o_, h))
    o, q = _bohm_oracle(denx, prm, plant)
    inner = np.sqrt(sum(c * c for c in xyz)) < 2.0 * sig
    want_scale = prm.h2o2mx / (2.0 * sig ** 4)
    err = max(float(np.abs(q[c] - want_scale * xyz[c])[inner].max()) for c in range(3))
    ref = want_scale * 2.0 * sig
    if plant is None:
        assert err < 1e-6 * ref, err / ref
        # quantum kinetic energy of the Gaussian: h2o2mx * 3/(4 sigma^2)
        ekq = float(np.sum(o.psix * o.hpsix)) * prm.dxyz
        assert abs(ekq - prm.h2o2mx * 3.0 / (4.0 * sig * sig)) < 1e-8 * ekq
    else:
        assert err > 1e-2 * ref, (plant, err / ref)


def _third_law_residual(o, prm):
    """(sum denx fconv_c dV) + (force on the He along c) for c = x, y, z, over
    the scale of the terms."""
    out = []
    for c, (fc, ax, h) in enumerate(zip((o.fx, o.fy, o.fz), "xyz", (prm.hx, prm.hy, prm.hz))):
        q = np.zeros_like(fc)  # the quantum part is zero where denx is zero
        on_imp = float(np.sum(o.denx * (fc * prm.m_imp - q))) * prm.dxyz
        on_he = -float(np.sum(o.den * orc.deriv(o.uimp, ax, 1, h))) * prm.dxyz
        scale = float(np.sum(np.abs(o.den * orc.deriv(o.uimp, ax, 1, h)))) * prm.dxyz
        out.append(abs(on_imp + on_he) / scale)
    return out


@pytest.mark.parametrize("plant", [None, "plus_vq", "swapped_axes", "lj_kernel"])
def test_convolution_force_obeys_newtons_third_law(t24, plant):
    prm, k, psi, r, v = t24
    kk = dict(k)
    o = orc.Oracle(prm, kk)
    o.set_state(psi, r, v)
    o.den = o.dens(o.psi)
    o.potenimp()
    # keep the impurity density but zero the quantum part: psix = hpsix = 0
    # leaves fx = fconv / m (sto7 multiplies a zero bracket)
    o.psix = np.zeros(prm.shape)
    o.hpsix = np.zeros(prm.shape)
    o.poten()
    if plant == "plus_vq":
        o.vq = -k["vq"]
    elif plant == "lj_kernel":
        o.vq = k["fvlj4"]
    elif plant == "swapped_axes":
        o.dxden, o.dyden = o.dyden, o.dxden
    o.forceimp()
    res = _third_law_residual(o, prm)
    if plant is None:
        assert max(res) < 1e-12, res
    else:
        assert max(res) > 1e-3, (plant, res)


def test_denpart_normalisation_is_the_fraction_inside():
    prm = orc.Params(nx=24, ny=24, nz=24, xmax=6.0, ymax=6.0, zmax=6.0, n_par=1000, ironx=3)
    g = np.random.default_rng(5)
    r = 0.8 * g.standard_normal((3, prm.n_par))   # clear of the ironing boundary rows
    r[:, :37] = 9.0 + g.random((3, 37))           # outside the box
    r[0, 37:50] = -6.0 - 0.5 * prm.hx - 1e-9       # just outside at the low edge
    k = dict(fvlj4=np.zeros(prm.hshape), wcgk=np.zeros(prm.hshape), kalfs=np.zeros(prm.hshape),
             vq=np.zeros(prm.hshape))
    o = orc.Oracle(prm, k)
    o.set_state(np.zeros(prm.shape, complex), r, np.zeros_like(r))
    o.denpart()
    inside = prm.n_par - 50
    assert abs(float(np.sum(o.denx)) * prm.dxyz - inside / prm.n_par) < 1e-14
    assert (o.denx >= 0.0).all()


def test_energy_terms_sum_to_total():
    """bench.py measures the ensemble's mixed-vs-FP64 energy differences
    against the sum of the magnitudes of the terms whose sum is etot
    (energy.f90:62-124): those terms add up to etot, and the scale is not
    smaller than |etot|."""
    import bench
    prm, o = _uniform()
    o.rimp = np.array([[0.3], [0.1], [-0.2]])
    o.vimp = np.array([[1.0], [0.5], [2.0]])
    e = o.energy()
    keys = ("ekin4", "elj4", "ealphas", "esolid", "ecor4", "ekinx", "ekinqx", "ehex")
    assert abs(sum(e[k] for k in keys) - e["etot"]) <= 1e-12 * bench._term_scale(e)
    assert bench._term_scale(e) >= abs(e["etot"])
    assert bench._term_scale(e) == sum(abs(e[k]) for k in keys)
