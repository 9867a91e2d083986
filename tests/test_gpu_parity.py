"""FP64 parity of the HIP time step with the CPU oracle, per step (steprk/steppc + Veval)."""
import numpy as np
import pytest

from conftest import oracle_params, rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-10   # FP64 parity bar: relative to the field's max magnitude


def _setup(hip, name, gkw=None, okw=None):
    import inputs
    from oracle import hetddft_oracle as orc
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl, **(gkw or {}))
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    o = orc.Oracle(oracle_params(wl, **(okw or {})), k)
    o.set_state(psi, r, v)
    g = hip.Instance(p)
    g.set_kernels(k)
    g.set_state(psi, r, v)
    return wl, o, g


FIELD_MAP = {"den": "den", "uimp": "uimp", "dencg": "dencg", "delj4": "delj4", "falfs": "falfs",
             "dxden": "dxden", "dyden": "dyden", "dzden": "dzden", "intxalf": "intxalf",
             "intyalf": "intyalf", "intzalf": "intzalf", "denx": "denx", "psix": "psix",
             "hpsix": "hpsix", "pot4": "pot4"}


@pytest.mark.parametrize("name", ["t24", "t32"])
def test_potential_fields(hip, name):
    wl, o, g = _setup(hip, name)
    print(f"potential fields {name}", flush=True)
    o.start()
    g.start()
    for gname, oname in FIELD_MAP.items():
        a = g.get_field(gname)
        b = getattr(o, oname)
        assert rel_max(a, b) < TOL, (gname, rel_max(a, b))
    _, _, _, acc = g.get_state()
    assert rel_max(acc, o.aimp) < TOL
    g.close()


@pytest.mark.parametrize("name", ["t24", "t32"])
def test_trajectory(hip, name):
    wl, o, g = _setup(hip, name)
    print(f"trajectory {name}", flush=True)
    o.start()
    g.start()
    for it in range(3 + wl.n_pc):
        eo = o.iterate(it, wl.deltat)
        eg = g.iterate(wl.deltat)
        psi, r, v, a = g.get_state()
        assert rel_max(psi, o.psi) < TOL, (it, rel_max(psi, o.psi))
        assert rel_max(r, o.rimp) < TOL, it
        assert rel_max(v, o.vimp) < TOL, it
        assert rel_max(a, o.aimp) < TOL, it
        assert rel_max(g.get_field("pot4"), o.pot4) < TOL, it
        if eo is not None:
            # the estimates are |p - c| residues of O(1) quantities: compare
            # them on the scale of psi, r and v
            scales = (np.abs(o.psi).max(), np.abs(o.rimp).max(), np.abs(o.vimp).max())
            for e_g, e_o, sc in zip(eg, eo, scales):
                assert abs(e_g - e_o) <= 1e-9 * abs(e_o) + 1e-15 * sc, (it, eg, eo)
        np.testing.assert_allclose(g.particle_means(), o.particle_means(), rtol=1e-10, atol=1e-12)
    g.close()


def test_energy(hip):
    wl, o, g = _setup(hip, "t32")
    o.start()
    g.start()
    for it in range(4):
        o.iterate(it, wl.deltat)
        g.iterate(wl.deltat)
    eo = o.energy()
    eg = g.energy()
    for key, val in eo.items():
        assert abs(eg[key] - val) <= 1e-10 * max(1.0, abs(val)), (key, eg[key], val)
    g.close()


def test_alpha_off_core_ot(hip):
    """core4='OT ' (no alpha_s term, poten.f90:94-105 skipped)."""
    wl, o, g = _setup(hip, "t24", gkw=dict(core_otc=0), okw=dict(core="OT"))
    o.start()
    g.start()
    assert rel_max(g.get_field("pot4"), o.pot4) < TOL
    for it in range(4):
        o.iterate(it, wl.deltat)
        g.iterate(wl.deltat)
    psi, r, v, a = g.get_state()
    assert rel_max(psi, o.psi) < TOL
    g.close()


def test_timec_lambda(hip):
    """Non-zero absorbing time constant timec = (Lambda(1+tanh((z-zm)/zs)), 1) (DFT4He3d.f90:773-782)."""
    tc = dict(lambdah=0.3, tzmean=2.0, tzsurf=1.5)
    wl, o, g = _setup(hip, "t24", gkw=tc, okw=tc)
    o.start()
    g.start()
    for it in range(5):
        o.iterate(it, wl.deltat)
        g.iterate(wl.deltat)
    psi, r, v, a = g.get_state()
    assert rel_max(psi, o.psi) < TOL
    g.close()


@pytest.mark.parametrize("tile,zc,xcd", [("1", "5", "1"), ("2", "4", "1"), ("3", "7", "0"), ("0", "3", "0")])
@pytest.mark.parametrize("name", ["t24", "t32"])
def test_trajectory_tile_shapes(hip, name, tile, zc, xcd, monkeypatch):
    """Every tile shape of the z-marching stencil engine, z chunks that do not
    divide nz, partial x tiles (24 < 32) and both block orders give the same
    trajectory as the oracle."""
    monkeypatch.setenv("HETDDFT_TILE", tile)
    monkeypatch.setenv("HETDDFT_ZC", zc)
    monkeypatch.setenv("HETDDFT_XCD", xcd)
    wl, o, g = _setup(hip, name)
    print(f"trajectory {name} tile {tile} zc {zc} xcd {xcd}", flush=True)
    o.start()
    g.start()
    for it in range(3 + 2):
        o.iterate(it, wl.deltat)
        eg = g.iterate(wl.deltat)
        psi, r, v, a = g.get_state()
        assert rel_max(psi, o.psi) < TOL, (it, rel_max(psi, o.psi))
        assert rel_max(r, o.rimp) < TOL, it
        assert rel_max(g.get_field("pot4"), o.pot4) < TOL, it
    g.close()


@pytest.mark.parametrize("cloud", ["cluster_and_strays", "uniform"])
def test_force_gather_any_cloud(hip, cloud):
    """The force gather evaluates F_c once per point of the box spanned by the
    particles' cell corners when that box holds at most 8 points per
    particle, and at each particle's own corners otherwise.  A tight cluster
    with a few particles outside the grid (zero force, potenimp.f90:104-105)
    takes the first path, a cloud spread over the whole box the second; both
    match the oracle over start-up and PC steps."""
    import inputs
    from oracle import hetddft_oracle as orc
    wl = inputs.WORKLOADS["t32"]
    p = inputs.params_for(wl)
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    rng = np.random.default_rng(11)
    if cloud == "uniform":
        r = rng.uniform(-0.95 * wl.xmax, 0.95 * wl.xmax, size=r.shape)
    else:
        r = 0.4 * rng.standard_normal(r.shape)
        r[0, :3] = 1.2 * wl.xmax      # beyond the last cell in x
        r[1, 3] = -1.1 * wl.xmax      # before the first cell in y
        r[2, 4] = wl.xmax - 0.5 * wl.h  # past the last grid point: outside [1, n-1]
    o = orc.Oracle(oracle_params(wl), k)
    o.set_state(psi, r, v)
    g = hip.Instance(p)
    g.set_kernels(k)
    g.set_state(psi, r, v)
    print(f"force gather {cloud}", flush=True)
    o.start()
    g.start()
    _, _, _, a = g.get_state()
    assert rel_max(a, o.aimp) < TOL
    if cloud != "uniform":
        assert np.all(a[:, :5] == 0.0)
    for it in range(3 + 2):
        o.iterate(it, wl.deltat)
        g.iterate(wl.deltat)
        psi_g, r_g, v_g, a = g.get_state()
        assert rel_max(a, o.aimp) < TOL, it
        assert rel_max(r_g, o.rimp) < TOL, it
        assert rel_max(psi_g, o.psi) < TOL, it
    g.close()


def test_force_fields_in_gather_box(hip):
    """The force convolutions -V*d_c den (potenimp.f90:39-50) are read only
    by the trilinear gather at the particles' cell corners (interpola3,
    potenimp.f90:71-116), so a potential evaluation produces their z planes
    that hold such corners: there they equal the oracle's full transforms,
    and the other planes keep the zeros of a fresh context."""
    from oracle import hetddft_oracle as orc
    wl, o, g = _setup(hip, "t32")
    print("force fields in the gather box", flush=True)
    o.start()
    g.start()
    p = o.p
    r = o.rimp
    a = [(r[0] + p.xmax) / p.hx + 1, (r[1] + p.ymax) / p.hy + 1, (r[2] + p.zmax) / p.hz + 1]
    ok = ((a[0] >= 1) & (a[0] < p.nx) & (a[1] >= 1) & (a[1] < p.ny) & (a[2] >= 1) & (a[2] < p.nz))
    iz = np.trunc(a[2][ok]).astype(np.int64)
    z0, z1 = int(iz.min()) - 1, int(iz.max())   # corners iz-1 and iz, 0-based
    assert 0 < z0 and z1 < p.nz - 1
    for name, d in (("fxconv", o.dxden), ("fyconv", o.dyden), ("fzconv", o.dzden)):
        full = orc.c2r(-(orc.r2c(d) * o.vq), p.shape, p.renor)
        f = g.get_field(name)
        assert rel_max(f[z0:z1 + 1], full[z0:z1 + 1]) < TOL, name
        assert np.all(f[:z0] == 0.0) and np.all(f[z1 + 1:] == 0.0), name
    g.close()


@pytest.mark.parametrize("ironx", [0, 1, 2, 3, 4])
def test_ironing_passes(hip, ironx):
    """ironx smoothing passes of the particle histogram (ironing.f90): the
    fields of the particle chain and the force match the oracle for every
    pass count (1-3 run in one launch, 0 and 4 one launch per pass)."""
    wl, o, g = _setup(hip, "t32", gkw=dict(ironx=ironx), okw=dict(ironx=ironx))
    print(f"ironing ironx={ironx}", flush=True)
    o.start()
    g.start()
    for f in ("denx", "psix", "hpsix", "uimp"):
        assert rel_max(g.get_field(f), getattr(o, f)) < TOL, f
    for it in range(3 + 1):
        o.iterate(it, wl.deltat)
        g.iterate(wl.deltat)
    psi, r, v, a = g.get_state()
    assert rel_max(psi, o.psi) < TOL
    assert rel_max(a, o.aimp) < TOL
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("ironx", [1, 2, 3])
def test_fused_ironing_equals_one_launch_per_pass(hip, monkeypatch, precision, ironx):
    """The fused ironing launch gives the same bits as one launch per pass,
    in both precisions, over the start-up and a PC step (the work region
    moves with the particles)."""
    import inputs
    wl = inputs.WORKLOADS["t32"]
    psi, r, v = inputs.initial_state(wl)
    p = inputs.params_for(wl, precision=precision, ironx=ironx)
    k = hip.make_kernels(p)
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("HETDDFT_IRON_FUSED", fused)
        print(f"ironing fused={fused} ironx={ironx} precision={precision}", flush=True)
        g = hip.Instance(p)
        g.set_kernels(k)
        g.set_state(psi, r, v)
        g.start()
        for _ in range(3 + 2):
            g.iterate(wl.deltat, want_errors=False)
        out.append([g.get_field(f) for f in ("denx", "psix", "hpsix")] + list(g.get_state()))
        g.close()
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_error_estimates_after_other_calls(hip):
    """The PC error estimates are reduced from device partials only when
    asked for.  Energy, diagnostics and particle-means calls made after a
    step reuse the partials buffer; the estimates read after them equal
    those returned by the step itself (a twin instance asks every step)."""
    import inputs
    wl = inputs.WORKLOADS["t24"]
    p = inputs.params_for(wl)
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    inst = []
    for _ in range(2):
        g = hip.Instance(p)
        g.set_kernels(k)
        g.set_state(psi, r, v)
        g.start()
        inst.append(g)
    a, b = inst
    print("error estimates after other calls", flush=True)
    for it in range(3 + 4):
        ea = a.iterate(wl.deltat, want_errors=True)
        b.iterate(wl.deltat, want_errors=False)
        if it >= 3:
            b.energy()
            b.diagnostics()
            b.particle_means()
            eb = b.last_errors()
            assert np.array_equal(np.asarray(ea), np.asarray(eb)), (it, ea, eb)
            assert np.array_equal(np.asarray(b.last_errors()), np.asarray(eb))
    a.close()
    b.close()


def test_particle_means_with_error_fetch(hip):
    """A host that fetches the PC error estimates and the particle means every
    step (DFT4He3d.f90:855-878) gets the means reduced in the same GPU round
    trip as the estimates; they equal the means of a twin instance that asks
    only for the means, and a state change in between (set_state + start, a
    checkpoint load) is never answered from the earlier reduction."""
    import inputs
    wl = inputs.WORKLOADS["t24"]
    p = inputs.params_for(wl)
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    a, b = hip.Instance(p), hip.Instance(p)
    for g in (a, b):
        g.set_kernels(k)
        g.set_state(psi, r, v)
        g.start()
    print("particle means with error fetch", flush=True)
    for it in range(3 + 4):
        a.iterate(wl.deltat, want_errors=True)
        b.iterate(wl.deltat, want_errors=False)
        ma, mb = a.particle_means(), b.particle_means()
        assert np.array_equal(ma, mb), (it, ma, mb)
    blob = a.checkpoint()
    r2 = r + 0.25
    a.set_state(psi, r2, v)
    a.start()
    c = hip.Instance(p)
    c.set_kernels(k)
    c.set_state(psi, r2, v)
    c.start()
    assert np.array_equal(a.particle_means(), c.particle_means())
    a.iterate(wl.deltat, want_errors=True)
    a.particle_means()
    a.restore(blob)
    assert np.array_equal(a.particle_means(), b.particle_means())
    for g in (a, b, c):
        g.close()

