"""Fresh starts built on the device (hetddft_init_droplet; mode 0 of
DFT4He3d.f90:585-605 with a closed-form density) against the oracle's
droplet_state:

* the workloads' droplets (64^3 with cavity, 128^3 with phase noise and a
  moving cloud, a rippled surface, a plane-wave boost): psi to 1e-12 of its
  maximum, positions to 1e-13, velocities exactly;
* a context started from the device-built state steps exactly as one
  started from the same state uploaded with hetddft_set_state;
* sigma <= 0 keeps the positions; invalid parameters are HETDDFT_E_ARG;
* the 32-instance ensemble set up on the device (one instance per grid size
  against the oracle) advances through the batch executor.
"""
import numpy as np
import pytest

from conftest import oracle_params, rel_max

pytestmark = pytest.mark.gpu


def _spec_wl(name):
    import inputs
    if name == "ripple":
        return inputs.Workload("ripple", 64, 0.45, 100, 2000, 1.0, v0=2.0, ripple=0.3, seed=5, cavity=4.2)
    return inputs.WORKLOADS[name]


def _compare_with_oracle(hip, wl, precision=0, lamda=(0.0, 0.0, 0.0)):
    import inputs
    from oracle import hetddft_oracle as orc
    spec = inputs.droplet_spec(wl)
    p = inputs.params_for(wl, precision=precision)
    prm = oracle_params(wl)
    with hip.Instance(p) as g:
        g.init_droplet(hip.droplet(lamda=lamda, **spec))
        psi, r, v, _ = g.get_state()
    dens = {k: spec[k] for k in ("cavity", "cavity_w", "ripple", "clm")}
    opsi, orimp, ovimp = orc.droplet_state(prm, spec["n_he"], spec["seed"], lamda=lamda,
                                           phase_noise=spec["phase_noise"], sigma=spec["sigma"], v0=spec["v0"],
                                           **dens)
    assert rel_max(psi, opsi) < 1e-12, rel_max(psi, opsi)
    assert rel_max(r, orimp) < 1e-13, rel_max(r, orimp)
    np.testing.assert_array_equal(v, ovimp)
    assert abs((np.abs(psi) ** 2).sum() * prm.dxyz - spec["n_he"]) < 1e-9 * spec["n_he"]


@pytest.mark.parametrize("name", ["t32", "c1", "c2", "ripple"])
def test_device_start_matches_the_oracle(hip, name):
    wl = _spec_wl(name)
    print(f"device droplet start vs oracle: {name} ({wl.n}^3)", flush=True)
    _compare_with_oracle(hip, wl)


def test_plane_wave_boost(hip):
    print("device droplet start with a plane-wave boost", flush=True)
    _compare_with_oracle(hip, _spec_wl("t32"), lamda=(0.4, -0.3, 0.25))


@pytest.mark.parametrize("precision", [0, 1])
def test_steps_from_device_start_equal_uploaded_start(hip, precision):
    import inputs
    wl = inputs.WORKLOADS["t32"]
    print(f"steps from a device start vs the same state uploaded, precision {precision}", flush=True)
    p = inputs.params_for(wl, precision=precision)
    k = hip.make_kernels(p)
    with hip.Instance(p) as a, hip.Instance(p) as b:
        a.set_kernels(k)
        b.set_kernels(k)
        a.init_droplet(hip.droplet(**inputs.droplet_spec(wl)))
        psi, r, v, _ = a.get_state()
        b.set_state(psi, r, v)
        a.start()
        b.start()
        for _ in range(3 + 4):
            ea = a.iterate(wl.deltat)
            eb = b.iterate(wl.deltat)
            assert ea == eb
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y)


def test_sigma_zero_keeps_positions_and_bad_parameters_fail(hip):
    import hetddft
    import inputs
    wl = inputs.WORKLOADS["t24"]
    psi, r, v = inputs.initial_state(wl)
    with hip.Instance(inputs.params_for(wl)) as g:
        g.set_state(psi, r, v)
        g.init_droplet(hip.droplet(n_he=12.0, seed=3, sigma=0.0, v0=(0.0, 1.0, 0.0)))
        _, r2, v2, _ = g.get_state()
        assert np.array_equal(r2, r)
        np.testing.assert_array_equal(v2[1], 7.63822291)
        np.testing.assert_array_equal(v2[0], 0.0)
        for bad in (dict(n_he=0.0), dict(n_he=12.0, a=0.0), dict(n_he=12.0, ripple=-1.0),
                    dict(n_he=12.0, cavity=2.0, cavity_w=0.0), dict(n_he=12.0, phase_noise=-0.1)):
            with pytest.raises(hetddft.HetddftError) as ei:
                g.init_droplet(hip.droplet(seed=1, **bad))
            assert ei.value.code == hetddft.E_ARG


@pytest.mark.timeout(600)
def test_ensemble_set_up_on_the_device(hip):
    """The c4 ensemble with every instance built on the device: one instance
    per grid size against the oracle, then RK start-up and PC steps through
    the batch executor with finite estimates and the norm kept."""
    import inputs
    from oracle import hetddft_oracle as orc
    wls = inputs.ensemble_c4()
    print("ensemble c4 set up on the device: 32 instances", flush=True)
    kc, insts = {}, []
    for wl in wls:
        p = inputs.params_for(wl, precision=1)
        if wl.n not in kc:
            kc[wl.n] = hip.make_kernels(p)
        g = hip.Instance(p)
        g.set_kernels(kc[wl.n])
        g.init_droplet(hip.droplet(**inputs.droplet_spec(wl)))
        insts.append(g)
    for i in range(4):                      # grids 48, 64, 80, 96
        wl = wls[i]
        spec = inputs.droplet_spec(wl)
        opsi, orimp, _ = orc.droplet_state(oracle_params(wl), spec["n_he"], spec["seed"], cavity=spec["cavity"],
                                           cavity_w=spec["cavity_w"], phase_noise=spec["phase_noise"],
                                           sigma=spec["sigma"], v0=spec["v0"])
        psi, r, _, _ = insts[i].get_state()
        assert rel_max(psi, opsi) < 1e-12 and rel_max(r, orimp) < 1e-13, wl.name
    for g in insts:
        g.start()
    norms = [g.energy()["norm"] for g in insts]
    b = hip.Batch(insts, [wl.deltat for wl in wls])
    err = b.iterate(3 + 4, want_errors=True)
    b.close()
    assert np.isfinite(err).all() and (err[:, 0] > 0).all()
    for g, n0, wl in zip(insts, norms, wls):
        assert abs(g.energy()["norm"] - n0) < 1e-9 * n0, wl.name
        g.close()
