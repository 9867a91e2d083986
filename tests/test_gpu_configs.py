"""The benchmark configurations at their real sizes on the GPU.

* 64^3 (N=100, 10^3 particles) and 128^3 (N=500, 10^4 particles, the metric
  workload), FP64: the HIP step against the oracle after every iteration of
  the reference loop (start, 3 RK-Gill steps, then PC steps; steppc.f90:1-155,
  DFT4He3d.f90:849-878): psi, den, pot4, r, v, a and the PC error estimates to
  1e-10 of each field's maximum.
* 64^3 mixed (FP32 convolutions and potential fields) against the FP64
  oracle over RK start-up and 100 PC steps: relative energy, norm and
  density-L2 errors every 10 PC steps.
* 256^3 (N=2000, 10^5 particles), both precisions: start, 3 RK and 5 PC
  steps with finite error estimates, the energy drift bounded, mixed within
  1e-6 of FP64, and the device memory of the context.
* the 32-instance ensemble (grids 48^3-96^3) through hetddft_batch_iterate:
  every instance bit-identical to its solo run, mixed within 1e-6 of FP64 per
  instance.

The oracle runs on the host's threads (oracle/threaded.py, the same values as
the single-threaded oracle).
"""
import os

import numpy as np
import pytest

from conftest import oracle_params, rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _pair(hip, name, precision=0):
    import inputs
    from oracle import hetddft_oracle as orc
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl, precision=precision)
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    o = orc.Oracle(oracle_params(wl), k)
    o.set_state(psi, r, v)
    g = hip.Instance(p)
    g.set_kernels(k)
    g.set_state(psi, r, v)
    return wl, o, g


def _compare(g, o, it, eo, eg):
    psi, r, v, a = g.get_state()
    assert rel_max(psi, o.psi) < TOL, (it, rel_max(psi, o.psi))
    assert rel_max(g.get_field("den"), o.den) < TOL, it
    assert rel_max(g.get_field("pot4"), o.pot4) < TOL, it
    assert rel_max(r, o.rimp) < TOL, it
    assert rel_max(v, o.vimp) < TOL, it
    assert rel_max(a, o.aimp) < TOL, it
    if eo is not None:
        scales = (np.abs(o.psi).max(), np.abs(o.rimp).max(), np.abs(o.vimp).max())
        for e_g, e_o, sc in zip(eg, eo, scales):
            assert abs(e_g - e_o) <= 1e-9 * abs(e_o) + 1e-15 * sc, (it, eg, eo)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,n_pc", [("c1", 5), ("c2", 2)])
def test_fp64_step_matches_the_oracle_at_config_size(hip, name, n_pc):
    from oracle import hetddft_oracle as orc
    from oracle.threaded import ThreadedOracle
    wl, o, g = _pair(hip, name)
    print(f"FP64 vs oracle at {name} ({wl.n}^3): start + 3 RK + {n_pc} PC", flush=True)
    with ThreadedOracle(orc, _workers()):
        o.start()
        g.start()
        _compare(g, o, -1, None, None)
        for it in range(3 + n_pc):
            eo = o.iterate(it, wl.deltat)
            eg = g.iterate(wl.deltat)
            _compare(g, o, it, eo, eg)
            print(f"  iteration {it} ok", flush=True)
        np.testing.assert_allclose(g.particle_means(), o.particle_means(), rtol=1e-10, atol=1e-12)
    g.close()


@pytest.mark.timeout(900)
def test_mixed_against_the_fp64_oracle_at_64cubed(hip):
    """Config 3, the mixed form of config 1, over config 1's whole run (RK
    start-up, then 100 PC steps) against the FP64 oracle: energy, norm and
    density-L2 errors every 10 PC steps, each within relative energy error
    1e-6, norm 1e-9 and density L2 1e-6 (energy.f90:12-129)."""
    from oracle import hetddft_oracle as orc
    from oracle.threaded import ThreadedOracle
    wl, o, g = _pair(hip, "c1", precision=1)
    print("mixed vs FP64 oracle at c1 (64^3), 100 PC steps", flush=True)
    with ThreadedOracle(orc, _workers()):
        o.start()
        g.start()
        for it in range(3 + 100):
            o.iterate(it, wl.deltat)
            g.iterate(wl.deltat, want_errors=False)
            if it > 2 and (it - 2) % 10 == 0:
                eo = o.energy()
                eg = g.energy()
                rel_e = abs(eg["etot"] - eo["etot"]) / abs(eo["etot"])
                rel_n = abs(eg["norm"] - eo["norm"]) / eo["norm"]
                d = g.get_field("den")
                rel_d = float(np.linalg.norm(d - o.den) / np.linalg.norm(o.den))
                print(f"  PC step {it - 2}: energy {rel_e:.2e}, norm {rel_n:.2e}, density L2 {rel_d:.2e}",
                      flush=True)
                assert rel_e <= 1e-6, rel_e
                assert rel_n <= 1e-9, rel_n
                assert rel_d <= 1e-6, rel_d
    g.close()


@pytest.mark.timeout(600)
def test_256cubed_steps_in_both_precisions(hip):
    import inputs
    wl = inputs.WORKLOADS["c5"]
    psi, r, v = inputs.initial_state(wl)
    res = {}
    for precision in (0, 1):
        print(f"256^3 N=2000 10^5 particles, precision {precision}", flush=True)
        p = inputs.params_for(wl, precision=precision)
        with hip.Instance(p) as g:
            g.init_kernels()
            g.set_state(psi, r, v)
            g.start()
            e0 = g.energy()
            for _ in range(3):
                assert g.iterate(wl.deltat) == (0.0, 0.0, 0.0)
            for _ in range(5):
                errs = g.iterate(wl.deltat)
                assert all(np.isfinite(errs)) and errs[0] < 1e-8, errs
            e1 = g.energy()
            dev = g.device_bytes
            den = g.get_field("den")
        assert abs(e1["etot"] - e0["etot"]) < 1e-3 * abs(e0["etot"]), (e0["etot"], e1["etot"])
        assert abs(e1["norm"] - e0["norm"]) < 1e-9 * e0["norm"]
        assert np.isfinite(den).all() and den.max() < 0.1
        # psi with 3 history slots, 2 H psi slots, pc, q, spare (8 complex
        # FP64 fields) dominate; the whole context stays below 10 GB
        assert 8 * 16 * 256 ** 3 < dev < 10e9, dev
        res[precision] = e1["etot"]
    assert abs(res[1] - res[0]) <= 1e-6 * abs(res[0]), res


def _energy_scale(e):
    """Sum of the magnitudes of the energy terms (K): the size of what cancels
    in the total."""
    return sum(abs(e[k]) for k in ("ekin4", "elj4", "ealphas", "esolid", "ecor4", "ekinx", "ekinqx", "ehex"))


@pytest.mark.timeout(900)
def test_ensemble_c4_batch_matches_solo(hip):
    """The 32-instance ensemble through the batch executor ends bit-identical
    to each instance run alone, in both precisions; mixed agrees with FP64
    per instance to 1e-6 of the energy scale (sum of term magnitudes: the
    total nearly cancels in the smallest droplets)."""
    import inputs
    wls = inputs.ensemble_c4()
    nsteps = 3 + 4
    etot = {}
    for precision in (0, 1):
        print(f"ensemble c4, precision {precision}: 32 instances, {nsteps} iterations", flush=True)
        kc = {}
        insts = []
        for wl in wls:
            p = inputs.params_for(wl, precision=precision)
            if wl.n not in kc:
                kc[wl.n] = hip.make_kernels(p)
            g = hip.Instance(p)
            g.set_kernels(kc[wl.n])
            psi, r, v = inputs.initial_state(wl)
            g.set_state(psi, r, v)
            g.start()
            insts.append(g)
        b = hip.Batch(insts, [wl.deltat for wl in wls])
        err = b.iterate(nsteps, want_errors=True)
        assert np.isfinite(err).all() and (err[:, 0] > 0).all()
        b.close()
        batch_states = [g.get_state() for g in insts]
        es = [g.energy() for g in insts]
        etot[precision] = es
        for g in insts:
            g.close()
        for i, wl in enumerate(wls):          # each instance alone
            p = inputs.params_for(wl, precision=precision)
            with hip.Instance(p) as g:
                g.set_kernels(kc[wl.n])
                psi, r, v = inputs.initial_state(wl)
                g.set_state(psi, r, v)
                g.start()
                for _ in range(nsteps):
                    g.iterate(wl.deltat, want_errors=False)
                for x, y in zip(g.get_state(), batch_states[i]):
                    assert np.array_equal(x, y), wl.name
    for wl, a, m in zip(wls, etot[0], etot[1]):
        assert abs(m["etot"] - a["etot"]) <= 1e-6 * _energy_scale(a), (wl.name, a["etot"], m["etot"])
