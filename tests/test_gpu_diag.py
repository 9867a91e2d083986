"""Device-side diagnostics (hetddft_diagnostics, hetddft_current) against the
oracle on the device's own state after a few steps, in both precisions."""
import numpy as np
import pytest

from conftest import oracle_params, rel_max

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["t32", "c1"])
@pytest.mark.parametrize("precision", [0, 1])
def test_diagnostics_match_oracle(hip, name, precision):
    import inputs
    from oracle import hetddft_oracle as orc
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl, precision=precision)
    print(f"diagnostics {name} precision={precision}", flush=True)
    with hip.Instance(p) as g:
        g.set_kernels(hip.make_kernels(p))
        psi, r, v = inputs.initial_state(wl)
        g.set_state(psi, r, v)
        g.start()
        for _ in range(5):
            g.iterate(wl.deltat, want_errors=False)
        d = g.diagnostics()
        j = g.current()
        gpsi, gr, gv, ga = g.get_state()
        pot4 = g.get_field("pot4")
        denx = g.get_field("denx")
    prm = oracle_params(wl)
    ref = orc.diagnostics(prm, gpsi, pot4, denx, gv, ga)
    norm = ref[0]
    assert abs(d["norm"] - norm) < 1e-13 * norm
    assert np.abs(d["cm_moments"] - ref[1:4]).max() < 1e-12 * norm * wl.xmax
    assert abs(d["mu"] - ref[4]) < 1e-11 * abs(ref[4])
    jscale = np.abs(np.stack(orc.current(prm, gpsi))).sum() * prm.dxyz
    assert np.abs(d["jtot"] - ref[5:8]).max() < 1e-12 * jscale
    assert np.abs(d["msx"] - ref[8:11]).max() < 1e-13 * ref[8:11].max()
    assert np.array_equal(d["vmin"], ref[11:17:2]) and np.array_equal(d["vmax"], ref[12:17:2])
    assert np.array_equal(d["amin"], ref[17:23:2]) and np.array_equal(d["amax"], ref[18:23:2])
    for a, b in zip(j, orc.current(prm, gpsi)):
        assert rel_max(a, b) < 1e-12
