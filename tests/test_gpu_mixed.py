"""Mixed precision (FP32 convolutions and potential fields) against the FP64 path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(hip, name, precision, nsteps):
    import inputs
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl, precision=precision)
    k = hip.make_kernels(p)
    psi, r, v = inputs.initial_state(wl)
    with hip.Instance(p) as g:
        g.set_kernels(k)
        g.set_state(psi, r, v)
        g.start()
        e0 = g.energy()
        for _ in range(nsteps):
            g.iterate(wl.deltat, want_errors=False)
        e1 = g.energy()
        den = g.get_field("den")
        psi1, r1, v1, a1 = g.get_state()
    return e0, e1, den, psi1, r1


@pytest.mark.parametrize("name,nsteps", [("t32", 8), ("c1", 13), ("c2", 23)])
def test_mixed_energy_within_1e6(hip, name, nsteps):
    print(f"mixed vs fp64 {name} {nsteps} steps", flush=True)
    a0, a1, da, pa, ra = _run(hip, name, 0, nsteps)
    b0, b1, db, pb, rb = _run(hip, name, 1, nsteps)
    for ea, eb in ((a0, b0), (a1, b1)):
        rel = abs(eb["etot"] - ea["etot"]) / abs(ea["etot"])
        assert rel <= 1e-6, (name, rel, ea["etot"], eb["etot"])
        assert abs(eb["norm"] - ea["norm"]) <= 1e-9 * ea["norm"]
    assert np.linalg.norm(db - da) / np.linalg.norm(da) < 1e-6
    assert np.abs(rb - ra).max() < 1e-6
