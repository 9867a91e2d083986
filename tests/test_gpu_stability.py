"""Long-run stability of the benchmark workload (128^3, N=500, 10^4 particles,
dt = 0.5 fs, He cavity around the impurity): 600 PC steps in both precisions
keep the PC error estimates finite and small, the total energy within 1% of
its start, the He density below the solid penalty's onset (den_m = 0.37), and
the two precisions' energies within 1e-6 of each other.  (Without the cavity
the same workload diverges near step 407.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_metric_workload_stays_stable(hip):
    import inputs
    wl = inputs.WORKLOADS["c2"]
    assert wl.cavity > 0.0
    psi, r, v = inputs.initial_state(wl)
    res = {}
    for precision in (0, 1):
        print(f"stability c2, precision {precision}", flush=True)
        p = inputs.params_for(wl, precision=precision)
        with hip.Instance(p) as g:
            g.init_kernels()
            g.set_state(psi, r, v)
            g.start()
            e0 = g.energy()["etot"]
            for it in range(3 + 600):
                g.iterate(wl.deltat, want_errors=False)
            errs = np.asarray(g.last_errors())
            e1 = g.energy()["etot"]
            den = g.get_field("den")
        assert np.all(np.isfinite(errs)) and errs[0] < 1e-8, errs
        assert abs(e1 - e0) < 0.01 * abs(e0), (e0, e1)
        assert den.max() < 0.2, den.max()
        res[precision] = e1
    assert abs(res[1] - res[0]) <= 1e-6 * abs(res[0]), res
