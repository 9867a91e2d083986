"""Failure path: a non-finite state is reported as HETDDFT_E_NONFINITE.

The reference's only health signal is the errHe/errimp/errvimp print of every
PC step (steppc.f90:109-112,133,147; DFT4He3d.f90:855-857).  The library
returns HETDDFT_E_NONFINITE when one of those estimates is not finite: from
the step that asks for them, from hetddft_last_errors, from hetddft_sync after
steps run without them, and from the ensemble executor.  The non-finite values
are planted from the host (set_state); every kernel runs to its normal end.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instance(hip, precision, plant=None):
    import inputs
    wl = inputs.WORKLOADS["t24"]
    p = inputs.params_for(wl, precision=precision)
    psi, r, v = inputs.initial_state(wl)
    psi, r = psi.copy(), r.copy()
    if plant == "psi":
        psi[wl.n // 2, wl.n // 2 + 1, wl.n // 2 - 1] = np.nan
    elif plant == "rimp":
        r[:, 3] = np.nan
    g = hip.Instance(p)
    g.set_kernels(hip.make_kernels(p))
    g.set_state(psi, r, v)
    g.start()
    return wl, g


def _expect_nonfinite(hip, fn):
    with pytest.raises(hip.HetddftError) as ei:
        fn()
    assert ei.value.code == hip.E_NONFINITE, ei.value
    assert "non-finite" in str(ei.value)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("plant", ["psi", "rimp"])
def test_step_reports_nonfinite(hip, precision, plant):
    print(f"failure path: NaN in {plant}, precision {precision}", flush=True)
    wl, g = _instance(hip, precision, plant)
    for _ in range(3):                        # RK start-up: no estimates
        assert g.iterate(wl.deltat) == (0.0, 0.0, 0.0)
    _expect_nonfinite(hip, lambda: g.iterate(wl.deltat))
    _expect_nonfinite(hip, g.last_errors)
    g.close()


@pytest.mark.parametrize("plant", ["psi", "rimp"])
def test_sync_reports_nonfinite_without_errors(hip, plant):
    print(f"failure path without errors: NaN in {plant}", flush=True)
    wl, g = _instance(hip, 0, plant)
    for _ in range(5):
        g.iterate(wl.deltat, want_errors=False)
    _expect_nonfinite(hip, g.sync)
    _expect_nonfinite(hip, g.last_errors)
    g.close()


def test_rimp_nan_is_dropped_from_the_histogram(hip):
    """A non-finite particle is dropped like one outside the box
    (denpart.f90:29-31): the He side and the other particles stay finite."""
    wl, g = _instance(hip, 0, "rimp")
    for _ in range(4):
        g.iterate(wl.deltat, want_errors=False)
    psi, r, v, a = g.get_state()
    assert np.isfinite(psi).all()
    assert np.isfinite(g.get_field("denx")).all()
    assert np.isfinite(g.get_field("pot4")).all()
    others = np.delete(r, 3, axis=1)
    assert np.isfinite(others).all() and not np.isfinite(r[:, 3]).any()
    g.close()


def test_finite_run_passes_sync(hip):
    wl, g = _instance(hip, 0)
    for _ in range(5):
        g.iterate(wl.deltat, want_errors=False)
    g.sync()
    e = g.last_errors()
    assert all(np.isfinite(e)) and e[0] > 0.0
    g.close()


def test_batch_reports_the_diverged_instance(hip):
    print("failure path: ensemble executor", flush=True)
    wl, good = _instance(hip, 0)
    _, bad = _instance(hip, 0, "psi")
    _, good2 = _instance(hip, 1)
    b = hip.Batch([good, bad, good2], [wl.deltat] * 3, nthreads=2)
    with pytest.raises(hip.HetddftError) as ei:
        b.iterate(steps=5, want_errors=False)
    assert ei.value.args[1] == hip.E_NONFINITE and ei.value.args[2] == 1, ei.value.args
    b.close()
    for g in (good, bad, good2):
        g.close()
