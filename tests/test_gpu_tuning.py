"""Launch-shape tuning: the autotuner restores the state it timed on, and no
launch shape (stencil tile and z chunk per kernel class, z-transform ky tile)
changes any result -- the state after every step is bit-identical to the
default shapes' (the errHe estimate sums per-workgroup partials, so it is
compared to rounding)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(hip, wl, precision, kernels):
    import inputs
    p = inputs.params_for(wl, precision=precision)
    g = hip.Instance(p)
    g.set_kernels(kernels)
    psi, r, v = inputs.initial_state(wl)
    g.set_state(psi, r, v)
    g.start()
    return g


def _same_state(a, b, what):
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y), what
    for f in ("den", "pot4", "denx", "uimp", "falfs"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), (what, f)


@pytest.mark.parametrize("name", ["t32", "c2"])
@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_restores_state_and_keeps_results(hip, name, precision):
    import inputs
    wl = inputs.WORKLOADS[name]
    print(f"autotune {name} precision={precision}", flush=True)
    p = inputs.params_for(wl, precision=precision)
    k = hip.make_kernels(p)
    a = _make(hip, wl, precision, k)
    b = _make(hip, wl, precision, k)
    for _ in range(5):
        a.iterate(wl.deltat, want_errors=False)
        b.iterate(wl.deltat, want_errors=False)
    t = a.autotune(wl.deltat, steps=2)
    assert all(0 <= t.tile[c] <= 3 for c in range(5))
    assert all(1 <= t.zc[c] <= wl.n for c in range(5))
    assert all(t.prefetch[c] in (0, 1) for c in range(5))
    assert t.ms_per_step > 0.0
    assert t.fft_z_tile in ((8, 16, 32) if wl.n in (64, 128, 256) else (0,))
    assert 1 <= t.fft_z_wgs <= 8
    got = a.get_tuning()
    assert list(got.tile) == list(t.tile) and list(got.zc) == list(t.zc)
    assert list(got.prefetch) == list(t.prefetch) and got.fft_z_tile == t.fft_z_tile
    assert got.fft_z_wgs == t.fft_z_wgs
    _same_state(a, b, "after autotune")
    for it in range(4):
        want = it == 3
        ea = a.iterate(wl.deltat, want_errors=want)
        eb = b.iterate(wl.deltat, want_errors=want)
        _same_state(a, b, f"step {it}")
    np.testing.assert_allclose(ea, eb, rtol=1e-12, atol=0)
    assert a.energy() == b.energy()
    # a second call for the same grid, precision and N_par reuses the result
    t2 = a.autotune(wl.deltat, steps=2)
    assert list(t2.tile) == list(t.tile) and list(t2.zc) == list(t.zc) and t2.fft_z_tile == t.fft_z_tile
    a.close()
    b.close()


@pytest.mark.parametrize("tile", [0, 1, 2, 3])
@pytest.mark.parametrize("zc", [1, 5, 24])
@pytest.mark.parametrize("pf", [1, 0])
@pytest.mark.parametrize("precision", [0, 1])
def test_every_stencil_shape_gives_the_same_steps(hip, tile, zc, pf, precision):
    """24^3: z chunks that do not divide nz, tiles wider than the grid; RK
    start-up and PC steps under one shape against the default shapes."""
    import inputs
    wl = inputs.WORKLOADS["t24"]
    print(f"shape tile={tile} zc={zc} prefetch={pf} precision={precision}", flush=True)
    p = inputs.params_for(wl, precision=precision)
    k = hip.make_kernels(p)
    a = _make(hip, wl, precision, k)
    b = _make(hip, wl, precision, k)
    t = b.get_tuning()
    for c in range(5):
        t.tile[c] = tile
        t.zc[c] = zc
        t.prefetch[c] = pf
    b.set_tuning(t)
    for it in range(3 + 3):
        a.iterate(wl.deltat, want_errors=False)
        b.iterate(wl.deltat, want_errors=False)
        _same_state(a, b, f"step {it}")
    np.testing.assert_allclose(a.last_errors(), b.last_errors(), rtol=1e-12, atol=0)
    a.close()
    b.close()


def test_set_tuning_rejects_what_the_grid_does_not_offer(hip):
    import inputs
    for name, z_ok in (("t32", ()), ("c1", (8, 16, 32))):
        wl = inputs.WORKLOADS[name]
        p = inputs.params_for(wl, precision=0)
        g = hip.Instance(p)
        base = g.get_tuning()
        assert base.fft_z_tile == (32 if z_ok else 0)
        for field, val in (("tile", 4), ("tile", -1), ("zc", 0), ("prefetch", 2)):
            t = g.get_tuning()
            getattr(t, field)[2] = val
            with pytest.raises(hip.HetddftError):
                g.set_tuning(t)
        for wgs, ok in ((9, False), (-1, False), (2, True)):
            t = g.get_tuning()
            t.fft_z_wgs = wgs
            if ok:
                g.set_tuning(t)
                assert g.get_tuning().fft_z_wgs == wgs
            else:
                with pytest.raises(hip.HetddftError):
                    g.set_tuning(t)
        for w in (8, 12, 16, 32):
            t = g.get_tuning()
            t.fft_z_tile = w
            if w in z_ok:
                g.set_tuning(t)
                assert g.get_tuning().fft_z_tile == w
            else:
                with pytest.raises(hip.HetddftError):
                    g.set_tuning(t)
        g.close()
