"""Kernels built on the device (hetddft_init_kernels) against the host
generator (hetddft_make_kernels, itself pinned against the oracle and the
reference's compiled fforma / initcg / FT_V_spline), and a trajectory run on
them."""
import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["t32", "c1"])
@pytest.mark.parametrize("precision", [0, 1])
def test_device_kernels_match_host(hip, name, precision):
    import inputs
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl, precision=precision)
    host = hip.make_kernels(p)
    print(f"device kernels {name} precision={precision}", flush=True)
    with hip.Instance(p) as g:
        g.init_kernels()
        dev = g.get_kernels()
    if precision == 1:
        host = {k: v.astype(np.float32).astype(np.float64) for k, v in host.items()}
    errs = {k: rel_max(dev[k], host[k]) for k in host}
    print("  rel_max", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert errs["wcgk"] < 1e-14 and errs["kalfs"] < 1e-14
    # fvlj4's bracket (sigma^6 sinx11 - sinx5) cancels more and more with p:
    # last-bit differences of the device sin/cos/pow grow there (as between the
    # host generator and the reference's S13ADF).  Tight where the density
    # spectrum lives, looser above 1/A.
    from conftest import oracle_params
    from oracle import hetddft_oracle as orc
    pm = orc.pmod(oracle_params(wl))
    scale = np.abs(host["fvlj4"]).max()
    d = np.abs(dev["fvlj4"] - host["fvlj4"])
    tight = 1e-12 if precision == 0 else 1.2e-7      # FP32 values: one ulp
    assert d[pm < 0.5].max() <= tight * scale, d[pm < 0.5].max() / scale
    assert errs["fvlj4"] < 1e-8
    # Vq is FP32-rounded: equal, or one FP32 ulp apart where the device sum
    # order moves a value across a rounding boundary
    ulp = np.spacing(np.abs(host["vq"]).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(dev["vq"] - host["vq"]) <= ulp)


def test_trajectory_on_device_kernels(hip):
    import inputs
    wl = inputs.WORKLOADS["t32"]
    p = inputs.params_for(wl)
    psi, r, v = inputs.initial_state(wl)
    out = []
    for dev in (False, True):
        with hip.Instance(p) as g:
            if dev:
                g.init_kernels()
            else:
                g.set_kernels(hip.make_kernels(p))
            g.set_state(psi, r, v)
            g.start()
            for _ in range(6):
                g.iterate(wl.deltat, want_errors=False)
            out.append(g.get_state())
    print("  trajectory psi / rimp rel_max", rel_max(out[1][0], out[0][0]), rel_max(out[1][1], out[0][1]), flush=True)
    assert rel_max(out[1][0], out[0][0]) < 1e-10
    assert rel_max(out[1][1], out[0][1]) < 1e-10
