"""Non-cubic grids: every compile-time transform length outside the whole-plane
kernels (x rows of 256 points, y and z lines of 256), odd mixes of lengths and
the 128 x 128 plane kernels, against the oracle in FP64 and against the FFT
definition in both precisions.

These grids reach the separate x / y / z passes that a 256^3 grid uses (a
256^3 plane does not fit in LDS), at sizes the oracle finishes in seconds.
"""
import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-10

# (nx, ny, nz): x rows of 256 (half length 128), y lines of 256, z lines of 256,
# and 128 x 128 planes (the register-resident whole-plane kernels of the
# 128^3 workloads) with every producer and epilogue of the potential; and
# lengths with no compile-time kernel (x half length 18, y 40, z 30: the
# run-time radix plans of every pass, including the force convolutions'
# plane selection in the run-time z and inverse passes)
GRIDS = [(256, 24, 32), (32, 256, 24), (24, 32, 256), (128, 128, 24), (36, 40, 30)]


def _state(nx, ny, nz, h, n_par, seed):
    """Woods-Saxon droplet of 12 atoms at the box centre (phase noise 1%) and a
    Gaussian particle cloud (sigma 0.5 A) moving along z, in oracle layouts."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) - nx // 2) * h
    y = (np.arange(ny) - ny // 2) * h
    z = (np.arange(nz) - nz // 2) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    R = 2.22 * 12 ** (1.0 / 3.0)
    rho = 0.0218 / (1.0 + np.exp((r - R) / 0.7))
    rho *= 12.0 / (rho.sum() * h ** 3)
    psi = np.sqrt(rho) * np.exp(1j * 0.01 * rng.standard_normal(rho.shape))
    rimp = 0.5 * rng.standard_normal((3, n_par))
    vimp = np.zeros((3, n_par))
    vimp[2] = 0.5 * 7.63822291
    return psi, rimp, vimp


@pytest.mark.parametrize("grid", GRIDS)
def test_noncubic_trajectory(hip, grid):
    from oracle import hetddft_oracle as orc
    nx, ny, nz = grid
    h = 0.5
    n_par = 64
    p = hip.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2,
                           n_par=n_par, ironx=3)
    k = hip.make_kernels(p)
    psi, rimp, vimp = _state(nx, ny, nz, h, n_par, seed=11)
    o = orc.Oracle(orc.Params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2,
                              n_par=n_par, ironx=3), k)
    o.set_state(psi, rimp, vimp)
    print(f"noncubic trajectory {grid}", flush=True)
    dt = 5e-4 / 7.63822291
    with hip.Instance(p) as g:
        g.set_kernels(k)
        g.set_state(psi, rimp, vimp)
        o.start()
        g.start()
        for name in ("pot4", "uimp", "dencg", "delj4", "falfs", "intxalf", "intyalf", "intzalf", "hpsix"):
            assert rel_max(g.get_field(name), getattr(o, name)) < TOL, name
        for it in range(3 + 2):
            o.iterate(it, dt)
            g.iterate(dt)
            gpsi, gr, gv, ga = g.get_state()
            assert rel_max(gpsi, o.psi) < TOL, (it, rel_max(gpsi, o.psi))
            assert rel_max(gr, o.rimp) < TOL, it
            assert rel_max(ga, o.aimp) < TOL, it


def _conv_ref(f, k):
    n = f.size
    return np.fft.irfftn(np.fft.rfftn(f) * k, s=f.shape, axes=(0, 1, 2), norm="forward") * (1.0 / n)


@pytest.mark.parametrize("shape", [(256, 64, 64), (128, 128, 256), (64, 256, 128), (256, 256, 256)])
@pytest.mark.parametrize("prec", [0, 1])
def test_convolve_large(hip, shape, prec):
    """Batched convolutions at the 256-point lengths (shape is nz, ny, nx),
    up to the full 256^3 grid of the large-droplet workload."""
    nz, ny, nx = shape
    p = hip.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * 0.15, ymax=ny * 0.15, zmax=nz * 0.15,
                           n_par=4, precision=prec)
    print(f"convolve large shape={shape} prec={prec}", flush=True)
    rng = np.random.default_rng(5)
    fields = [rng.standard_normal(shape) for _ in range(2)]
    # a separable real kernel, even in every frequency: unambiguous for c2r
    kz = np.exp(-np.abs(np.fft.fftfreq(nz)) * 3.0)
    ky = np.exp(-np.abs(np.fft.fftfreq(ny)) * 2.0)
    kx = 1.0 / (1.0 + 10.0 * np.fft.rfftfreq(nx) ** 2)
    kern = kz[:, None, None] * ky[None, :, None] * kx[None, None, :]
    with hip.Instance(p) as g:
        outs = g.convolve(fields, kern)
    tol = 1e-12 if prec == 0 else 2e-5
    for f, o in zip(fields, outs):
        assert rel_max(o, _conv_ref(f, kern)) < tol
