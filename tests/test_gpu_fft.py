"""Batched 3D r2c/c2r convolution on the GPU vs numpy.fft (the FFT layer of fft.f90)."""
import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu

SHAPES = [(24, 24, 24), (32, 32, 32), (48, 48, 48), (64, 64, 64), (40, 24, 30), (16, 18, 20),
          (80, 80, 80), (96, 96, 96), (128, 128, 128)]


def _hermitian_kernel(rng, shape):
    """Random real kernel with K(kx,ky,kz) = K(kx,-ky,-kz), as every kernel of
    the path (a function of |p|) has; c2r of the product is then well defined."""
    nz, ny, nx = shape
    a = rng.standard_normal(nx // 2 + 1)
    b = rng.standard_normal(ny // 2 + 1)[np.minimum(np.arange(ny), ny - np.arange(ny))]
    c = rng.standard_normal(nz // 2 + 1)[np.minimum(np.arange(nz), nz - np.arange(nz))]
    return c[:, None, None] * b[None, :, None] * a[None, None, :]


def _conv_ref(f, k):
    n = f.size
    return np.fft.irfftn(np.fft.rfftn(f) * k, s=f.shape, axes=(0, 1, 2), norm="forward") * (1.0 / n)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [0, 1])
def test_convolve_random(hip, shape, prec):
    nz, ny, nx = shape
    p = hip.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * 0.25, ymax=ny * 0.25, zmax=nz * 0.25,
                           n_par=4, precision=prec)
    print(f"convolve shape={shape} prec={prec}", flush=True)
    rng = np.random.default_rng(3)
    fields = [rng.standard_normal(shape) for _ in range(3)]
    kern = _hermitian_kernel(rng, shape)
    with hip.Instance(p) as g:
        outs = g.convolve(fields, kern)
    tol = 1e-12 if prec == 0 else 2e-5
    for f, o in zip(fields, outs):
        assert rel_max(o, _conv_ref(f, kern)) < tol


@pytest.mark.parametrize("shape", [(32, 32, 32), (24, 40, 30)])
def test_convolve_identity_roundtrip(hip, shape):
    nz, ny, nx = shape
    p = hip.default_params(nx=nx, ny=ny, nz=nz, xmax=8.0, ymax=8.0, zmax=8.0, n_par=4)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(shape)
    with hip.Instance(p) as g:
        (o,) = g.convolve([f], np.ones((nz, ny, nx // 2 + 1)))
    assert np.abs(o - f).max() < 1e-13 * np.abs(f).max()
