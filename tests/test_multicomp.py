"""Multi-component dats, AoS and SoA builds: a 2- and a 3-component dat set from their
indices, a stencil reading neighbours of different components, a sum and a max over
components, a periodic halo transfer of a 2-component dat, and the stencil again from an
API v1 kernel (pointer arguments, OPS_ACC_MD<n> accessors); against numpy, and the device
against the oracle bit for bit."""
import numpy as np
import pytest

from app_util import argv, lib, p

KINDS = ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)]
APPS = ["multicomp", "multicomp_soa"]


def run(app, kind, nx, ny):
    L = lib(app, kind)
    n = nx * ny
    m = (nx - 2) * (ny - 2)
    out = np.full(n * 2 + n * 3 + m + 2 + (nx + 2) * ny * 2 + m, np.nan)
    argc, av = argv()
    assert L.multicomp_run(argc, av, nx, ny, p(out)) == 0
    return out


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("app", APPS)
@pytest.mark.parametrize("shape", [(7, 5), (130, 67)])
def test_multicomp(app, kind, shape):
    nx, ny = shape
    out = run(app, kind, nx, ny)
    y, x = np.mgrid[0:ny, 0:nx]
    v = np.stack([x, 10 * y], axis=-1).astype(float)                  # [y][x][c]
    w = np.stack([100 * c + x + 7 * y for c in range(3)], axis=-1).astype(float)
    k = 0
    assert np.array_equal(out[k:k + nx * ny * 2], v.ravel()); k += nx * ny * 2
    assert np.array_equal(out[k:k + nx * ny * 3], w.ravel()); k += nx * ny * 3
    st = (v[1:-1, 2:, 0] + v[:-2, 1:-1, 1] + w[1:-1, :-2, 2] - w[2:, 1:-1, 0])
    assert np.array_equal(out[k:k + (nx - 2) * (ny - 2)], st.ravel()); k += (nx - 2) * (ny - 2)
    assert out[k] == v.sum() and out[k + 1] == w[..., 2].max(); k += 2
    wide = np.concatenate([v[:, -1:], v, v[:, :1]], axis=1)            # periodic along x
    assert np.array_equal(out[k:k + wide.size], wide.ravel()); k += wide.size
    assert np.array_equal(out[k:], st.ravel())                         # API v1 kernel


@pytest.mark.gpu
@pytest.mark.parametrize("app", APPS)
def test_gpu_equals_oracle(app):
    a, b = run(app, "gpu", 61, 33), run(app, "cpu", 61, 33)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
