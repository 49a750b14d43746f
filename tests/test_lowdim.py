"""Low-dimensional reductions: a 3-D field reduced with MAX, MIN and INC onto planes and
lines (strided stencils with stride 0 in the collapsed dimensions), against numpy
reductions of the same field; a second INC adds onto the first (same access as the last
write, so no re-initialisation); the reduced plane and line read back over 3-D."""
import numpy as np
import pytest

from app_util import argv, lib, p

KINDS = ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)]


def field(nx, ny, nz):
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    return ((x * 37 + y * 101 + z * 211) % 1009 - 500).astype(float)  # [z][y][x]


def run(kind, nx, ny, nz):
    L = lib("lowdim", kind)
    sizes = [nx * ny, ny * nz, nx * nz, nx, ny, nz]
    out = np.full(4 * sum(sizes) + nx * ny * nz, np.nan)
    argc, av = argv()
    assert L.lowdim_run(argc, av, nx, ny, nz, p(out)) == 0
    parts, k = [], 0
    for _ in range(4):
        mode = []
        for s in sizes:
            mode.append(out[k:k + s])
            k += s
        parts.append(mode)
    return parts, out[k:]


def expected(f, op):
    # planes XY (reduce z), YZ (reduce x), XZ (reduce y); lines X, Y, Z; x fastest
    return [op(f, axis=0).ravel(), op(f, axis=2).ravel(), op(f, axis=1).ravel(),
            op(f, axis=(0, 1)), op(f, axis=(0, 2)), op(f, axis=(1, 2))]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("shape", [(10, 10, 10), (37, 5, 19), (64, 48, 33)])
def test_lowdim_reductions(kind, shape):
    nx, ny, nz = shape
    f = field(nx, ny, nz)
    (mx, mn, s1, s2), bc = run(kind, nx, ny, nz)
    for got, want in zip(mx, expected(f, np.max)):
        assert np.array_equal(got, want)
    for got, want in zip(mn, expected(f, np.min)):
        assert np.array_equal(got, want)
    for got, want in zip(s1, expected(f, np.sum)):
        assert np.array_equal(got, want)  # integer values: exact in any order
    for got, want in zip(s2, expected(f, np.sum)):
        assert np.array_equal(got, 2 * want)
    xy = 2 * f.sum(axis=0)                 # [y][x]
    zl = 2 * f.sum(axis=(1, 2))            # [z]
    assert np.array_equal(bc.reshape(nz, ny, nx), xy[None, :, :] + zl[:, None, None])


@pytest.mark.gpu
def test_gpu_equals_oracle():
    a, ba = run("gpu", 45, 23, 17)
    b, bb = run("cpu", 45, 23, 17)
    for ma, mb in zip(a, b):
        for x, y in zip(ma, mb):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    assert np.array_equal(ba, bb)
