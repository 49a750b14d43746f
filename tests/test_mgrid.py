"""Multigrid transfers (restrict and prolong stencils, dats of grid strides 1, 2, 4): the
application's own checks, the transferred fields against closed forms, and the device
against the oracle bit for bit."""
import numpy as np
import pytest

from app_util import argv, lib, p

KINDS = ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)]


def run(kind, fac):
    L = lib("mgrid", kind)
    n = 24 * fac
    errs = np.full(2, -7, dtype=np.int32)
    out = [np.zeros(n * n), np.zeros(n * n // 4), np.zeros(n * n // 4), np.zeros(n * n // 16)]
    argc, av = argv()
    assert L.mgrid_run(argc, av, fac, p(errs), *(p(a) for a in out)) == 0
    return errs, out


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("fac", [1, 3, 16])
def test_transfers_match_closed_forms(kind, fac):
    errs, (pf, pm, rm, rc) = run(kind, fac)
    n, m, c = 24 * fac, 12 * fac, 6 * fac
    assert errs.tolist() == [0, 0]
    y, x = np.mgrid[0:n, 0:n]
    # coarse populated with x + c y; prolonged twice by 2 (index / 2 each time)
    assert np.array_equal(pf.reshape(n, n), (x // 4 + c * (y // 4)).astype(float))
    ym, xm = np.mgrid[0:m, 0:m]
    assert np.array_equal(pm.reshape(m, m), (xm // 2 + c * (ym // 2)).astype(float))
    # fine populated with x + n y; restricted twice by 2 (index * 2 each time)
    assert np.array_equal(rm.reshape(m, m), (2 * xm + n * 2 * ym).astype(float))
    yc, xc = np.mgrid[0:c, 0:c]
    assert np.array_equal(rc.reshape(c, c), (4 * xc + n * 4 * yc).astype(float))


@pytest.mark.gpu
def test_gpu_equals_oracle():
    eg, og = run("gpu", 5)
    ec, oc = run("cpu", 5)
    assert eg.tolist() == ec.tolist() == [0, 0]
    for a, b in zip(og, oc):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
