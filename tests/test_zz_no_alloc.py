"""The allocation-free tridiagonal entry point (ops_tridMultiDimBatch_noalloc): two calls in
a row, the second on a smaller box, against the queueing entry points with the sequential
sweeps (bitwise) and with the default partitioned solve (1e-12); the static scratch limit."""
import numpy as np
import pytest

from app_util import argv, lib, p

CASES = [
    # ndim, solvedim, size, first box, second (smaller) box, inc
    (3, 0, (70, 9, 5), (0, 70, 0, 9, 0, 5), (3, 40, 1, 6, 0, 4), 0),
    (3, 1, (33, 40, 7), (0, 33, 0, 40, 0, 7), (2, 20, 5, 30, 1, 6), 1),
    (3, 2, (17, 9, 64), (0, 17, 0, 9, 0, 64), (0, 16, 0, 8, 3, 50), 1),
    (2, 0, (300, 20), (0, 300, 0, 20), (10, 100, 0, 11), 0),
]


def run2(kind, case, mode, seed=4):
    ndim, sd, size, r1, r2, inc = case
    L = lib("apitest", kind)
    n = int(np.prod(size))
    rng = np.random.default_rng(seed)
    a, c = rng.uniform(-1, 0, n), rng.uniform(-1, 0, n)
    b = 2.5 + rng.uniform(0, 1, n)
    d, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    o1, o2 = np.zeros(n), np.zeros(n)
    argc, av = argv()
    ia = lambda v: np.array(v, dtype=np.int32)
    rc = L.api_tridiag2(argc, av, ndim, sd, p(ia(list(size) + [1] * (3 - ndim))), p(ia(r1)), p(ia(r2)),
                        p(a), p(b), p(c), p(d), p(u), inc, mode, p(o1), p(o2))
    return rc, o1, o2


@pytest.mark.parametrize("case", CASES)
def test_oracle_noalloc_equals_queueing_form(case):
    rc0, a1, a2 = run2("cpu", case, 0)
    rc1, b1, b2 = run2("cpu", case, 1)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_noalloc_equals_existing_entry_points(case, monkeypatch):
    rc, n1, n2 = run2("gpu", case, 1)
    assert rc == 0
    _, d1, d2 = run2("gpu", case, 0)  # default partitioned solve
    monkeypatch.setenv("OPS_TRID_ALGO", "thomas")
    _, t1, t2 = run2("gpu", case, 0)  # the same sequential sweeps through the queue
    assert np.array_equal(n1, t1) and np.array_equal(n2, t2)
    for x, y in ((n1, d1), (n2, d2)):
        assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(y))
    _, o1, o2 = run2("cpu", case, 1)  # and bitwise the oracle's sequential solve
    assert np.array_equal(n1, o1) and np.array_equal(n2, o2)


@pytest.mark.gpu
def test_gpu_noalloc_rejects_a_solve_beyond_its_scratch():
    # 200 x 200 x 120 points need 2 * 8 B each of c', d': 76.8 MB > 64 MiB
    case = (3, 1, (200, 200, 120), (0, 200, 0, 200, 0, 120), (0, 10, 0, 10, 0, 10), 0)
    rc, _, _ = run2("gpu", case, 1)
    assert rc != 0
    # a later call within the limit still works
    ok = (3, 0, (70, 9, 5), (0, 70, 0, 9, 0, 5), (3, 40, 1, 6, 0, 4), 0)
    rc, n1, _ = run2("gpu", ok, 1)
    assert rc == 0 and np.isfinite(n1).all()
