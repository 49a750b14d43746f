"""API-level parity: halo transfers (the reference's mblock groups and 3-D orientations,
mixed precision), batched tridiagonal solves, reductions and the host random generator.
The oracle host backend is checked against closed forms and independent computations;
the GPU runtime is checked against the oracle bitwise."""
import math

import numpy as np
import pytest

from app_util import argv, lib, p

KINDS = ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)]


# ---------------------------------------------------------------- mblock halo groups

def mblock(kind, mask):
    L = lib("apitest", kind)
    a0 = np.zeros((24, 24))
    a1 = np.zeros((24, 24))
    argc, av = argv()
    assert L.api_mblock(argc, av, mask, p(a0), p(a1)) == 0
    return a0, a1


def mblock_expected(mask):
    """Halo values from the geometry of each group (apps/c/mblock/mblock.cpp:80-158),
    written out face by face. Arrays are [y + 2, x + 2] of the 24 x 24 allocation."""
    d = [np.zeros((24, 24)), np.zeros((24, 24))]
    for b in (0, 1):
        for y in range(20):
            for x in range(20):
                d[b][y + 2, x + 2] = x + 20 * y

    def at(b, x, y):
        return d[b][y + 2, x + 2]

    def put(b, x, y, v):
        d[b][y + 2, x + 2] = v

    for g in range(5):
        if not mask & (1 << g):
            continue
        if g == 0:    # block 1 sits right of block 0
            for i in range(2):
                for y in range(20):
                    put(1, -2 + i, y, at(0, 18 + i, y))
            for i in range(2):
                for y in range(20):
                    put(0, 20 + i, y, at(1, i, y))
        elif g == 1:  # block 1 sits above block 0
            for j in range(2):
                for x in range(20):
                    put(1, x, -2 + j, at(0, x, 18 + j))
            for j in range(2):
                for x in range(20):
                    put(0, x, 20 + j, at(1, x, j))
        elif g == 2:  # block 1 left of block 0, upside down
            for i in range(2):
                for y in range(20):
                    put(1, 20 + i, 19 - y, at(0, i, y))
            for i in range(2):
                for y in range(20):
                    put(0, -2 + i, y, at(1, 18 + i, 19 - y))
        elif g == 3:  # block 1 below block 0, mirrored in x
            for j in range(2):
                for x in range(20):
                    put(1, 19 - x, 20 + j, at(0, x, j))
            for j in range(2):
                for x in range(20):
                    put(0, x, -2 + j, at(1, 19 - x, 18 + j))
        elif g == 4:  # block 0's right face is block 1's bottom face, axes swapped
            for i in range(2):
                for y in range(20):
                    put(1, y, -2 + i, at(0, 18 + i, y))
            for i in range(2):
                for y in range(20):
                    put(0, 20 + i, y, at(1, y, i))
    return d


@pytest.mark.parametrize("mask", [1, 2, 4, 8, 16, 31])
def test_oracle_mblock_halo_groups(mask):
    a0, a1 = mblock("cpu", mask)
    e0, e1 = mblock_expected(mask)
    assert np.array_equal(a0, e0) and np.array_equal(a1, e1)


@pytest.mark.gpu
@pytest.mark.parametrize("mask", [1, 2, 4, 8, 16, 31])
def test_gpu_mblock_halo_groups(mask):
    a0, a1 = mblock("gpu", mask)
    e0, e1 = mblock_expected(mask)
    assert np.array_equal(a0, e0) and np.array_equal(a1, e1)


# ---------------------------------------------------------------- 3-D orientations

CASES3D = [
    # sizes and halos of A and B, types, iter, from_base, to_base, from_dir, to_dir
    dict(sa=(6, 5, 4), sb=(6, 5, 4), ta=0, tb=0, it=(2, 5, 4), fb=(4, 0, 0), tbse=(-2, 0, 0),
         fd=(1, 2, 3), td=(1, 2, 3)),
    dict(sa=(6, 5, 4), sb=(5, 4, 6), ta=0, tb=0, it=(5, 4, 2), fb=(0, 0, 2), tbse=(0, 0, -2),
         fd=(1, 2, 3), td=(2, 3, 1)),
    dict(sa=(7, 3, 5), sb=(5, 7, 3), ta=0, tb=1, it=(7, 3, 2), fb=(0, 0, 3), tbse=(0, 0, 0),
         fd=(1, -2, 3), td=(3, 1, -2)),
    dict(sa=(4, 4, 4), sb=(4, 4, 4), ta=1, tb=0, it=(4, 1, 4), fb=(0, 3, 0), tbse=(0, 4, 0),
         fd=(-1, 2, -3), td=(1, 2, 3)),
]


def halo3d(kind, c):
    L = lib("apitest", kind)
    dm, dp = (-2, -2, -2), (2, 2, 2)
    na = [s + 4 for s in c["sa"]]
    nb = [s + 4 for s in c["sb"]]
    i = np.arange(na[0] * na[1] * na[2])
    inA = (1.0 + (i % na[0]) + 100.0 * ((i // na[0]) % na[1]) + 10000.0 * (i // (na[0] * na[1])))
    inA = inA.astype(np.float64)
    out = np.zeros(nb[0] * nb[1] * nb[2])
    ia = lambda v: np.array(v, dtype=np.int32)
    argc, av = argv()
    rc = L.api_halo3d(argc, av, p(ia(c["sa"])), p(ia(dm)), p(ia(dp)), c["ta"], p(ia(c["sb"])),
                      p(ia(dm)), p(ia(dp)), c["tb"], p(ia(c["it"])), p(ia(c["fb"])),
                      p(ia(c["tbse"])), p(ia(c["fd"])), p(ia(c["td"])), p(inA), p(out))
    assert rc == 0
    return inA.reshape(na[::-1]), out.reshape(nb[::-1])


@pytest.mark.parametrize("case", range(len(CASES3D)))
def test_oracle_halo3d_moves_every_value_once(case):
    # each value of the source window lands exactly once; the rest of B stays zero
    c = CASES3D[case]
    A, B = halo3d("cpu", c)
    n = int(np.prod(c["it"]))
    assert np.count_nonzero(B) == n
    got = np.sort(B[B != 0])
    # the source window in A, in the from dataset's own coordinates
    ranges = []
    for d in range(3):
        ax = abs(c["fd"][d]) - 1
        lo = c["fb"][d]
        ranges.append(range(lo + 2, lo + 2 + c["it"][ax]))
    src = A[np.ix_(ranges[2], ranges[1], ranges[0])]
    assert np.array_equal(got, np.sort(src.ravel()))


def halo3d_expected(c, A):
    """B after the transfer, element by element, from the orientation rule of
    ops_halo_transfer (ops/c/src/core/ops_device_singlenode_common.cpp:286-336 with
    ops_halo_copy_tobuf/frombuf, doc/opsapi.md:192-212): iteration point q (one coordinate
    per iteration axis) is read in dat dimension i of the source at
    s_i + q[|dir_i| - 1] when dir_i > 0 and at s_i + iter[|dir_i| - 1] - 1 - q[|dir_i| - 1]
    when dir_i < 0, with s_i = base_i - d_m_i (allocation index; dat base 0), and written
    at the destination's point given by the same rule; a value crossing precisions goes
    through the narrower type. Arrays are [z, y, x] of the allocations (2-deep halos)."""
    nb = [s + 4 for s in c["sb"]]
    B = np.zeros(nb[::-1])
    it = c["it"]
    q = np.indices(it).reshape(3, -1)  # q[a] = coordinate along iteration axis a

    def side(base, dirs):
        idx = []
        for i in range(3):
            a = abs(dirs[i]) - 1
            s = base[i] + 2
            idx.append(s + q[a] if dirs[i] > 0 else s + it[a] - 1 - q[a])
        return idx

    fx, fy, fz = side(c["fb"], c["fd"])
    tx, ty, tz = side(c["tbse"], c["td"])
    v = A[fz, fy, fx]
    if c["ta"] != c["tb"]:
        v = v.astype(np.float32).astype(np.float64)
    B[tz, ty, tx] = v
    return B


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("case", range(len(CASES3D)))
def test_halo3d_positional(kind, case):
    # every element of B, halo and interior, at the position the orientation rule gives
    c = CASES3D[case]
    A, B = halo3d(kind, c)
    assert np.array_equal(B, halo3d_expected(c, A))


def test_halo3d_expected_detects_a_flipped_axis():
    # the restatement is sensitive to the orientation: flipping one axis of the destination
    # or of the source changes the expected B
    for case in (1, 2):
        c = dict(CASES3D[case])
        A, _ = halo3d("cpu", c)
        ref = halo3d_expected(c, A)
        flipped = dict(c, td=tuple(-d if k == 0 else d for k, d in enumerate(c["td"])))
        assert not np.array_equal(halo3d_expected(flipped, A), ref)
        if case == 2:
            swapped = dict(c, fd=(c["fd"][0], -c["fd"][1], c["fd"][2]))
            assert not np.array_equal(halo3d_expected(swapped, A), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES3D)))
def test_gpu_halo3d_equals_oracle(case):
    c = CASES3D[case]
    _, Bc = halo3d("cpu", c)
    _, Bg = halo3d("gpu", c)
    assert np.array_equal(Bc, Bg)


@pytest.mark.parametrize("kind", KINDS)
def test_halo_outside_allocation_is_rejected(kind):
    # the z window of B would be [3, 6) in a dat whose allocation ends at z = 4
    c = dict(CASES3D[2])
    c["tbse"] = (0, 0, 3)
    with pytest.raises(AssertionError):
        halo3d(kind, c)


def test_oracle_halo3d_identity_orientation_is_a_box_copy():
    c = CASES3D[0]
    A, B = halo3d("cpu", c)
    # A x in [4, 6) -> B x in [-2, 0), all y, z (+2 for the halo offset in the arrays)
    assert np.array_equal(B[2:6, 2:7, 0:2], A[2:6, 2:7, 6:8])


# ---------------------------------------------------------------- tridiagonal batches

def tri_inputs(size, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(size))
    a = rng.uniform(-1, 0, n)
    c = rng.uniform(-1, 0, n)
    b = 2.5 + rng.uniform(0, 1, n)  # diagonally dominant
    d = rng.uniform(-1, 1, n)
    u = rng.uniform(-1, 1, n)
    return a, b, c, d, u


def tri_run(kind, ndim, sd, size, rng_box, inc, seed=1):
    L = lib("apitest", kind)
    a, b, c, d, u = tri_inputs(size, seed)
    out = np.zeros(int(np.prod(size)))
    argc, av = argv()
    sz = np.array(list(size) + [1] * (3 - ndim), dtype=np.int32)
    rg = np.array(rng_box, dtype=np.int32)
    assert L.api_tridiag(argc, av, ndim, sd, p(sz), p(rg), p(a), p(b), p(c), p(d), p(u), inc, p(out)) == 0
    return (a, b, c, d, u), out


def tri_dense(inputs, ndim, sd, size, box, inc):
    a, b, c, d, u = [x.reshape(tuple(size[::-1])) for x in inputs]
    res = (u if inc else d).copy()
    sl = tuple(slice(box[2 * k], box[2 * k + 1]) for k in range(ndim))[::-1]
    ax = ndim - 1 - sd  # numpy axis of the solve dimension
    A, B, C, D = (np.moveaxis(x[sl], ax, -1) for x in (a, b, c, d))
    R = np.moveaxis(res[sl], ax, -1)
    m = A.shape[-1]
    for idx in np.ndindex(A.shape[:-1]):
        M = np.diag(B[idx]) + np.diag(A[idx][1:], -1) + np.diag(C[idx][:-1], 1)
        x = np.linalg.solve(M, D[idx])
        R[idx] = R[idx] + x if inc else x
    return res.ravel()


TRI = [
    (1, 0, (37,), (3, 30), 0),
    (2, 0, (17, 9), (0, 17, 2, 9), 0),
    (2, 1, (17, 9), (1, 16, 0, 9), 1),
    (3, 0, (9, 7, 5), (0, 9, 1, 6, 0, 5), 0),
    (3, 1, (9, 7, 5), (2, 9, 0, 7, 1, 5), 0),
    (3, 2, (9, 7, 5), (0, 9, 0, 7, 0, 5), 1),
    (3, 0, (70, 3, 2), (0, 70, 0, 3, 0, 2), 0),   # more than one 64-line tile row of x-lines
]


@pytest.mark.parametrize("case", TRI)
def test_oracle_tridiag_matches_dense_solve(case):
    ndim, sd, size, box, inc = case
    inputs, out = tri_run("cpu", ndim, sd, size, box, inc)
    ref = tri_dense(inputs, ndim, sd, size, box, inc)
    assert np.max(np.abs(out - ref)) < 1e-13


# line lengths around the partition boundaries of the partitioned solve (16 partitions of
# ceil(n/16) rows: empty, one-row and two-row partitions) and its 256-row limit
TRI += [(2, 0, (n, 3), (0, n, 0, 3), n % 2) for n in (1, 2, 3, 15, 16, 17, 31, 33, 47, 65, 250, 256)]
TRI += [(3, 2, (20, 3, 257), (0, 20, 0, 3, 0, 257), 1), (3, 1, (33, 100, 2), (0, 33, 0, 100, 0, 2), 0)]


def test_partitioned_restatement_matches_dense_solve():
    # the numpy restatement of the partitioned algorithm against dense solves, line lengths
    # 1..256 (every partition shape), diagonally dominant and ADI-like coefficients
    from trid_ref import part_solve
    rng = np.random.default_rng(8)
    for n in list(range(1, 40)) + [63, 64, 65, 127, 128, 129, 200, 255, 256]:
        a = rng.uniform(-1, 0, (5, n))
        c = rng.uniform(-1, 0, (5, n))
        b = 2.5 + rng.uniform(0, 1, (5, n))
        d = rng.uniform(-1, 1, (5, n))
        for red in ("pcr", "thomas"):
            x = part_solve(a, b, c, d, reduced=red)
            for i in range(5):
                M = np.diag(b[i]) + np.diag(a[i][1:], -1) + np.diag(c[i][:-1], 1)
                ref = np.linalg.solve(M, d[i])
                assert np.max(np.abs(x[i] - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref))), (n, red)
    # a dropped coupling term (the partition's last row ignoring the next partition) fails
    n = 64
    a, c = -0.5 * np.ones((1, n)), -0.5 * np.ones((1, n))
    b, d = 2.0 * np.ones((1, n)), rng.uniform(-1, 1, (1, n))
    x = part_solve(a, b, c, d)
    M = np.diag(b[0]) + np.diag(a[0][1:], -1) + np.diag(c[0][:-1], 1)
    cut = M.copy()
    cut[3, 4] = 0.0  # rows 0-3 and 4-7 form the first two partitions
    assert np.max(np.abs(x[0] - np.linalg.solve(M, d[0]))) < 1e-14
    assert np.max(np.abs(x[0] - np.linalg.solve(cut, d[0]))) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("case", TRI)
def test_gpu_tridiag_sequential_sweeps_equal_oracle(case, monkeypatch):
    # OPS_TRID_ALGO=thomas: the sequential sweeps, bitwise the oracle's sequential solve
    monkeypatch.setenv("OPS_TRID_ALGO", "thomas")
    ndim, sd, size, box, inc = case
    _, oc = tri_run("cpu", ndim, sd, size, box, inc)
    _, og = tri_run("gpu", ndim, sd, size, box, inc)
    assert np.array_equal(oc, og)


@pytest.mark.gpu
@pytest.mark.parametrize("case", TRI)
def test_gpu_tridiag_partitioned(case):
    # the default solve: bitwise the restatement of its algorithm where it applies (lines of
    # at most 256 rows; longer lines take the sequential sweeps, bitwise the oracle), and
    # within 1e-12 of the oracle's sequential solve
    from trid_ref import part_dims
    ndim, sd, size, box, inc = case
    inputs, oc = tri_run("cpu", ndim, sd, size, box, inc)
    _, og = tri_run("gpu", ndim, sd, size, box, inc)
    n = box[2 * sd + 1] - box[2 * sd]
    exp = part_dims(inputs, ndim, sd, size, box, inc) if n <= 256 else oc
    assert np.array_equal(og, exp)
    assert np.max(np.abs(og - oc)) <= 1e-12 * np.max(np.abs(oc))


# ---------------------------------------------------------------- reductions

def reduce_run(kind, nx, ny, seed=3):
    L = lib("apitest", kind)
    rng = np.random.default_rng(seed)
    v = rng.uniform(-5, 5, nx * ny)
    w = rng.integers(-1000, 1000, nx * ny).astype(np.int32)
    out = np.zeros(8)
    argc, av = argv()
    assert L.api_reduce(argc, av, nx, ny, p(v), p(w), p(out)) == 0
    return v, w, out


def reduce_graph_run(kind, nx, ny, passes, seed=4):
    L = lib("apitest", kind)
    v = np.random.default_rng(seed).integers(-50, 50, nx * ny).astype(np.float64)
    out = np.zeros(passes)
    argc, av = argv()
    assert L.api_reduce_graph(argc, av, nx, ny, p(v), passes, p(out)) == 0
    return v, out


@pytest.mark.parametrize("kind", KINDS)
def test_reduction_read_after_every_pass(kind):
    # passes of {v += 1; sum(v)} read on the host after each: eagerly, replayed from a
    # captured graph (on the GPU), and eagerly again; integer values, so every sum is exact
    nx, ny, passes = 300, 201, 7
    v, out = reduce_graph_run(kind, nx, ny, passes)
    expect = [v.sum() + (k + 1) * nx * ny for k in range(passes)]
    assert out.tolist() == expect


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("shape", [(2, 3), (7, 5), (300, 201)])
def test_reductions(kind, shape):
    nx, ny = shape
    v, w, out = reduce_run(kind, nx, ny)
    assert abs(out[0] - math.fsum(v)) <= 1e-12 * max(1.0, math.fsum(abs(v)))
    assert out[1] == v.min() and out[2] == v.max()
    assert out[3] == int(w.astype(np.int64).sum()) and out[4] == w.min() and out[5] == w.max()
    assert abs(out[6] - math.fsum(v * v)) <= 1e-12 * math.fsum(v * v)
    assert out[7] == nx * ny


# ---------------------------------------------------------------- one dat, two arguments

@pytest.mark.parametrize("kind", KINDS)
def test_dat_read_and_written_through_two_arguments(kind):
    """v = 2 v + 1 with v passed as both the read and the written argument (the device
    build's kernel variant without no-alias pointers), then u = 2 v + 1 into another dat
    (the no-alias variant)."""
    L = lib("apitest", kind)
    nx, ny = 67, 19
    v = np.random.default_rng(5).integers(-1000, 1000, nx * ny).astype(np.float64)
    out = np.zeros(2 * nx * ny)
    argc, av = argv()
    assert L.api_alias(argc, av, nx, ny, p(v), p(out)) == 0
    assert np.array_equal(out[:nx * ny], 2 * v + 1)
    assert np.array_equal(out[nx * ny:], 2 * (2 * v + 1) + 1)


# ---------------------------------------------------------------- random numbers

def mt_uniform(seed, n):
    """std::mt19937(seed) + std::uniform_real_distribution<double>(0, 1) as libstdc++'s
    generate_canonical forms it: two 32-bit draws, (g1 + g2 * 2^32) / 2^64, rounded in
    double at the addition."""
    bg = np.random.RandomState(seed)._bit_generator
    raw = bg.random_raw(2 * n).astype(np.float64)
    s = raw[0::2] + raw[1::2] * 4294967296.0
    r = s / 18446744073709551616.0
    return np.where(r >= 1.0, np.nextafter(1.0, 0.0), r)


@pytest.mark.parametrize("kind", KINDS)
def test_random_uniform_is_mt19937_canonical(kind):
    L = lib("apitest", kind)
    nx, ny = 13, 7
    out = np.zeros((nx + 2) * (ny + 2))
    argc, av = argv()
    assert L.api_random(argc, av, 1234, nx, ny, p(out)) == 0
    # the whole allocation, halos included, in memory order (ops_lib_core.cpp:2404-2416)
    assert np.array_equal(out, mt_uniform(1234, out.size))


def mt_normal(seed, n):
    """std::normal_distribution<double>(0, 1) over std::mt19937(seed), as libstdc++ forms it
    (bits/random.tcc, the Marsaglia polar method): draw x = 2u - 1 and y = 2v - 1 from the
    canonical uniforms until 0 < r2 = x*x + y*y <= 1, m = sqrt(-2 log(r2) / r2), return y*m
    and keep x*m for the next call. One distribution object per fill, so a fill of odd
    length drops its last saved value. math.log / math.sqrt are the C library's."""
    u = iter(mt_uniform(seed, 4 * n + 64))
    out, saved = [], None
    while len(out) < n:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            x = 2.0 * next(u) - 1.0
            y = 2.0 * next(u) - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        m = math.sqrt(-2.0 * math.log(r2) / r2)
        saved = x * m
        out.append(y * m)
    return np.array(out)


def test_mt_normal_statistics():
    # the restatement itself: standard normal moments over 20000 draws, and the pair
    # structure (the saved value follows its partner)
    z = mt_normal(99, 20000)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    assert abs(np.mean(z ** 4) - 3.0) < 0.15


@pytest.mark.parametrize("kind", KINDS)
def test_random_normal_is_libstdcxx_polar(kind):
    L = lib("apitest", kind)
    nx, ny = 13, 7
    out = np.zeros((nx + 2) * (ny + 2))
    argc, av = argv()
    assert L.api_random_normal(argc, av, 4321, nx, ny, p(out)) == 0
    # the whole allocation, halos included, in memory order (ops_lib_core.cpp:2452-2474)
    assert np.array_equal(out, mt_normal(4321, out.size))


# ---------------------------------------------------------------- loop batches

def batch_sizes(n):
    return [(3 + (37 * j) % 70, 2 + (11 * j) % 9) for j in range(n)]


def batch_run(kind, n, mode):
    L = lib("apitest", kind)
    out = np.zeros(sum(a * b for a, b in batch_sizes(n)))
    argc, av = argv()
    assert L.api_batch(argc, av, n, mode, p(out)) == 0
    res, off = [], 0
    for nx, ny in batch_sizes(n):
        res.append(out[off:off + nx * ny].reshape(ny, nx))
        off += nx * ny
    return res


def batch_expected(n, solve):
    """O_j = c_j S_j + x + 1000 y with S_j = j + x/7 - y/3, c_j = 0.5 + j; with `solve`, O_1's
    rows then go through the dense solve of tridiag(-1, 4, -1) along x."""
    res = []
    for j, (nx, ny) in enumerate(batch_sizes(n)):
        y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
        o = (0.5 + j) * (j + x / 7.0 - y / 3.0) + x + 1000.0 * y
        if solve and j == 1:
            M = 4.0 * np.eye(nx) - np.eye(nx, k=1) - np.eye(nx, k=-1)
            o = np.linalg.solve(M, o.T).T
        res.append(o)
    return res


@pytest.mark.parametrize("mode", [0, 2])
def test_oracle_batch_sequence_closed_form(mode):
    n = 11
    got = batch_run("cpu", n, mode)
    for g, e in zip(got, batch_expected(n, mode & 2)):
        assert np.max(np.abs(g - e)) <= 1e-12 * np.max(np.abs(e))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", range(8))
def test_gpu_batch_sequence_equals_oracle(mode, monkeypatch):
    # 11 (more than one batch of 8) and 17 launches of one kernel of mixed extents, its
    # WRITE and RW stubs interleaved, batched or not, with a tridiagonal solve of a batched
    # launch's output in between (sequential sweeps, as the oracle solves), eagerly or
    # replayed from a graph: bitwise the oracle's
    monkeypatch.setenv("OPS_TRID_ALGO", "thomas")
    for n in (11, 17):
        cpu = batch_run("cpu", n, mode & 2)
        gpu = batch_run("gpu", n, mode)
        for a, b in zip(cpu, gpu):
            assert np.array_equal(a, b)


# ---------------------------------------------------------------- loop chains

def chain_run(kind, nx, ny, tiles, mode):
    L = lib("apitest", kind)
    out = np.zeros(4 * nx * ny + 1)
    argc, av = argv()
    assert L.api_chain(argc, av, nx, ny, tiles, mode, p(out)) == 0
    return out


def chain_numpy(nx, ny, mode):
    """The api_chain sequence in program order (apps/apitest: api_ch_init, l1..l5), on
    [y][x] arrays with two halo layers."""
    h = 2
    y, x = np.mgrid[-h:ny + h, -h:nx + h]
    # C's % truncates toward zero (negative in the halo): np.fmod, not Python's %
    A = 1.0 + 0.01 * x - 0.003 * y * y + 0.25 * np.fmod(x + 3 * y, 7)
    B = 0.5 - 0.02 * x * y + 0.125 * np.fmod(2 * x + y, 5)
    C = np.zeros_like(A)
    D = np.zeros_like(A)

    def at(F, dx, dy, y0, y1):  # F(x + dx, y + dy) over rows y0..y1-1 and every x
        return F[h + y0 + dy:h + y1 + dy, h + dx:h + nx + dx]

    def put(F, v, y0, y1):
        F[h + y0:h + y1, h:h + nx] = v

    w = 2.0
    put(C, at(A, 0, -1, 0, ny) + w * at(A, 0, 1, 0, ny) + at(B, 0, 0, 0, ny) + 0.25 * at(A, 1, 0, 0, ny), 0, ny)
    put(A, 0.5 * at(C, 0, 0, 1, ny) + at(C, 0, 2, 1, ny) - 0.125 * at(C, -1, 0, 1, ny), 1, ny)
    put(D, at(A, 0, -2, 0, ny - 1) - at(A, 0, 1, 0, ny - 1) + at(C, 0, -1, 0, ny - 1), 0, ny - 1)
    total = float(np.sum(at(D, 0, 0, 0, ny))) if mode & 2 else 0.0
    put(B, at(B, 0, 0, 2, ny) + at(D, 0, 1, 2, ny) - at(D, 0, -1, 2, ny), 2, ny)
    put(C, at(D, 0, 0, 0, ny) + at(B, 0, 2, 0, ny) - at(B, 0, -2, 0, ny), 0, ny)
    inner = [F[h:h + ny, h:h + nx].ravel() for F in (A, B, C, D)]
    return np.concatenate(inner + [np.array([total])])


@pytest.mark.parametrize("mode", [0, 7])
def test_oracle_chain_sequence_equals_numpy(mode):
    # the host backend runs chained loops whole and in order: the numpy restatement of the
    # sequence, operation for operation (bitwise), the sum to rounding
    nx, ny = 23, 19
    got = chain_run("cpu", nx, ny, 4, mode)
    want = chain_numpy(nx, ny, mode)
    m = 4 * nx * ny
    assert np.array_equal(got[:m].view(np.uint64), want[:m].view(np.uint64))
    assert abs(got[m] - want[m]) <= 1e-12 * max(1.0, abs(want[m]))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(37, 29), (300, 257)])
@pytest.mark.parametrize("tiles", [2, 3, 7, 16, 40])
@pytest.mark.parametrize("mode", [1, 3, 7])
def test_gpu_chain_bitwise_equals_loops_in_order(shape, tiles, mode):
    # chained loops run band by band with skewed boundaries: every field bitwise the loops
    # run in order (the host backend, and the GPU without the chain); the reduction between
    # l3 and l4 is not chained and sums in the GPU's order either way
    nx, ny = shape
    want = chain_run("cpu", nx, ny, tiles, mode)
    eager = chain_run("gpu", nx, ny, tiles, mode & ~1)
    got = chain_run("gpu", nx, ny, tiles, mode)
    m = 4 * nx * ny
    assert np.array_equal(eager[:m].view(np.uint64), want[:m].view(np.uint64))
    assert np.array_equal(got[:m].view(np.uint64), want[:m].view(np.uint64))
    assert got[m] == eager[m]
    assert abs(got[m] - want[m]) <= 1e-12 * max(1.0, abs(want[m]))
