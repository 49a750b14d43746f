"""Batched 3-D ADI (preproc 7-point loop + x, y, z tridiagonal solves, the z solve adding
into u). The oracle is checked against per-line dense solves of the same scheme written
out in numpy, a constant fixed point, linearity and axis-permutation symmetry; the GPU
runtime against the oracle: bitwise with the sequential sweeps (OPS_TRID_ALGO=thomas), and
within 1e-12 with the default partitioned solve, including the 48-block configuration."""
import os

import numpy as np
import pytest

from app_util import argv, lib, p
from ops_amd.apps.adi.config import adi_sizes


def run(kind, sizes, steps, mode=1, seed=5, lam=1.0, u0=None, fetch=None, threads=8):
    L = lib("adi", kind)
    if kind == "cpu":
        L.ops_cpu_set_threads(threads)
    sz = np.array([v for s in sizes for v in s], dtype=np.int32)
    argc, av = argv()
    assert L.adi_init(argc, av, len(sizes), p(sz), mode, seed, lam) == 0
    if u0 is not None:
        for k, a in enumerate(u0):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert L.adi_set(k, p(a)) == 0
    assert L.adi_step(steps) == 0
    out = []
    for k in (range(len(sizes)) if fetch is None else fetch):
        n = L.adi_fetch(k, None)
        a = np.zeros(n)
        assert L.adi_fetch(k, p(a)) == n
        nx, ny, nz = sizes[k]
        out.append(a.reshape(nz, ny, nx))
    assert L.adi_finalize() == 0
    return out


def dense_step(u, lam=1.0):
    """One step of the scheme from first principles: du = lam * (7-point Laplacian) inside,
    0 on the boundary layer; then (I - lam/2 D2) solves along x, y and z, each row of a
    boundary-layer cell being the identity; u += the z result."""
    nz, ny, nx = u.shape
    bnd = np.zeros(u.shape, bool)
    bnd[0, :, :] = bnd[-1, :, :] = bnd[:, 0, :] = bnd[:, -1, :] = bnd[:, :, 0] = bnd[:, :, -1] = True
    du = np.zeros_like(u)
    c = u[1:-1, 1:-1, 1:-1]
    lap = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1] +
           u[:-2, 1:-1, 1:-1] + u[2:, 1:-1, 1:-1] - 6.0 * c)
    du[1:-1, 1:-1, 1:-1] = lam * lap

    def solve(d, axis):
        d = np.moveaxis(d, axis, -1).copy()
        b = np.moveaxis(bnd, axis, -1)
        m = d.shape[-1]
        for idx in np.ndindex(d.shape[:-1]):
            M = np.eye(m)
            for i in range(m):
                if not b[idx][i]:
                    M[i, i] = 1.0 + lam
                    M[i, i - 1] = M[i, i + 1] = -0.5 * lam
            d[idx] = np.linalg.solve(M, d[idx])
        return np.moveaxis(d, -1, axis)

    d1 = solve(du, 2)
    d2 = solve(d1, 1)
    d3 = solve(d2, 0)
    return u + d3


def test_oracle_one_step_matches_dense_line_solves():
    sizes = [(6, 5, 7)]
    u0 = run("cpu", sizes, 0)[0]
    u1 = run("cpu", sizes, 1)[0]
    assert np.max(np.abs(u1 - dense_step(u0))) < 1e-13


def test_oracle_reference_init_matches_dense_two_steps():
    sizes = [(5, 6, 4)]
    u0 = run("cpu", sizes, 0, mode=0)[0]
    assert u0[0, 0, 0] == 1.0 and u0[2, 2, 2] == 0.0
    u2 = run("cpu", sizes, 2, mode=0)[0]
    assert np.max(np.abs(u2 - dense_step(dense_step(u0)))) < 1e-13


def test_oracle_constant_field_is_a_fixed_point():
    sizes = [(9, 8, 7), (5, 11, 6)]
    u0 = [np.full((s[2], s[1], s[0]), 0.5) for s in sizes]
    out = run("cpu", sizes, 3, u0=u0)
    for a, b in zip(out, u0):
        assert np.array_equal(a, b)


def test_oracle_scheme_is_linear():
    sizes = [(10, 7, 9)]
    rng = np.random.default_rng(11)
    a = rng.uniform(0, 1, (9, 7, 10))
    b = rng.uniform(0, 1, (9, 7, 10))
    ra = run("cpu", sizes, 2, u0=[a])[0]
    rb = run("cpu", sizes, 2, u0=[b])[0]
    rab = run("cpu", sizes, 2, u0=[a + b])[0]
    assert np.max(np.abs(rab - (ra + rb))) < 1e-13


def test_oracle_axis_permutation_symmetry():
    # the x, y and z line operators commute on a cube, so a symmetric start stays symmetric
    out = run("cpu", [(8, 8, 8)], 5, mode=0)[0]
    for perm in [(0, 2, 1), (1, 0, 2), (2, 1, 0), (1, 2, 0)]:
        assert np.max(np.abs(out - out.transpose(perm))) < 1e-14
    assert not np.allclose(out, 1.0)


def test_adi_sizes_generator():
    s = adi_sizes()
    assert len(s) == 48 and all(64 <= v <= 256 for t in s for v in t)
    assert s == adi_sizes()  # deterministic
    assert len({v for t in s for v in t}) > 40


FIELD_BAR = 1e-12  # FP64 fields, relative to the field's largest value


def close(a, b):
    return np.max(np.abs(a - b)) <= FIELD_BAR * np.max(np.abs(b))


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["thomas", "partitioned"])
def test_gpu_uneven_blocks_equal_oracle(monkeypatch, algo):
    if algo == "thomas":
        monkeypatch.setenv("OPS_TRID_ALGO", "thomas")
    sizes = [(33, 17, 9), (16, 40, 7), (70, 3, 20)]
    cpu = run("cpu", sizes, 4)
    gpu = run("gpu", sizes, 4)
    for a, b in zip(cpu, gpu):
        assert np.array_equal(a, b) if algo == "thomas" else close(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["thomas", "partitioned"])
def test_gpu_autotuned_uneven_blocks_equal_oracle(monkeypatch, algo):
    # the autotuner tries its workgroup shapes on the first launches (blocks of different
    # sizes share one kernel); the fields stay the oracle's
    if algo == "thomas":
        monkeypatch.setenv("OPS_TRID_ALGO", "thomas")
    monkeypatch.setenv("OPS_AUTOTUNE", "1")
    sizes = [(33, 17, 9), (16, 40, 7), (70, 3, 20), (64, 24, 40)]
    gpu = run("gpu", sizes, 12)
    monkeypatch.delenv("OPS_AUTOTUNE")
    cpu = run("cpu", sizes, 12)
    for a, b in zip(cpu, gpu):
        assert np.array_equal(a, b) if algo == "thomas" else close(a, b)


GAMMA = 0x9E3779B97F4A7C15  # splitmix64 increment: the state after k draws is seed + k * GAMMA


def skip_seed(seed, sizes, k):
    """splitmix64 seed that starts block k's draws: the field is drawn block after block, x
    fastest, so block k alone is generated by a generator advanced by the cells of blocks
    0..k-1 (splitmix64's state is a counter, so skipping ahead is one multiply)."""
    skip = sum(a * b * c for a, b, c in sizes[:k])
    return (seed + skip * GAMMA) % (1 << 64)


def test_skip_seed_reproduces_the_continued_stream():
    # two small blocks drawn together equal the second block drawn alone from its skip seed
    sizes = [(5, 4, 3), (6, 2, 7)]
    both = run("cpu", sizes, 0)
    alone = run("cpu", sizes[1:], 0, seed=skip_seed(5, sizes, 1))
    assert np.array_equal(both[1], alone[0])
    assert not np.array_equal(run("cpu", sizes[1:], 0)[0], both[1])


# the 48-block benchmark configuration: batch boundaries of the 16-system launches (15/16,
# 31/32), the first and the last block
ADI_BLOCKS = (0, 15, 16, 31, 32, 47)


@pytest.fixture(scope="module")
def adi48_gpu():
    # the GPU run with the default partitioned solve and with the sequential sweeps
    sizes = adi_sizes()
    out = {}
    for algo in ("partitioned", "thomas"):
        old = os.environ.pop("OPS_TRID_ALGO", None)
        try:
            if algo == "thomas":
                os.environ["OPS_TRID_ALGO"] = "thomas"
            out[algo] = dict(zip(ADI_BLOCKS, run("gpu", sizes, 100, fetch=list(ADI_BLOCKS))))
        finally:
            os.environ.pop("OPS_TRID_ALGO", None)
            if old is not None:
                os.environ["OPS_TRID_ALGO"] = old
    return sizes, out


@pytest.mark.gpu
@pytest.mark.parametrize("k", ADI_BLOCKS)
def test_gpu_48_block_configuration_equals_oracle(adi48_gpu, k):
    # the benchmark configuration, 100 steps: block k of the GPU run against the oracle
    # running block k alone from its skip-ahead seed: bitwise with the sequential sweeps,
    # within 1e-12 with the partitioned solve
    sizes, gpu = adi48_gpu
    cpu = run("cpu", [sizes[k]], 100, seed=skip_seed(5, sizes, k), threads=16)[0]
    assert np.array_equal(gpu["thomas"][k], cpu)
    assert close(gpu["partitioned"][k], cpu)


@pytest.mark.gpu
def test_gpu_graph_replay_equals_eager(monkeypatch):
    # the step is recorded once and replayed; replays separated by fetches (host
    # synchronisation) must match eager launches after every replay
    sizes = [(40, 33, 17), (9, 65, 30)]

    def staged(graph):
        monkeypatch.setenv("OPS_ADI_GRAPH", "1" if graph else "0")
        L = lib("adi", "gpu")
        sz = np.array([v for s in sizes for v in s], dtype=np.int32)
        argc, av = argv()
        assert L.adi_init(argc, av, len(sizes), p(sz), 1, 9, 1.0) == 0
        snaps = []
        for n in (3, 2, 4):
            assert L.adi_step(n) == 0
            for k in range(len(sizes)):
                a = np.zeros(L.adi_fetch(k, None))
                L.adi_fetch(k, p(a))
                snaps.append(a)
        assert L.adi_finalize() == 0
        return snaps

    for a, b in zip(staged(True), staged(False)):
        assert np.array_equal(a, b)
