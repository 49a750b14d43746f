"""Laplace2D Jacobi: oracle pinned to the reference, GPU path against the oracle."""
import numpy as np
import pytest

import oracle
from helpers import first_diff, max_rel, ptr, sha256


def run_oracle(nx, ny, iters, mode, threads=8):
    A = np.zeros((ny + 2, nx + 2))
    An = np.zeros_like(A)
    e = oracle.lib().oracle_laplace(nx, ny, iters, mode, ptr(A), ptr(An), threads)
    return A, An, e


# ---------------------------------------------------------------- oracle pins (CPU)

def test_oracle_bitwise_equals_reference_step7(golden):
    fx = golden("laplace_step7.json")
    A, An, e = run_oracle(fx["nx"], fx["ny"], fx["iters"], fx["mode"])
    assert sha256(A) == fx["dats"]["A"]["sha256"]
    assert sha256(An) == fx["dats"]["Anew"]["sha256"]
    assert e == fx["reductions"]["error"]["values"][0]


def test_oracle_reference_known_answer(golden):
    # apps/c/laplace2d_tutorial/step7/laplace2d.cpp:134-137
    fx = golden("laplace_step7.json")
    _, _, e = run_oracle(fx["nx"], fx["ny"], fx["iters"], fx["mode"])
    assert abs(100.0 * (e / fx["golden_error"]) - 100.0) < 0.001


def test_oracle_independent_of_thread_count():
    a1, _, e1 = run_oracle(300, 200, 40, 0, threads=1)
    a8, _, e8 = run_oracle(300, 200, 40, 0, threads=8)
    assert first_diff(a1, a8) is None and e1 == e8


def test_oracle_mirror_symmetry_config1():
    # top-boundary problem is symmetric in x; a swapped axis or a wrong neighbour breaks it
    A, _, _ = run_oracle(128, 96, 200, 0)
    top = A[-1, 1:-1]
    if np.array_equal(top, top[::-1]):
        assert np.array_equal(A[:, 1:-1], A[:, 1:-1][:, ::-1])
    else:
        # sin(pi) is 1.2e-16, not 0: the mirror image agrees to rounding level
        assert np.max(np.abs(A[:, 1:-1] - A[:, 1:-1][:, ::-1])) < 1e-14
    assert not np.allclose(A[1:-1, 1:-1], A[1:-1, 1:-1][::-1, :])


def test_oracle_fixed_point_is_harmonic():
    # converged interior satisfies the discrete Laplace equation (mean of neighbours)
    A, _, e = run_oracle(24, 24, 4000, 0)
    assert e < 1e-12
    c = A[1:-1, 1:-1]
    nb = 0.25 * (A[1:-1, 2:] + A[1:-1, :-2] + A[2:, 1:-1] + A[:-2, 1:-1])
    assert np.max(np.abs(c - nb)) < 1e-11


# ---------------------------------------------------------------- GPU path

def run_gpu(nx, ny, iters, mode):
    import ops_amd
    L = ops_amd.app("laplace2d")
    argc, argv = ops_amd.argv()
    assert L.laplace2d_init(argc, argv, nx, ny, mode) == 0
    err = np.zeros(1)
    assert L.laplace2d_iterate(iters, ptr(err)) == 0
    A = np.zeros((ny + 2, nx + 2))
    assert L.laplace2d_fetch_full(ptr(A)) == 0
    assert L.laplace2d_finalize() == 0
    return A, float(err[0])


@pytest.mark.gpu
def test_gpu_bitwise_equals_reference_step7(golden):
    fx = golden("laplace_step7.json")
    A, e = run_gpu(fx["nx"], fx["ny"], fx["iters"], fx["mode"])
    assert sha256(A) == fx["dats"]["A"]["sha256"], first_diff(A, run_oracle(fx["nx"], fx["ny"], fx["iters"], 1)[0])
    assert e == fx["reductions"]["error"]["values"][0]


@pytest.mark.gpu
def test_gpu_config1_bitwise_vs_oracle():
    # 1024 x 1024 interior, top boundary sin(pi x / 1023), 1000 iterations
    A, e = run_gpu(1024, 1024, 1000, 0)
    Ao, _, eo = run_oracle(1024, 1024, 1000, 0)
    assert first_diff(A, Ao) is None
    assert e == eo


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 2), (7, 3), (65, 5), (130, 257), (3, 67)])
def test_gpu_ragged_shapes_bitwise(shape):
    nx, ny = shape
    A, e = run_gpu(nx, ny, 13, 1)
    Ao, _, eo = run_oracle(nx, ny, 13, 1)
    assert first_diff(A, Ao) is None and e == eo


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (9, 1)])
def test_gpu_edge_dat_extent_is_rejected(shape):
    # a block extent of 1 makes an edge dat of size 1 in that dimension (ops_lib_core.cpp:
    # 474-481), so the one-cell halo loops address outside it: the loop raises instead
    import ctypes
    import ops_amd
    L = ops_amd.app("laplace2d")
    R = ops_amd.runtime()
    R.ops_get_last_error.restype = ctypes.c_char_p
    argc, argv = ops_amd.argv()
    assert L.laplace2d_init(argc, argv, shape[0], shape[1], 1) != 0
    assert b"outside its allocation" in R.ops_get_last_error()
    assert L.laplace2d_finalize() == 0
    A, e = run_gpu(16, 16, 5, 1)  # the runtime is usable again afterwards
    Ao, _, eo = run_oracle(16, 16, 5, 1)
    assert first_diff(A, Ao) is None and e == eo


@pytest.mark.gpu
def test_gpu_device_boundary_loops_with_idx():
    # mode 2 evaluates the tutorial boundaries on the GPU from ops_arg_idx (device sin)
    A, e = run_gpu(200, 150, 30, 2)
    Ao, _, eo = run_oracle(200, 150, 30, 1)
    assert max_rel(A, Ao, floor=1e-300) < 1e-13
    assert abs(e - eo) <= 1e-13 * abs(eo)
