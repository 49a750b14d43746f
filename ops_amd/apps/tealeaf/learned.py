"""Learned warm start for the TeaLeaf implicit solve: data from the method's own solves,
training on the device, and solves that start from the network's prediction.

The network and its kernels are in native/tealeaf_unet.hip (C-ABI in tealeaf_unet.h); the
TeaLeaf driver exposes tealeaf_capture (store the last solve as a training sample) and
tealeaf_set_guess (evaluate the network before every solve). Training pairs come from
seeded rectangle decks (decks.tea_deck): each deck runs `steps_per_deck` implicit steps and
every step's solve becomes one sample, (density, energy, u0, residual of u0) -> u.
"""
import ctypes
import tempfile
import time
from pathlib import Path

import numpy as np

from .decks import tea_deck

# the network's convolutions (native/tealeaf_unet.hip): input channels from the upsampled
# source, from the same-resolution source, output channels, resolution level (1/2^level)
UNET_LAYERS = [(0, 16, 16, 0), (0, 16, 32, 1), (0, 32, 48, 2), (0, 48, 64, 3),
               (64, 48, 48, 2), (48, 32, 32, 1), (32, 16, 16, 0), (0, 16, 16, 0)]


def forward_flops_per_pixel():
    """multiply-adds x 2 of one forward pass per input pixel: sum over the 3 x 3
    convolutions of 2 * 9 * Cin * Cout at each layer's resolution"""
    return sum(2 * 9 * (c1 + c2) * co / 4 ** lv for c1, c2, co, lv in UNET_LAYERS)


def training_flops_per_pixel():
    """forward + weight gradient (the same products) + input gradient (every layer but the
    first)"""
    c1, c2, co, lv = UNET_LAYERS[0]
    return 3 * forward_flops_per_pixel() - 2 * 9 * (c1 + c2) * co / 4 ** lv


def _run_deck(L, text, solver):
    import ops_amd
    t = tempfile.TemporaryDirectory()
    deck = Path(t.name) / "tea.in"
    deck.write_text(text)
    argc, argv = ops_amd.argv()
    if L.tealeaf_init(argc, argv, str(deck).encode()) != 0:
        t.cleanup()
        raise RuntimeError("tealeaf_init failed")
    if L.tealeaf_set_solver(solver) != 0:
        t.cleanup()
        raise RuntimeError("tealeaf_set_solver failed")
    return t


def generate(L, count, cells=1024, steps_per_deck=5, seed0=20261203, dt=0.004, solver=3, eps=1e-10):
    """Fill the device dataset with `count` samples at cells^2; returns seconds taken."""
    t0 = time.perf_counter()
    if L.tea_unet_dataset_reserve(count, cells, cells) != 0:
        raise RuntimeError("tea_unet_dataset_reserve failed")
    idx, k = 0, 0
    while idx < count:
        t = _run_deck(L, tea_deck(cells, steps_per_deck, seed=seed0 + k, eps=eps, dt=dt), solver)
        try:
            for _ in range(steps_per_deck):
                if idx >= count:
                    break
                if L.tealeaf_capture_next(idx) != 0 or L.tealeaf_step(1) != 0:
                    raise RuntimeError("tealeaf step/capture failed")
                idx += 1
        finally:
            L.tealeaf_finalize()
            t.cleanup()
        k += 1
    return time.perf_counter() - t0


def train(L, epochs, batch=4, lr=1e-3, seed=1, beta=0.0):
    """Adam over the dataset; returns (loss per epoch, seconds). beta weights the conduction
    part of the loss norm (0: relative mean squared error, 1: energy norm)."""
    loss = np.zeros(epochs, np.float32)
    t0 = time.perf_counter()
    if L.tea_unet_train(epochs, batch, ctypes.c_float(lr), ctypes.c_float(beta), seed,
                        loss.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise RuntimeError("tea_unet_train failed")
    return loss, time.perf_counter() - t0


def solve_run(L, text, solver, steps, net_cells=0, fetch=(), mode=1, starts=None):
    """Run `steps` implicit steps of a deck, with the learned warm start when net_cells > 0
    (mode 1: kept when its residual beats the standard start's, mode 2: always kept, mode 3:
    always kept and the network's change also corrects the multigrid preconditioner, mode
    4: that correction alone), or from given starting fields (starts(k) -> the field of
    step k, halo included, in tealeaf_fetch's layout).
    Returns per-step iterations, standard / warm initial residuals, warm start kept, its
    Galerkin scaling and energy-norm drop, the final CG residual ratio, seconds, and the
    fields named in `fetch` (halo included) after the last step."""
    import ops_amd
    R = ops_amd.runtime()
    t = _run_deck(L, text, solver)
    try:
        if L.tealeaf_set_guess(mode if net_cells else 0, net_cells) != 0:
            raise RuntimeError("tealeaf_set_guess failed")
        st, gs = np.zeros(7), np.zeros(5)
        out = {"iters": [], "initial": [], "warm": [], "kept": [], "alpha": [], "energy_drop": [], "rel": []}
        R.ops_hip_device_sync()
        t0 = time.perf_counter()
        cnt = L.tealeaf_fetch(b"u", None)
        for k in range(steps):
            if starts is not None:
                f = np.ascontiguousarray(starts(k), dtype=np.float64)
                if L.tealeaf_set_start(f.ctypes.data_as(ctypes.c_void_p), cnt) != 0:
                    raise RuntimeError("tealeaf_set_start failed")
            if L.tealeaf_step(1) != 0:
                raise RuntimeError("tealeaf_step failed")
            L.tealeaf_state(st.ctypes.data_as(ctypes.c_void_p))
            L.tealeaf_guess_state(gs.ctypes.data_as(ctypes.c_void_p))
            out["iters"].append(int(st[2]))
            out["rel"].append(float(st[3] / st[4]))
            out["initial"].append(float(gs[0]))
            out["warm"].append(float(gs[1]))
            out["kept"].append(int(gs[2]))
            out["alpha"].append(float(gs[3]))
            out["energy_drop"].append(float(gs[4]))
        R.ops_hip_device_sync()
        out["s"] = time.perf_counter() - t0
        for name in fetch:
            n = L.tealeaf_fetch(name.encode(), None)
            v = np.zeros(n)
            L.tealeaf_fetch(name.encode(), v.ctypes.data_as(ctypes.c_void_p))
            out[name] = v
    finally:
        L.tealeaf_set_guess(0, 0)
        L.tealeaf_finalize()
        t.cleanup()
    return out


def best_case(L, text, solver, steps, errors=(0.1, 0.01, 0.001), seed=1):
    """What any warm start could save: the solve of each step starts from its converged
    solution (from a plain run of the same deck) with a smooth error of relative size e
    (in the 2-norm of the step's change u - u0: a sum of six random low-frequency sine
    products, the kind of error a network's smooth prediction leaves), for each e in
    `errors`. Returns {"plain": per-step iterations, e: per-step iterations}."""
    base = solve_run(L, text, solver, steps, fetch=())
    us, u0s = [], []
    # the converged u and the u0 of every step, from a second plain run that fetches both
    import ops_amd
    t = _run_deck(L, text, solver)
    try:
        cnt = L.tealeaf_fetch(b"u", None)
        for _ in range(steps):
            if L.tealeaf_step(1) != 0:
                raise RuntimeError("tealeaf_step failed")
            u, v = np.zeros(cnt), np.zeros(cnt)
            L.tealeaf_fetch(b"u", u.ctypes.data_as(ctypes.c_void_p))
            L.tealeaf_fetch(b"u0", v.ctypes.data_as(ctypes.c_void_p))
            us.append(u)
            u0s.append(v)
    finally:
        L.tealeaf_finalize()
        t.cleanup()
    m = int(round(np.sqrt(cnt)))
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(m), np.arange(m))
    out = {"plain": base["iters"]}
    for e in errors:
        def starts(k):
            f = np.zeros((m, m))
            for _ in range(6):
                kx, ky = rng.integers(1, 6, 2)
                ph = rng.uniform(0.0, 2.0 * np.pi, 2)
                f += np.sin(np.pi * kx * X / m + ph[0]) * np.sin(np.pi * ky * Y / m + ph[1])
            d = us[k] - u0s[k]
            f = f.ravel() * (np.linalg.norm(d) / np.linalg.norm(f))
            return us[k] + e * f
        out[e] = solve_run(L, text, solver, steps, starts=starts)["iters"]
    return out
