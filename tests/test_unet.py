"""Warm-start U-Net: the numpy oracle's backward pass against finite differences (float64),
the device forward pass and every parameter gradient against the oracle (BF16 points
rounded alike), and training on TeaLeaf solves lowering the loss and the initial residual
of the implicit solve."""
import sys
from pathlib import Path

import ctypes

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from oracle import unet_ref as U  # noqa: E402


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    lay, n = U.layout()
    p = np.zeros(n)
    for wo, bo, co, cin in lay:
        p[wo:bo] = rng.normal(0, scale * np.sqrt(2.0 / (9 * cin)), bo - wo)
        p[bo:bo + co] = rng.normal(0, 0.05, co)
    return p


def random_input(seed, B, H, W):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, H, W, 16))
    x[..., :3] = rng.uniform(-1, 1, (B, H, W, 3))
    return x


def test_oracle_backward_matches_finite_differences():
    p = random_params(1)
    x = random_input(2, 1, 8, 8)
    rng = np.random.default_rng(3)
    go = rng.normal(0, 1, (1, 8, 8))
    g = U.backward(p, x, go, bf=False)

    def loss(q):
        return float(np.sum(U.forward(q, x, bf=False)[0] * go))

    lay, n = U.layout()
    idx = []
    for wo, bo, co, cin in lay:
        idx += list(rng.choice(np.arange(wo, bo), 3, replace=False)) + [bo + int(rng.integers(co))]
    for i in idx:
        e = np.zeros(n)
        e[i] = 1e-6
        fd = (loss(p + e) - loss(p - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# (2, 32, 24): widths that are not multiples of a wavefront's row segment at any level;
# (1, 64, 512): widths 512 .. 64, multiples of it at every level
@pytest.mark.gpu
@pytest.mark.parametrize("B,H,W", [(2, 32, 24), (1, 64, 512)])
def test_gpu_forward_and_gradients_match_oracle(B, H, W):
    import ops_amd
    L = ops_amd.app("tealeaf")
    p = random_params(5).astype(np.float32)
    x = random_input(6, B, H, W).astype(np.float32)
    go = np.random.default_rng(7).normal(0, 1, (B, H, W)).astype(np.float32)
    assert L.tea_unet_init(1) == 0
    n = L.tea_unet_num_params()
    assert n == U.layout()[1]
    assert L.tea_unet_set_params(p.ctypes.data_as(ctypes.c_void_p)) == 0
    pred = np.zeros((B, H, W), np.float32)
    assert L.tea_unet_forward_host(B, H, W, x.ctypes.data_as(ctypes.c_void_p), pred.ctypes.data_as(ctypes.c_void_p)) == 0
    ref_pred, _ = U.forward(p.astype(np.float64), x.astype(np.float64), bf=True)
    assert rel(pred, ref_pred) < 2e-2
    gp = np.zeros(n, np.float32)
    assert L.tea_unet_backward_host(B, H, W, x.ctypes.data_as(ctypes.c_void_p), go.ctypes.data_as(ctypes.c_void_p),
                                    gp.ctypes.data_as(ctypes.c_void_p)) == 0
    ref = U.backward(p.astype(np.float64), x.astype(np.float64), go.astype(np.float64), bf=True)
    lay, _ = U.layout()
    for l, (wo, bo, co, cin) in enumerate(lay):
        assert rel(gp[wo:bo], ref[wo:bo]) < 3e-2, ("weights", l, rel(gp[wo:bo], ref[wo:bo]))
        assert rel(gp[bo:bo + co], ref[bo:bo + co]) < 3e-2, ("bias", l)
    assert L.tea_unet_free() == 0


# Small-scale version of the learned warm start: 64^2 decks with dt = 1 have the
# conduction number rx = dt / dx^2 = 41 of the 1024^2 decks at dt = 0.004.
SMALL, SMALL_DT = 64, 1.0
HELD_OUT = 20261203 + 5000  # deck seed outside the training decks


@pytest.fixture(scope="module")
def trained():
    import ops_amd
    from ops_amd.apps.tealeaf import learned
    L = ops_amd.app("tealeaf")
    assert L.tea_unet_init(11) == 0
    learned.generate(L, 384, cells=SMALL, steps_per_deck=4, seed0=20261203, dt=SMALL_DT)
    loss, _ = learned.train(L, 60, batch=8, lr=2e-3, seed=3)
    yield L, loss
    L.tea_unet_free()


@pytest.mark.gpu
def test_training_lowers_the_loss(trained):
    L, loss = trained
    assert np.all(np.isfinite(loss))
    assert loss[-5:].mean() < 0.1 * loss[:2].mean(), loss


def apply_A(x, kx, ky, rx, ry):
    """the implicit operator on an (n+2)^2 field, [y][x] (tea_leaf_cg_calc_w_reduce_kernel)"""
    n = x.shape[0] - 2
    c = np.s_[1:n + 1, 1:n + 1]
    kxc, kxp, kyc, kyp = kx[c], kx[1:n + 1, 2:n + 2], ky[c], ky[2:n + 2, 1:n + 1]
    out = np.zeros_like(x)
    out[c] = ((1 + ry * (kyp + kyc) + rx * (kxp + kxc)) * x[c] - ry * (kyp * x[2:, 1:n + 1] + kyc * x[:n, 1:n + 1])
              - rx * (kxp * x[1:n + 1, 2:] + kxc * x[1:n + 1, :n]))
    return out


def first_step_energy_ratio(L, cells):
    """energy-norm error of the learned warm start over that of the plain start u = u0, on the
    first solve of the held-out deck: sqrt(1 - drop / delta^T A delta), delta = u - u0 of
    the converged plain solve (A delta = r0, so delta^T A delta = delta^T r0)"""
    from ops_amd.apps.tealeaf import learned
    from ops_amd.apps.tealeaf.decks import tea_deck
    deck = tea_deck(cells, 1, seed=HELD_OUT, eps=1e-12, dt=SMALL_DT, max_iters=20000)
    names = ("u0", "u", "vector_Kx", "vector_Ky")
    plain = learned.solve_run(L, deck, 1, 1, fetch=names)
    u0, u, kx, ky = (plain[k].reshape(cells + 2, cells + 2) for k in names)
    rx = ry = SMALL_DT / (10.0 / cells) ** 2
    r0 = u0 - apply_A(u0, kx, ky, rx, ry)
    energy = float(np.sum((u - u0) * r0))
    warm = learned.solve_run(L, deck, 1, 1, net_cells=SMALL, mode=2)
    return np.sqrt(max(0.0, 1.0 - warm["energy_drop"][0] / energy)), plain, warm


@pytest.mark.gpu
@pytest.mark.parametrize("cells,bound", [(SMALL, 0.5), (2 * SMALL, 0.75)])
def test_warm_start_cuts_the_first_solve_error(trained, cells, bound):
    """On a deck outside the training seeds, the trained network's warm start cuts the
    energy-norm error of the first solve's starting guess (where the state is the
    discontinuous initial layout) against the plain start (below `bound` x), and lowers the
    starting residual; the trained network removes at least 1.5x the squared energy error
    that the same network untrained removes (a Galerkin-scaled random smooth field already
    removes some). At 128^2 the network sees
    2 x 2 averages (the inference-at-a-multiple-of-the-training-size path), which cannot
    carry the cell-scale part of the change, hence the looser bound."""
    L, _ = trained
    ratio, plain, warm = first_step_energy_ratio(L, cells)
    assert warm["alpha"][0] > 0.0
    assert ratio < bound, ratio
    assert warm["warm"][0] < plain["initial"][0], (warm["warm"], plain["initial"])
    n = L.tea_unet_num_params()
    keep = np.zeros(n, np.float32)
    assert L.tea_unet_get_params(keep.ctypes.data_as(ctypes.c_void_p)) == 0
    try:
        assert L.tea_unet_init(11) == 0  # same initial weights as before training
        untrained, _, _ = first_step_energy_ratio(L, cells)
    finally:
        assert L.tea_unet_set_params(keep.ctypes.data_as(ctypes.c_void_p)) == 0
    assert 1.0 - ratio ** 2 > 1.5 * (1.0 - untrained ** 2), (untrained, ratio)


@pytest.mark.gpu
@pytest.mark.parametrize("solver", [0, 1, 3])
def test_warm_started_solves_reach_the_same_field(trained, solver):
    """Every solver accepts the warm start, stops on the plain start's threshold and ends
    on the field of its plain run. The warm start is kept only where its residual is below
    the plain start's (mode 1)."""
    from ops_amd.apps.tealeaf import learned
    from ops_amd.apps.tealeaf.decks import tea_deck
    L, _ = trained
    deck = tea_deck(SMALL, 3, seed=HELD_OUT + 1, eps=1e-10, dt=SMALL_DT, max_iters=20000)
    plain = learned.solve_run(L, deck, solver, 3, fetch=("u",))
    warm = learned.solve_run(L, deck, solver, 3, net_cells=SMALL, fetch=("u",))
    assert all(w == -1 for w in plain["warm"])
    for w, i, k in zip(warm["warm"], warm["initial"], warm["kept"]):
        assert k == (1 if w < i else 0)
    assert all(r < 1e-10 for r in warm["rel"]) and all(r < 1e-10 for r in plain["rel"])
    assert np.max(np.abs(warm["u"] - plain["u"])) <= 1e-8 * np.max(np.abs(plain["u"]))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [3, 4])
def test_corrected_preconditioner_reaches_the_same_field(trained, mode):
    """The network's change as a rank-one A-orthogonal correction of the multigrid
    preconditioner, with the warm start (mode 3) or from the plain start (mode 4): the
    multigrid PCG still stops on the plain start's threshold and ends on the field of its
    plain run."""
    from ops_amd.apps.tealeaf import learned
    from ops_amd.apps.tealeaf.decks import tea_deck
    L, _ = trained
    deck = tea_deck(SMALL, 3, seed=HELD_OUT + 1, eps=1e-10, dt=SMALL_DT, max_iters=20000)
    plain = learned.solve_run(L, deck, 3, 3, fetch=("u",))
    corr = learned.solve_run(L, deck, 3, 3, net_cells=SMALL, fetch=("u",), mode=mode)
    assert all(k == (1 if mode == 3 else 0) for k in corr["kept"])
    assert all(r < 1e-10 for r in corr["rel"])
    assert np.max(np.abs(corr["u"] - plain["u"])) <= 1e-8 * np.max(np.abs(plain["u"]))
