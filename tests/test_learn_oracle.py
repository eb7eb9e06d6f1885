"""Pins of the learned warm start's CPU oracle (oracle/learn_oracle.py): bf16
rounding against torch, grids against a per-axis linear search and a frame
invariance, gradients against finite differences, the pose correction against its
round trip."""
import numpy as np
import pytest

import learn_oracle as L


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([(rng.normal(size=20000) * 10.0 ** rng.integers(-8, 8, 20000)).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3e38, -3e38, 1e-40], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(L.bf16_bits(x), ref)
    assert np.isnan(L.bf16_value(L.bf16_bits(np.array([np.nan], np.float32)))[0])


def test_log1p_table():
    import torch
    lut = L.log1p_lut()
    c = np.arange(65536)
    ref = torch.from_numpy(np.log1p(c).astype(np.float32)).to(torch.bfloat16).float().numpy()
    assert np.array_equal(L.bf16_value(lut), ref)
    assert lut[0] == 0


def _linear_search_cell(v, lo, step, n):
    for i in range(n):
        if lo + step * i <= v < lo + step * (i + 1):
            return i
    return -1


def test_bev_index_against_linear_search():
    rng = np.random.default_rng(1)
    pts = np.column_stack([rng.uniform(-18, 18, 400), rng.uniform(-18, 18, 400), rng.uniform(-4, 11, 400)])
    # exact cell boundaries and edges
    pts = np.vstack([pts, [[-16.0, -16.0, -2.5], [15.999, 15.5, 9.49], [16.0, 0, 0], [0, 0, 9.5], [np.nan, 0, 0]]])
    got = L.bev_index(pts[:, 0], pts[:, 1], pts[:, 2])
    for p, g in zip(pts, got):
        ix = _linear_search_cell(p[0], -16.0, 0.5, 64)
        iy = _linear_search_cell(p[1], -16.0, 0.5, 64)
        iz = _linear_search_cell(p[2], -2.5, 1.5, 8)
        want = -1 if min(ix, iy, iz) < 0 else (iz * 64 + iy) * 64 + ix
        assert g == want, (p, g, want)


def test_scan_grid_counts():
    rng = np.random.default_rng(2)
    s = np.zeros((3000, 4), np.float32)
    s[:, :3] = rng.uniform(-20, 20, (3000, 3))
    grid = L.scan_grid(s)
    idx = L.bev_index(s[:, 0], s[:, 1], s[:, 2])
    counts = np.round(np.expm1(L.bf16_value(grid).astype(np.float64)))
    # log1p in bf16 is exact enough to recover small counts
    small = np.bincount(idx[idx >= 0], minlength=L.CELLS)
    assert np.all(np.abs(counts - small) <= np.maximum(1, 0.01 * small))
    assert (grid != 0).sum() == (small > 0).sum()


def test_map_grid_equals_scan_grid_of_transformed_map():
    """The map grid at guess G is the scan grid of the map expressed in G's frame."""
    rng = np.random.default_rng(3)
    m = np.zeros((20000, 4), np.float32)
    m[:, :2] = rng.uniform(-30, 30, (20000, 2))
    m[:, 2] = rng.uniform(-3, 12, 20000)
    G = L.pose_matrix(np.array([1.3, -2.2, 0.4, 0.01, -0.02, 0.7]))
    q = (np.linalg.inv(G) @ np.column_stack([m[:, :3].astype(np.float64), np.ones(len(m))]).T).T[:, :3]
    a = L.map_grid(m, G)
    # bin the transformed points in float64 directly: identical except where the two
    # evaluations of q straddle a cell boundary
    idx = L.bev_index(q[:, 0], q[:, 1], q[:, 2])
    b = L._grid_from_index(idx, L.log1p_lut())
    assert (a != b).sum() <= 3
    assert (a != 0).sum() > 1000


def test_correction_round_trip():
    rng = np.random.default_rng(4)
    for _ in range(50):
        G = L.pose_matrix(np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-0.3, 0.3, 3)]))
        T = G @ L.pose_matrix(np.concatenate([rng.uniform(-3, 3, 3), rng.uniform(-0.2, 0.2, 3)]))
        p6 = L.correction_params(G, T)
        assert np.allclose(L.apply_correction(G, p6), T, atol=1e-9)


def _small_params(rng, f0=12, f1=8, f2=5):
    return {"W1": rng.normal(size=(f1, f0)) * 0.5, "b1": rng.normal(size=f1) * 0.1,
            "W2": rng.normal(size=(f2, f1)) * 0.5, "b2": rng.normal(size=f2) * 0.1,
            "W3": rng.normal(size=(6, f2)) * 0.5, "b3": rng.normal(size=6) * 0.1}


@pytest.mark.parametrize("name", ["W1", "b1", "W2", "b2", "W3", "b3"])
def test_mlp_gradients_finite_differences(name):
    rng = np.random.default_rng(5)
    P = _small_params(rng)
    x = rng.normal(size=(7, 12))
    T = rng.normal(size=(7, 6))
    w = rng.uniform(0.5, 2.0, 6)
    _, g, _ = L.mlp_backward(x, P, T, w, bf16=False)
    h = 1e-6
    for _ in range(8):
        i = tuple(rng.integers(0, s) for s in P[name].shape)
        Pp = {k: v.copy() for k, v in P.items()}
        Pm = {k: v.copy() for k, v in P.items()}
        Pp[name][i] += h
        Pm[name][i] -= h
        lp = L.mlp_loss(L.mlp_forward(x, Pp, False)["Y"], T, w)
        lm = L.mlp_loss(L.mlp_forward(x, Pm, False)["Y"], T, w)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[name][i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[name][i])


def test_mlp_bf16_mode_stays_close_to_exact():
    rng = np.random.default_rng(6)
    P = _small_params(rng, 64, 32, 16)
    x = L.bf16_round(rng.normal(size=(16, 64)).astype(np.float32)).astype(np.float64)
    a = L.mlp_forward(x, P, True)["Y"]
    b = L.mlp_forward(x, P, False)["Y"]
    assert np.abs(a - b).max() < 0.05 * np.abs(b).max()


def test_adam_first_step_is_signed_lr():
    g = np.array([0.3, -2.0, 1e-3])
    w, m, v = L.adam_step(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, lr=0.01, eps=0.0)
    assert np.allclose(w, -0.01 * np.sign(g))
