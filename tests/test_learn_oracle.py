"""Pins of the learned warm start's CPU oracle (oracle/learn_oracle.py): bf16
rounding against torch, grids against a per-axis linear search and a frame
invariance, gradients against finite differences, the pose correction against its
round trip; the registration's scores against a cell-by-cell count, its choice
among ties, its parabola and small-angle series against closed forms and libm,
and its result against planted shifts and yaws; the leveling's cells against a
linear search, its plane against least squares, and its result against planted
tilts of level and of tilted true poses."""
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


# ---------------------------------------------------------------- registration

def _cells(rows):
    """(8, 64, 64) bool of (512,) uint64 rows"""
    r = np.asarray(rows, np.uint64).reshape(L.NZ, L.NY)
    return np.stack([(r >> np.uint64(x)) & np.uint64(1) for x in range(L.NX)], axis=-1).astype(bool)


def test_register_scores_brute_force():
    """score[k, dy, dx] = number of cells c with scan(c) and map_k(c + (dx, dy)),
    counted cell by cell on random occupancies (a shift of the wrong sign or axis
    fails)."""
    rng = np.random.default_rng(11)
    S = rng.uniform(size=(L.NZ, L.NY, L.NX)) < 0.08
    Ms = [rng.uniform(size=(L.NZ, L.NY, L.NX)) < 0.08 for _ in range(2)]
    srows = L.occupancy_rows(S.reshape(-1))
    mrows = np.stack([L.occupancy_rows(M.reshape(-1)) for M in Ms])
    assert np.array_equal(_cells(srows), S)
    R = 3
    sc = L.register_scores(srows, mrows, R)
    for k, M in enumerate(Ms):
        for dy in range(-R, R + 1):
            for dx in range(-R, R + 1):
                n = 0
                for y in range(L.NY):
                    for x in range(L.NX):
                        yy, xx = y + dy, x + dx
                        if 0 <= yy < L.NY and 0 <= xx < L.NX:
                            n += int(np.sum(S[:, y, x] & M[:, yy, xx]))
                assert sc[k, dy + R, dx + R] == n, (k, dy, dx)


def test_register_pick_ties_and_empty():
    rp = dict(radius=2, yaws=3, yaw_step=0.01, refine_yaw=1)
    sc = np.zeros((3, 5, 5), np.int64)
    assert L.register_pick(sc, rp) == (0, 1, 0, 0)
    sc[1, 2, 3] = sc[1, 2, 1] = 4          # dx = +1 and dx = -1 tie: positive first
    assert L.register_pick(sc, rp) == (4, 1, 0, 1)
    sc[0, 2, 2] = 4                        # the same score at the centre shift of another yaw
    assert L.register_pick(sc, rp) == (4, 1, 0, 1)
    sc[2, 0, 0] = 5
    assert L.register_pick(sc, rp) == (5, 2, -2, -2)


def test_parabola_vertex_and_small_angles():
    import math
    for x0 in (-0.5, -0.25, 0.0, 0.125, 0.375, 0.5):
        f = [-(x - x0) ** 2 + 7.0 for x in (-1.0, 0.0, 1.0)]
        assert L._parabola(*f) == x0
    assert L._parabola(1.0, 1.0, 1.0) == 0.0
    for d in np.linspace(-0.025, 0.025, 41):
        c, s = L._small_cos_sin(float(d))
        assert abs(c - math.cos(d)) <= 2.3e-16 and abs(s - math.sin(d)) <= 2.3e-16


def _scene(A):
    mp = A.make_map(A.scene_cfg(4, extent=80.0), 80000)
    G = L.pose_matrix([3.0, -5.0, 1.6, 0.0, 0.0, 0.4])
    Gi = np.linalg.inv(G)
    p = np.column_stack([mp[:, :3].astype(np.float64), np.ones(len(mp))]) @ Gi.T
    scan = np.column_stack([p[:, :3], np.ones(len(p))]).astype(np.float32)
    return mp, G, scan


@pytest.mark.parametrize("shift,steps", [((1.0, -1.5, 0.0), 0), ((0.0, 0.0, 0.75), 2), ((-2.0, 0.5, -0.4), -3),
                                         ((2.5, 2.0, 1.2), 3), ((-2.5, -3.0, 0.0), 1)])
def test_register_recovers_planted_offsets(A, shift, steps):
    """A scan made of the map's own points seen from the true pose T, and the guess
    G = T T(-d) Rz(-steps yaw_step), d whole cells on x and y and whole height bins
    on z, the yaw within the searched ones (the outermost without a neighbour to
    refine with): the registration must return T within 3 cm on x and y, 1 cm on z
    and 0.05 degrees (observed: 1 cm, 0.2 mm, 0.008 degrees; the shift applied
    along the guess's axes instead of the searched yaw's would miss by 8 cm at 3
    steps, a height shift of the wrong sign by twice d_z).  Without the height
    search the height stays as it is."""
    mp, T, scan = _scene(A)
    rp = L.register_params_default()
    d = np.array([shift[0], shift[1], shift[2], 0.0, 0.0, 0.0])
    G = T @ L.pose_matrix(-d) @ L.pose_matrix([0, 0, 0, 0, 0, -steps * rp["yaw_step"]])
    gl, gw, v = L.register(scan, mp, G, G, rp)
    assert np.array_equal(gl, gw) and v > 1000
    p = L.correction_params(gl, T)
    assert np.hypot(p[0], p[1]) < 0.03 and abs(p[2]) < 0.01, p
    assert abs(p[5]) < np.radians(0.05), p
    gl2, _, _ = L.register(scan, mp, G, G, dict(rp, z_search=0, level=0))
    assert abs(L.correction_params(gl2, T)[2] - shift[2]) < 1e-9


def test_height_shift_rules():
    sc = np.zeros(2 * L.ZH_RANGE + 1, np.int64)
    assert L.z_shift(sc) == 0.0
    sc[L.ZH_RANGE + 4] = sc[L.ZH_RANGE - 4] = 10     # a tie: the positive shift
    assert L.z_shift(sc) == 4 * L.ZH_BIN
    sc[L.ZH_RANGE + 3] = 10                          # the same score nearer zero
    assert abs(L.z_shift(sc) - 3.5 * L.ZH_BIN) < 1e-15   # parabola through 0, 10, 10: +0.5
    hs = np.zeros(L.ZH_BINS, np.int64)
    hm = np.zeros(L.ZH_BINS, np.int64)
    hs[100], hm[107] = 3, 5                          # map heights 7 bins above the scan's
    assert L.z_scores(hs, hm)[L.ZH_RANGE + 7] == 15 and L.z_scores(hs, hm).sum() == 15


def test_register_without_common_cells_keeps_the_guess(A):
    mp, T, scan = _scene(A)
    rp = L.register_params_default()
    far = T.copy()
    far[0, 3] += 500.0
    gl, gw, v = L.register(scan, mp, far, far, rp)
    assert v == 0 and np.array_equal(gl, far) and np.array_equal(gw, far)
    high = scan.copy()
    high[:, 2] += 40.0
    gl, _, v = L.register(high, mp, T, T, rp)
    assert v == 0 and np.array_equal(gl, T)


def test_level_cells_against_linear_search():
    rng = np.random.default_rng(2)
    x, y, z = rng.uniform(-18, 18, 3000), rng.uniform(-18, 18, 3000), rng.uniform(-10, 10, 3000)
    x[:5] = [np.nan, 15.999999, -16.0, 16.0, 0.0]
    z[5], z[6], z[7] = 8.0, -8.0, np.nan
    out = L.level_cells(x, y, z)
    ref = np.full(L.LEV_N * L.LEV_N, np.inf)
    for xi, yi, zi in zip(x, y, z):
        i, j = _linear_search_cell(xi, -16.0, 1.0, 32), _linear_search_cell(yi, -16.0, 1.0, 32)
        if i >= 0 and j >= 0 and -8.0 <= zi < 8.0:
            ref[j * 32 + i] = min(ref[j * 32 + i], zi)
    assert np.array_equal(out, ref)


def test_level_plane_against_least_squares():
    """The tree sums and Cramer's rule give numpy's least-squares slopes; cells on
    one row leave the system singular (no tilt)."""
    rng = np.random.default_rng(3)
    i = np.arange(L.LEV_N * L.LEV_N)
    x, y = (i % 32) - 15.5, (i // 32) - 15.5
    for _ in range(6):
        keep = rng.uniform(size=i.size) < rng.uniform(0.05, 0.95)
        d = rng.normal(0, 0.02) * x + rng.normal(0, 0.02) * y + rng.normal(0, 0.5) + rng.normal(0, 0.01, i.size)
        a, b = L.level_solve(L.level_sums(np.where(keep, d, 0.0), keep))
        ref = np.linalg.lstsq(np.column_stack([x[keep], y[keep], np.ones(keep.sum())]), d[keep], rcond=None)[0]
        assert abs(a - ref[0]) < 1e-12 and abs(b - ref[1]) < 1e-12
    row = i < 32
    assert L.level_solve(L.level_sums(np.where(row, 0.1 * x, 0.0), row)) is None


def _tilt(G, T):
    """roll and pitch of T^-1 G"""
    E = np.linalg.inv(T) @ G
    return np.arctan2(E[2, 1], E[2, 2]), np.arcsin(np.clip(-E[2, 0], -1.0, 1.0))


@pytest.mark.parametrize("truth,err", [((0.0, 0.0), (0.4, -0.3)), ((1.5, -1.0), (-0.45, 0.2)),
                                       ((-2.0, 0.5), (0.3, 0.45)), ((0.7, 2.2), (0.0, -0.4))])
def test_level_recovers_planted_tilt(A, truth, err):
    """A scan made of the map's own points seen from a true pose T with the given
    roll and pitch (degrees; the scene's ground is level, so a tilted T sees it
    tilted), and the guess T turned by a further roll and pitch, 20 cm off on x:
    leveling returns T's roll and pitch within 0.02 degrees (the error planted is
    0.3-0.45), whatever T's own tilt, and leaves the translation alone; a tilt
    of the wrong sign would double the error."""
    mp = A.make_map(A.scene_cfg(4, extent=80.0), 80000)
    T = L.pose_matrix([3.0, -5.0, 1.6, np.radians(truth[0]), np.radians(truth[1]), 0.4])
    p = np.column_stack([mp[:, :3].astype(np.float64), np.ones(len(mp))]) @ np.linalg.inv(T).T
    scan = np.column_stack([p[:, :3], np.ones(len(p))]).astype(np.float32)
    G = T @ L.pose_matrix([0.2, 0.0, 0.0, np.radians(err[0]), np.radians(err[1]), 0.0])
    gl, gw, kept, pr = L.level(scan, mp, G, G)
    assert np.array_equal(gl, gw) and kept > 300 and pr is not None
    r, q = _tilt(gl, T)
    assert abs(r) < np.radians(0.02) and abs(q) < np.radians(0.02), np.degrees([r, q])
    assert np.array_equal(gl[:3, 3], G[:3, 3])


def test_level_keeps_a_guess_without_common_cells(A):
    mp, T, scan = _scene(A)
    far = T.copy()
    far[0, 3] += 500.0
    gl, gw, kept, pr = L.level(scan, mp, far, far)
    assert pr is None and kept == 0 and np.array_equal(gl, far) and np.array_equal(gw, far)
