"""CPU oracle of the learned warm start (numpy).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py may import this module, as the
checker of the GPU engine; the product never calls it.

The reference has no learned component: its NDT starts from the guess it is given
(src/CPU/ndt_mapping/kernel.cpp:1187-1206).  This module restates the engine's own
definitions (include/ndt_learn.h):

  bev_index / scan_grid / map_grid    the two 64 x 64 x 8 occupancy grids, log1p
                                      counts in bf16 (bit patterns)
  correction_params / apply_correction   the pose correction the model predicts
  mlp_forward / mlp_backward          the bf16 MLP and its gradients, float64
                                      arithmetic with the engine's bf16 roundings
                                      (or none, for finite-difference checks)
  adam_step                           Adam (Kingma & Ba 2015)
  register                            the BEV registration before the model: yaw-
                                      rotated occupancy of the map, popcount scores
                                      of every shift, the best candidate and its
                                      parabolic refinement

Pins (tests/test_learn_oracle.py): bf16 rounding against torch's bfloat16
conversion; grids against a per-axis linear search on small inputs and against
the invariance scan grid(G^-1 map) = map grid(G); gradients against central
finite differences in float64; the correction against its round trip; the
registration's scores against a cell-by-cell count, its tie rule on constructed
score tables, its parabola and small-angle series against closed forms and libm,
and its result against planted shifts and yaws.  No function is left unpinned.
"""
from __future__ import annotations

import numpy as np

NX, NY, NZ = 64, 64, 8
CELLS = NX * NY * NZ
X0, Y0, Z0, DZ = -16.0, -16.0, -2.5, 1.5
CORR_SCALE = np.array([3.0, 3.0, 1.0, 0.05, 0.05, 0.2])


# ---------------------------------------------------------------- bf16

def bf16_bits(x) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even, NaN kept NaN."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x) -> np.ndarray:
    return bf16_value(bf16_bits(x))


def log1p_lut() -> np.ndarray:
    c = np.arange(65536, dtype=np.float64)
    return bf16_bits(np.log1p(c).astype(np.float32))


# ---------------------------------------------------------------- grids

def bev_index(qx, qy, qz) -> np.ndarray:
    """Cell of each point in the grid frame (float64 arithmetic), -1 outside."""
    fx = np.floor((np.asarray(qx, np.float64) - X0) * 2.0)
    fy = np.floor((np.asarray(qy, np.float64) - Y0) * 2.0)
    fz = np.floor((np.asarray(qz, np.float64) - Z0) / DZ)
    with np.errstate(invalid="ignore"):
        ok = (fx >= 0) & (fx < NX) & (fy >= 0) & (fy < NY) & (fz >= 0) & (fz < NZ)
    idx = np.full(fx.shape, -1, np.int64)
    idx[ok] = ((fz[ok].astype(np.int64) * NY + fy[ok].astype(np.int64)) * NX + fx[ok].astype(np.int64))
    return idx


def _grid_from_index(idx, lut) -> np.ndarray:
    counts = np.bincount(idx[idx >= 0], minlength=CELLS)
    return lut[np.minimum(counts, 65535)]


def scan_grid(scan_xyzw, lut=None) -> np.ndarray:
    """Grid of a scan in its sensor frame: (32768,) bf16 bits."""
    lut = log1p_lut() if lut is None else lut
    s = np.asarray(scan_xyzw, np.float32)
    return _grid_from_index(bev_index(s[:, 0], s[:, 1], s[:, 2]), lut)


def map_grid(map_xyzw, guess_local, lut=None) -> np.ndarray:
    """Grid of the map points around a guess (map-local 4x4), in the guess frame
    q = R^T (p - t), evaluated term by term as the kernel does."""
    lut = log1p_lut() if lut is None else lut
    m = np.asarray(map_xyzw, np.float32)
    G = np.asarray(guess_local, np.float64).reshape(4, 4)
    dx = m[:, 0].astype(np.float64) - G[0, 3]
    dy = m[:, 1].astype(np.float64) - G[1, 3]
    dz = m[:, 2].astype(np.float64) - G[2, 3]
    qx = G[0, 0] * dx + G[1, 0] * dy + G[2, 0] * dz
    qy = G[0, 1] * dx + G[1, 1] * dy + G[2, 1] * dz
    qz = G[0, 2] * dx + G[1, 2] * dy + G[2, 2] * dz
    return _grid_from_index(bev_index(qx, qy, qz), lut)


# ---------------------------------------------------------------- pose correction

def pose_matrix(p) -> np.ndarray:
    """R = Rx(roll) Ry(pitch) Rz(yaw), t = p[0:3]."""
    cx, sx = np.cos(p[3]), np.sin(p[3])
    cy, sy = np.cos(p[4]), np.sin(p[4])
    cz, sz = np.cos(p[5]), np.sin(p[5])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    M = np.eye(4)
    M[:3, :3] = Rx @ Ry @ Rz
    M[:3, 3] = p[:3]
    return M


def correction_params(G, T) -> np.ndarray:
    """(x, y, z, roll, pitch, yaw) of C = G^-1 T."""
    G = np.asarray(G, np.float64).reshape(4, 4)
    T = np.asarray(T, np.float64).reshape(4, 4)
    C = np.linalg.inv(G) @ T
    R = C[:3, :3]
    roll = np.arctan2(-R[1, 2], R[2, 2])
    pitch = np.arcsin(np.clip(R[0, 2], -1.0, 1.0))
    yaw = np.arctan2(-R[0, 1], R[0, 0])
    return np.array([C[0, 3], C[1, 3], C[2, 3], roll, pitch, yaw])


def apply_correction(G, p6) -> np.ndarray:
    return np.asarray(G, np.float64).reshape(4, 4) @ pose_matrix(np.asarray(p6, np.float64))


# ---------------------------------------------------------------- MLP

def _rnd(x, on):
    return bf16_round(np.asarray(x, np.float32)).astype(np.float64) if on else np.asarray(x, np.float64)


def mlp_forward(x, P, bf16=True) -> dict:
    """x: (B, f0) float64 (or bf16 values); P: dict W1 b1 W2 b2 W3 b3 (float).
    With bf16, weight matrices and hidden activations are rounded as the engine
    rounds them; accumulation is exact (float64)."""
    X = np.asarray(x, np.float64)
    W1, W2, W3 = (_rnd(P[k], bf16) for k in ("W1", "W2", "W3"))
    Z1 = X @ W1.T + np.asarray(P["b1"], np.float64)
    H1 = _rnd(np.maximum(Z1, 0.0), bf16)
    Z2 = H1 @ W2.T + np.asarray(P["b2"], np.float64)
    H2 = _rnd(np.maximum(Z2, 0.0), bf16)
    Y = H2 @ W3.T + np.asarray(P["b3"], np.float64)
    return dict(X=X, W1=W1, W2=W2, W3=W3, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Y=Y)


def mlp_loss(Y, T, w=None) -> float:
    w = np.ones(Y.shape[1]) if w is None else np.asarray(w, np.float64)
    return float(np.sum(w * (Y - T) ** 2) / Y.shape[0])


def mlp_backward(x, P, T, w=None, bf16=True) -> tuple:
    """Loss L = (1/B) sum_b sum_o w_o (Y - T)^2 and its gradients.  With bf16, the
    backpropagated signals dZ2 and dZ1 are rounded before the products that use
    them, as the engine stores them in bf16; the bias gradients use the unrounded
    values."""
    f = mlp_forward(x, P, bf16)
    B = f["X"].shape[0]
    T = np.asarray(T, np.float64)
    w = np.ones(f["Y"].shape[1]) if w is None else np.asarray(w, np.float64)
    dY = 2.0 * w * (f["Y"] - T) / B
    g = {"W3": dY.T @ f["H2"], "b3": dY.sum(0)}
    dZ2 = (dY @ f["W3"]) * (f["H2"] > 0)
    g["b2"] = dZ2.sum(0)
    dZ2r = _rnd(dZ2, bf16)
    g["W2"] = dZ2r.T @ f["H1"]
    dZ1 = (dZ2r @ f["W2"]) * (f["H1"] > 0)
    g["b1"] = dZ1.sum(0)
    g["W1"] = _rnd(dZ1, bf16).T @ f["X"]
    return mlp_loss(f["Y"], T, w), g, f


def adam_step(w, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
    """One Adam update at step t (1-based) with decoupled weight decay wd
    (AdamW); returns new (w, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return w - lr * (mh / (np.sqrt(vh) + eps) + wd * w), m, v


# ---------------------------------------------------------------- registration

def register_params_default() -> dict:
    """radius 6 cells (+-3 m), 7 yaws 1 degree apart, yaw refined, height searched,
    roll and pitch leveled"""
    return dict(radius=6, yaws=7, yaw_step=3.14159265358979323846 / 180.0, refine_yaw=1, z_search=1, level=1)


ZH_BINS, ZH_RANGE, ZH_Z0, ZH_BIN = 320, 30, -4.0, 0.05


def z_histogram(z) -> np.ndarray:
    """(320,) int64 counts of heights in 5 cm bins over [-4, 12) m, bin
    floor((z + 4) * 20) (NaN outside)"""
    with np.errstate(invalid="ignore"):
        f = np.floor((np.asarray(z, np.float64) - ZH_Z0) * 20.0)
        ok = (f >= 0) & (f < ZH_BINS)
    return np.bincount(f[ok].astype(np.int64), minlength=ZH_BINS).astype(np.int64)


def z_scores(hs, hm) -> np.ndarray:
    """(61,) score[k + 30] = sum_b scan(b) map(b + k), k = -30 .. 30"""
    out = np.zeros(2 * ZH_RANGE + 1, np.int64)
    for k in range(-ZH_RANGE, ZH_RANGE + 1):
        lo, hi = max(0, -k), ZH_BINS - max(0, k)
        out[k + ZH_RANGE] = int(np.sum(hs[lo:hi] * hm[lo + k:hi + k]))
    return out


def z_shift(sc) -> float:
    """height shift (m) of the best score: highest, then smallest |k|, negative
    first; refined by a parabola through its neighbours; 0 when every score is 0"""
    jb = ZH_RANGE
    for j in range(2 * ZH_RANGE + 1):
        k, kb = j - ZH_RANGE, jb - ZH_RANGE
        if sc[j] > sc[jb] or (sc[j] == sc[jb] and 2 * abs(k) + (k < 0) < 2 * abs(kb) + (kb < 0)):
            jb = j
    if sc[jb] == 0:
        return 0.0
    o = _parabola(float(sc[jb - 1]), float(sc[jb]), float(sc[jb + 1])) if 0 < jb < 2 * ZH_RANGE else 0.0
    return (float(jb - ZH_RANGE) + o) * ZH_BIN


def yaw_table(rp) -> tuple:
    """cosine, sine and tangent of yaw k = (k - (yaws - 1) / 2) * yaw_step (the C
    library's)"""
    import math
    K = rp["yaws"]
    th = [float(k - (K - 1) // 2) * rp["yaw_step"] for k in range(K)]
    return [math.cos(t) for t in th], [math.sin(t) for t in th], [math.tan(t) for t in th]


def occupancy_rows(grid_bits) -> np.ndarray:
    """(512,) uint64: row (z, y) of a grid's occupied cells, bit x = cell x"""
    occ = (np.asarray(grid_bits) != 0).reshape(NZ * NY, NX)
    w = np.zeros(NZ * NY, np.uint64)
    for x in range(NX):
        w |= occ[:, x].astype(np.uint64) << np.uint64(x)
    return w


def map_occupancy_yaws(map_xyzw, guess_local, cs) -> np.ndarray:
    """(K, 512) uint64: occupancy of the map points around the guess in the frame
    turned by yaw k: q = R^T (p - t) as map_grid computes it, then
    q_k = (qx + t_k qy, qy - t_k qx, qz) with t_k = tan(yaw_k) (the rotation by
    -yaw_k scaled by 1 / cos(yaw_k)); cells floor((2 qx + 32) + t_k (2 qy)) and
    floor((2 qy + 32) - t_k (2 qx)), z as bev_index."""
    t_ = cs[2]
    m = np.asarray(map_xyzw, np.float32)
    G = np.asarray(guess_local, np.float64).reshape(4, 4)
    dx = m[:, 0].astype(np.float64) - G[0, 3]
    dy = m[:, 1].astype(np.float64) - G[1, 3]
    dz = m[:, 2].astype(np.float64) - G[2, 3]
    qx = G[0, 0] * dx + G[1, 0] * dy + G[2, 0] * dz
    qy = G[0, 1] * dx + G[1, 1] * dy + G[2, 1] * dz
    qz = G[0, 2] * dx + G[1, 2] * dy + G[2, 2] * dz
    fz = np.floor((qz - Z0) / DZ)
    ax, bx, ay, by = 2.0 * qx + 32.0, 2.0 * qy, 2.0 * qy + 32.0, 2.0 * qx
    out = np.zeros((len(t_), NZ * NY), np.uint64)
    for k in range(len(t_)):
        idx = bev_index_xyz_floor(np.floor(ax + t_[k] * bx), np.floor(ay - t_[k] * by), fz)
        occ = np.zeros(CELLS, bool)
        occ[idx[idx >= 0]] = True
        out[k] = occupancy_rows(occ)
    return out


def bev_index_xyz_floor(fx, fy, fz) -> np.ndarray:
    with np.errstate(invalid="ignore"):
        ok = (fx >= 0) & (fx < NX) & (fy >= 0) & (fy < NY) & (fz >= 0) & (fz < NZ)
    idx = np.full(fx.shape, -1, np.int64)
    idx[ok] = (fz[ok].astype(np.int64) * NY + fy[ok].astype(np.int64)) * NX + fx[ok].astype(np.int64)
    return idx


def register_scores(srows, mrows, R) -> np.ndarray:
    """(K, 2R+1, 2R+1) int64: score[k, dy+R, dx+R] = sum over cells c of
    scan(c) AND map_k(c + (dx, dy)), cells outside the grid empty."""
    S = np.asarray(srows, np.uint64).reshape(NZ, NY)
    W = 2 * R + 1
    out = np.zeros((mrows.shape[0], W, W), np.int64)
    for k in range(mrows.shape[0]):
        M = mrows[k].reshape(NZ, NY)
        for dy in range(-R, R + 1):
            Mr = np.zeros_like(M)
            if dy >= 0:
                Mr[:, :NY - dy] = M[:, dy:]
            else:
                Mr[:, -dy:] = M[:, :NY + dy]
            for dx in range(-R, R + 1):
                Ms = (Mr >> np.uint64(dx)) if dx >= 0 else (Mr << np.uint64(-dx))
                out[k, dy + R, dx + R] = int(np.bitwise_count(S & Ms).sum())
    return out


def _parabola(a, b, c) -> float:
    d = a - 2.0 * b + c
    return 0.5 * (a - c) / d if d < 0.0 else 0.0


def _small_cos_sin(d) -> tuple:
    d2 = d * d
    c = 1.0 - d2 / 2.0 + d2 * d2 / 24.0 - d2 * d2 * d2 / 720.0 + d2 * d2 * d2 * d2 / 40320.0
    s = d - d * d2 / 6.0 + d * d2 * d2 / 120.0 - d * d2 * d2 * d2 / 5040.0
    return c, s


def register_pick(sc, rp) -> tuple:
    """(score, k, dy, dx) of the best candidate: highest score, then the lowest
    rank (|k - kc|, |dy|, |dx|, each negative before positive)."""
    K, R = sc.shape[0], rp["radius"]
    kc = (K - 1) // 2
    best = None
    for k in range(K):
        for dy in range(-R, R + 1):
            for dx in range(-R, R + 1):
                rank = ((((abs(k - kc) * 2 + (k < kc)) * 16 + abs(dy)) * 2 + (dy < 0)) * 16 + abs(dx)) * 2 + (dx < 0)
                key = (int(sc[k, dy + R, dx + R]), -rank)
                if best is None or key > best[0]:
                    best = (key, k, dy, dx)
    return best[0][0], best[1], best[2], best[3]


def map_heights(map_xyzw, guess_local) -> np.ndarray:
    """height histogram of the map points under the grid's footprint
    (|qx| < 16, |qy| < 16 in the guess frame, q as map_grid computes it)"""
    m = np.asarray(map_xyzw, np.float32)
    G = np.asarray(guess_local, np.float64).reshape(4, 4)
    dx = m[:, 0].astype(np.float64) - G[0, 3]
    dy = m[:, 1].astype(np.float64) - G[1, 3]
    dz = m[:, 2].astype(np.float64) - G[2, 3]
    qx = G[0, 0] * dx + G[1, 0] * dy + G[2, 0] * dz
    qy = G[0, 1] * dx + G[1, 1] * dy + G[2, 1] * dz
    qz = G[0, 2] * dx + G[1, 2] * dy + G[2, 2] * dz
    with np.errstate(invalid="ignore"):
        foot = (np.abs(qx) < -X0) & (np.abs(qy) < -Y0)
    return z_histogram(qz[foot])


def register(scan_xyzw, map_xyzw, guess_local, guess_world, rp, lut=None) -> tuple:
    """Registered (map-local guess, world guess, best score) of one (scan, guess)
    pair: G' = G Rz(yaw_k) T(d) [Rz(delta)], d the best shift refined by parabolas
    through its neighbours' scores (half a cell at most), delta likewise between
    the neighbouring yaws when refine_yaw; unchanged when the best score is 0.
    With z_search the result then moves along its z axis by the best height shift
    (z_scores of the scan's heights and the map's under the footprint of G); with
    level it is then leveled (`level`) against the map around it."""
    GL = np.asarray(guess_local, np.float64).reshape(4, 4).copy()
    GW = np.asarray(guess_world, np.float64).reshape(4, 4).copy()
    if rp["radius"] == 0:
        return GL, GW, 0
    dz = 0.0
    if rp.get("z_search", 0):
        dz = z_shift(z_scores(z_histogram(np.asarray(scan_xyzw, np.float32)[:, 2].astype(np.float64)),
                              map_heights(map_xyzw, GL)))

    def lift(Gs):
        out = []
        for G in Gs:
            g = G.reshape(-1).copy()
            if dz != 0.0:
                for r in range(3):
                    g[4 * r + 3] = g[4 * r + 3] + g[4 * r + 2] * dz
            out.append(g.reshape(4, 4))
        return out
    cs = yaw_table(rp)
    srows = occupancy_rows(scan_grid(scan_xyzw, lut))
    sc = register_scores(srows, map_occupancy_yaws(map_xyzw, GL, cs), rp["radius"])
    v, k, dy, dx = register_pick(sc, rp)
    if v == 0:
        gl, gw = lift([GL, GW])
        if rp.get("level", 0):
            gl, gw = level(scan_xyzw, map_xyzw, gl, gw)[:2]
        return gl, gw, 0
    R, K = rp["radius"], len(cs[0])
    S = sc[k].astype(np.float64)
    b = S[dy + R, dx + R]
    ox = _parabola(S[dy + R, dx + R - 1], b, S[dy + R, dx + R + 1]) if abs(dx) < R else 0.0
    oy = _parabola(S[dy + R - 1, dx + R], b, S[dy + R + 1, dx + R]) if abs(dy) < R else 0.0
    ck, sk = cs[0][k], cs[1][k]
    c, s = ck, sk
    if rp["refine_yaw"] and 0 < k < K - 1:
        o = _parabola(float(sc[k - 1, dy + R, dx + R]), b, float(sc[k + 1, dy + R, dx + R]))
        cd, sd = _small_cos_sin(o * rp["yaw_step"])
        c = ck * cd - sk * sd
        s = sk * cd + ck * sd
    d0, d1 = (float(dx) + ox) * 0.5, (float(dy) + oy) * 0.5
    out = []
    for G in (GL, GW):
        g = [float(v_) for v_ in G.reshape(-1)]
        for r in range(3):
            a, bb = g[4 * r], g[4 * r + 1]
            ak, bk = a * ck + bb * sk, bb * ck - a * sk
            g[4 * r + 3] = g[4 * r + 3] + ak * d0 + bk * d1
            g[4 * r] = a * c + bb * s
            g[4 * r + 1] = bb * c - a * s
        out.append(np.array(g).reshape(4, 4))
    out = lift(out)
    if rp.get("level", 0):
        out = level(scan_xyzw, map_xyzw, out[0], out[1])[:2]
    return out[0], out[1], int(v)


# ---------------------------------------------------------------- leveling
LEV_N, LEV_R, LEV_Z0, LEV_Z1 = 32, 16.0, -8.0, 8.0
LEV_TRIM, LEV_MIN_CELLS, LEV_MAX_ANGLE = 0.25, 32, 0.035


def level_cells(x, y, z) -> np.ndarray:
    """(1024,) float64 lowest height per 1 m cell of [-16, 16)^2 (cell (j, i) =
    (floor(y + 16), floor(x + 16)), index j * 32 + i), heights in [-8, 8) only;
    +inf where the cell holds no point"""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    with np.errstate(invalid="ignore"):
        fi, fj = np.floor(x + LEV_R), np.floor(y + LEV_R)
        ok = (fi >= 0) & (fi < LEV_N) & (fj >= 0) & (fj < LEV_N) & (z >= LEV_Z0) & (z < LEV_Z1)
    out = np.full(LEV_N * LEV_N, np.inf)
    np.minimum.at(out, fj[ok].astype(np.int64) * LEV_N + fi[ok].astype(np.int64), z[ok])
    return out


def _wave_tree(v) -> float:
    """sum of 64 values as a 64-lane butterfly leaves it in lane 0"""
    a = np.asarray(v, np.float64)
    while a.shape[0] > 1:
        h = a.shape[0] // 2
        a = a[:h] + a[h:]
    return float(a[0])


def level_sums(d, keep) -> list:
    """the nine sums n, x, y, xx, xy, yy, d, xd, yd over the kept cells (cell
    centres x = i - 15.5, y = j - 15.5): per 64-cell block a butterfly tree, the 16
    block sums then added in block order"""
    i = np.arange(LEV_N * LEV_N)
    x = (i % LEV_N).astype(np.float64) - 15.5
    y = (i // LEV_N).astype(np.float64) - 15.5
    k = np.asarray(keep, bool)
    dd = np.where(k, np.asarray(d, np.float64), 0.0)
    one = k.astype(np.float64)
    terms = [one, one * x, one * y, one * x * x, one * x * y, one * y * y, dd, dd * x, dd * y]
    out = []
    for t in terms:
        s = 0.0
        for w in range(LEV_N * LEV_N // 64):
            s = s + _wave_tree(t[64 * w:64 * w + 64]) if w else _wave_tree(t[:64])
        out.append(s)
    return out


def level_solve(S) -> tuple:
    """slopes (a, b) of the least-squares plane d = a x + b y + c from the sums, by
    Cramer's rule; None when the system is singular"""
    n, sx, sy, sxx, sxy, syy, sd, sxd, syd = S
    m0 = syy * n - sy * sy
    m1 = sxy * n - sy * sx
    m2 = sxy * sy - syy * sx
    det = sxx * m0 - sxy * m1 + sx * m2
    if not det > 0.0:
        return None
    a = (sxd * m0 - sxy * (syd * n - sy * sd) + sx * (syd * sy - syy * sd)) / det
    b = (sxx * (syd * n - sy * sd) - sxd * m1 + sx * (sxy * sd - syd * sx)) / det
    return a, b


def level(scan_xyzw, map_xyzw, guess_local, guess_world) -> tuple:
    """Leveled (map-local guess, world guess, kept cells, (pitch, roll)) of a
    registered guess.  Per 1 m cell of [-16, 16)^2 the lowest scan point (sensor
    frame) and the lowest map point in the guess frame (q = R^T (p - t), heights in
    [-8, 8)); d = map - scan where both exist.  With the cells within 0.25 m of the
    lower median of d (at least 32 cells, and 32 kept), the plane d = a x + b y + c
    gives the tilt of the guess against the scan: pitch a, roll -b (a rotation of
    the guess by Rx(roll) Ry(pitch) lifts the map in its frame by about
    pitch x - roll y).  The guess turns by (Rx(roll) Ry(pitch))^-1, each angle's
    cosine and sine from the series the registration uses, unless an angle exceeds
    0.035 rad; otherwise it is returned unchanged (pitch and roll then None)."""
    GL = np.asarray(guess_local, np.float64).reshape(4, 4).copy()
    GW = np.asarray(guess_world, np.float64).reshape(4, 4).copy()
    s = np.asarray(scan_xyzw, np.float32)
    zs = level_cells(s[:, 0], s[:, 1], s[:, 2])
    m = np.asarray(map_xyzw, np.float32)
    dx = m[:, 0].astype(np.float64) - GL[0, 3]
    dy = m[:, 1].astype(np.float64) - GL[1, 3]
    dz = m[:, 2].astype(np.float64) - GL[2, 3]
    qx = GL[0, 0] * dx + GL[1, 0] * dy + GL[2, 0] * dz
    qy = GL[0, 1] * dx + GL[1, 1] * dy + GL[2, 1] * dz
    qz = GL[0, 2] * dx + GL[1, 2] * dy + GL[2, 2] * dz
    zm = level_cells(qx, qy, qz)
    both = np.isfinite(zs) & np.isfinite(zm)
    n = int(both.sum())
    if n < LEV_MIN_CELLS:
        return GL, GW, 0, None
    with np.errstate(invalid="ignore"):
        d = np.where(both, zm - zs, np.inf)
    med = np.sort(d)[(n - 1) // 2]
    keep = both & (np.abs(d - med) < LEV_TRIM)
    if int(keep.sum()) < LEV_MIN_CELLS:
        return GL, GW, int(keep.sum()), None
    ab = level_solve(level_sums(np.where(keep, d, 0.0), keep))
    if ab is None:
        return GL, GW, int(keep.sum()), None
    pitch, roll = ab[0], -ab[1]
    if not (abs(pitch) <= LEV_MAX_ANGLE and abs(roll) <= LEV_MAX_ANGLE):
        return GL, GW, int(keep.sum()), None
    ct, st = _small_cos_sin(pitch)
    cf, sf = _small_cos_sin(roll)
    # M = Ry(-pitch) Rx(-roll)
    M = [[ct, st * sf, -(st * cf)], [0.0, cf, sf], [st, -(ct * sf), ct * cf]]
    out = []
    for G in (GL, GW):
        g = [float(v) for v in G.reshape(-1)]
        for r in range(3):
            a0, a1, a2 = g[4 * r], g[4 * r + 1], g[4 * r + 2]
            for k in range(3):
                g[4 * r + k] = a0 * M[0][k] + a1 * M[1][k] + a2 * M[2][k]
        out.append(np.array(g).reshape(4, 4))
    return out[0], out[1], int(keep.sum()), (pitch, roll)
