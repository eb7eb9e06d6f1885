"""GPU parity: the HIP engine through its C ABI against the CPU oracle on the same
seeded inputs.  Tolerances are stated per test; integer / index results (voxel
keys, counts, pair counts, iteration counts, flags) must match exactly."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg0(A, orc):
    scene = A.scene_cfg(1)
    mp = A.make_map(scene, 100000)
    return scene, mp, orc.NdtGrid(mp)


@pytest.fixture(scope="module")
def gpu_grid(A, ctx, cfg0):
    return A.Grid(ctx, cfg0[1], A.params(A.NDT_MODE_NDT))


def test_ndt_grid_matches_oracle(gpu_grid, cfg0):
    gv = gpu_grid.voxels()
    ov = cfg0[2].voxels()
    assert np.array_equal(gv["key"], ov["key"]) and np.array_equal(gv["n"], ov["n"])
    assert np.array_equal(gv["valid"], ov["valid"])
    # same float64 operations in the same order on both sides
    assert np.array_equal(gv["mean"], ov["mean"])
    assert np.allclose(gv["icov"], ov["icov"], rtol=1e-12, atol=0)
    info = gpu_grid.info()
    assert info["valid_voxels"] == int(ov["valid"].sum())


@pytest.mark.parametrize("point,summ,tol", [(0, 0, 1e-4), (0, 1, 1e-4), (0, 2, 1e-4), (1, 2, 1e-10)])
def test_derivative_pass(A, orc, gpu_grid, cfg0, point, summ, tol):
    """One pass: score, gradient and Hessian relative to their largest entry; fp32
    per-point arithmetic within 1e-4, fp64 within 1e-10.  Pair counts exact."""
    scene, mp, og = cfg0
    for s in range(3):
        scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(2000), s)
        p = np.r_[guess[:3, 3], orc.ndt_euler(guess)] + np.array([0.02, -0.01, 0.0, 0.01, -0.02, 0.004])
        so, go, Ho, npo = og.derivatives(scan, p)
        prm = A.params(A.NDT_MODE_NDT, point_precision=point, sum_precision=summ)
        sg, gg, Hg, npg = gpu_grid.derivatives(scan, p, prm=prm)
        assert npg == npo
        assert abs(sg - so) <= tol * abs(so)
        assert np.abs(gg - go).max() <= tol * np.abs(go).max()
        assert np.abs(Hg - Ho).max() <= tol * np.abs(Ho).max()


def _compare_runs(res, ref, dt_tol, dr_tol, orc):
    """Scans whose GPU run differs from the oracle's: pose beyond the tolerances, or
    a different iteration count, derivative-pass count (a line-search divergence
    that lands on the same pose) or converged flag."""
    bad = []
    for i, (r, o) in enumerate(zip(res, ref)):
        dt, dr = orc.pose_error(r["final_f64"], o["final_f64"])
        if not (dt <= dt_tol and dr <= dr_tol and r["iterations"] == o["iterations"] and r["passes"] == o["passes"]
                and r["converged"] == o["converged"] and r["status"] == 0):
            bad.append((i, dt, dr, r["iterations"], o["iterations"], r["passes"], o["passes"], r["converged"],
                        o["converged"]))
    return bad


def test_align_fp64_points_matches_oracle(A, orc, gpu_grid, cfg0):
    """fp64 per-point arithmetic: same final transform within 1e-3 m / 1e-4 rad, same
    iteration count and converged flag as the float64 oracle on every scan."""
    scene, mp, og = cfg0
    scans = [A.make_scan(scene, mp, A.scan_cfg(2000), s) for s in range(12)]
    prm = A.params(A.NDT_MODE_NDT, point_precision=A.NDT_POINT_FP64)
    res = [gpu_grid.align(s[0], s[2], prm=prm) for s in scans]
    ref = og.align_many([s[0] for s in scans], np.stack([s[2] for s in scans]), threads=8)
    assert _compare_runs(res, ref, 1e-3, 1e-4, orc) == []


def test_hessian_at_accept_matches_oracle(A, orc, gpu_grid, cfg0):
    """Line-search loop passes without the Hessian and one more pass at the accepted
    step (the reference's order): the same iterations, passes and flags as the
    oracle run the same way, and the same poses as the default mode (the Hessian
    at the accepted step is the same sum either way), with exactly the extra
    passes the oracle makes."""
    scene, mp, og = cfg0
    scans = [A.make_scan(scene, mp, A.scan_cfg(2000, 1.0, 3.0), 700 + s) for s in range(10)]
    pts, gs = [s[0] for s in scans], np.stack([s[2] for s in scans])
    res = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT, hessian_at_accept=1))
    full = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT))
    oa = orc.NdtGrid(mp, orc.ndt_params(hessian_at_accept=1))
    ref = oa.align_many(pts, gs, threads=8)
    ref_full = og.align_many(pts, gs, threads=8)
    assert _compare_runs(res, ref, 1e-3, 1e-4, orc) == []
    extra = 0
    for r, f, o, of in zip(res, full, ref, ref_full):
        assert np.array_equal(r["final_f64"], f["final_f64"])
        assert r["passes"] - f["passes"] == o["passes"] - of["passes"]
        extra += r["passes"] - f["passes"]
    assert extra > 0


def test_align_fp32_precision_study(A, orc, gpu_grid, cfg0):
    """float32 per-point arithmetic (fp64 sums) is not a parity mode: per-pair float32
    rounding (relative 1e-6, systematic per voxel) changes the line-search decisions
    of some of these trajectories, which the float64 oracle itself shows to be
    sensitive (a 1e-9 m shift of the guess moves some results by centimetres).
    Characterised here: at least half of the scans land within 1 cm / 0.05 deg of the
    oracle, and every float32 run converges."""
    scene, mp, og = cfg0
    scans = [A.make_scan(scene, mp, A.scan_cfg(2000), s) for s in range(12)]
    ref = og.align_many([s[0] for s in scans], np.stack([s[2] for s in scans]), threads=8)
    prm = A.params(A.NDT_MODE_NDT, point_precision=A.NDT_POINT_FP32, sum_precision=A.NDT_SUM_FP64)
    res = [gpu_grid.align(s[0], s[2], prm=prm) for s in scans]
    close = 0
    for r, o in zip(res, ref):
        dt, dr = orc.pose_error(r["final_f64"], o["final_f64"])
        close += dt <= 0.01 and dr <= np.deg2rad(0.05)
        assert r["status"] == 0
    assert close >= 6
    assert sum(r["converged"] for r in res) >= sum(o["converged"] for o in ref) - 1


def test_batch_matches_single(A, orc, gpu_grid, cfg0):
    scene, mp, og = cfg0
    sizes = [500, 3000, 1200, 2500, 800, 4000, 1500, 2000]
    scans = [A.make_scan(scene, mp, A.scan_cfg(n), 40 + k) for k, n in enumerate(sizes)]
    prm = A.params(A.NDT_MODE_NDT, point_precision=A.NDT_POINT_FP64)
    rb = gpu_grid.align_batch([s[0] for s in scans], np.stack([s[2] for s in scans]), prm=prm)
    ref = og.align_many([s[0] for s in scans], np.stack([s[2] for s in scans]), threads=8)
    assert _compare_runs(rb, ref, 1e-3, 1e-4, orc) == []


def test_schedulers_and_batch_composition_agree(A, gpu_grid, cfg0):
    """Work queue and rounds reduce the same partial sums in the same order, so their
    results are bit-identical (at two chunk sizes), and a scan's result does not
    depend on the other scans of its batch."""
    scene, mp, _ = cfg0
    sizes = [700, 5000, 1300, 2600, 900, 4100, 1600, 2100, 3000]
    scans = [A.make_scan(scene, mp, A.scan_cfg(n), 60 + k) for k, n in enumerate(sizes)]
    pts = [s[0] for s in scans]
    gs = np.stack([s[2] for s in scans])
    runs = []
    for sched in (A.NDT_SCHED_QUEUE, A.NDT_SCHED_ROUNDS):
        for ppw in (0, 256):
            prm = A.params(A.NDT_MODE_NDT, scheduler=sched, points_per_wg=ppw)
            runs.append((sched, ppw, gpu_grid.align_batch(pts, gs, prm=prm)))
    alone = gpu_grid.align_batch(pts[3:5], gs[3:5], prm=A.params(A.NDT_MODE_NDT))
    for (sa, pa, ra), (sb, pb, rb) in [(runs[0], runs[2]), (runs[1], runs[3])]:
        for x, y in zip(ra, rb):
            assert np.array_equal(x["final_f64"], y["final_f64"]) and x["passes"] == y["passes"], (sa, sb, pa)
    for x, y in zip(alone, runs[0][2][3:5]):
        assert np.array_equal(x["final_f64"], y["final_f64"]) and x["passes"] == y["passes"]


@pytest.mark.parametrize("fine_pass", [0, 12, 6])
def test_late_pass_chunking_agrees(A, gpu_grid, cfg0, fine_pass):
    """From pass fine_pass of a run on (0: the 24th) a scan's chunks are a quarter
    the size.  Some scans run past it (guesses +-4 m / +-15 deg off for the default,
    +-3 m / +-10 deg otherwise): queue and rounds still give bit-identical results
    (both derive the chunking from the pass number), a scan's result does not
    depend on its batch, 0 is 24, and another fine_pass moves results only at the
    level of rounding (on starts whose results are stable to that level)."""
    scene, mp, _ = cfg0
    sizes = [2600, 4100, 1800, 3300, 2200, 5200, 1500, 3900]
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 300 + k) for k, n in enumerate(sizes)]
    pts = [m[0] for m in made]
    if fine_pass:
        gs = A.perturb_poses(np.stack([m[1] for m in made]), np.random.default_rng(5), xy=3.0, yaw_deg=10.0)
    else:
        gs = A.perturb_poses(np.stack([m[1] for m in made]), np.random.default_rng(19), xy=4.0, yaw_deg=15.0)
    for ppw in (0, 1024):
        kw = dict(points_per_wg=ppw, fine_pass=fine_pass)
        q = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT, scheduler=A.NDT_SCHED_QUEUE, **kw))
        r = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT, scheduler=A.NDT_SCHED_ROUNDS, **kw))
        assert max(x["passes"] for x in q) > (fine_pass or 24), [x["passes"] for x in q]
        for x, y in zip(q, r):
            assert np.array_equal(x["final_f64"], y["final_f64"]) and x["passes"] == y["passes"], ppw
        k = int(np.argmax([x["passes"] for x in q]))
        alone = gpu_grid.align_batch(pts[k:k + 1], gs[k:k + 1],
                                     prm=A.params(A.NDT_MODE_NDT, scheduler=A.NDT_SCHED_QUEUE, **kw))
        assert np.array_equal(alone[0]["final_f64"], q[k]["final_f64"]) and alone[0]["passes"] == q[k]["passes"]
        if fine_pass:
            d = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT, points_per_wg=ppw))
            assert max(np.abs(x["final_f64"][:3, 3] - y["final_f64"][:3, 3]).max() for x, y in zip(q, d)) < 1e-3
        else:
            e = gpu_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT, points_per_wg=ppw, fine_pass=24))
            for x, y in zip(q, e):
                assert np.array_equal(x["final_f64"], y["final_f64"]) and x["passes"] == y["passes"], ppw


def test_deterministic(A, gpu_grid, cfg0):
    scene, mp, _ = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(2000), 3)
    a = gpu_grid.align(scan, guess)
    b = gpu_grid.align(scan, guess)
    assert np.array_equal(a["final_f64"], b["final_f64"]) and a["passes"] == b["passes"]


def test_scan_outside_map(A, gpu_grid, orc, cfg0):
    """No neighbours anywhere: zero gradient and Hessian, zero Newton step, converged
    at the guess after no iterations (as the oracle)."""
    scene, mp, og = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(500), 0)
    far = guess.copy()
    far[2, 3] += 500.0
    r = gpu_grid.align(scan, far)
    o = og.align(scan, far)
    assert r["iterations"] == o["iterations"] == 0 and r["converged"] == o["converged"]
    assert np.allclose(r["final_f64"], far, atol=1e-9)


def test_invalid_inputs(A, ctx, gpu_grid):
    with pytest.raises(A.NdtError) as e:
        gpu_grid.align(np.zeros((0, 4), np.float32), np.eye(4))
    assert e.value.code == 1
    with pytest.raises(A.NdtError):
        A.Grid(ctx, np.full((10, 4), np.nan, np.float32), A.params(A.NDT_MODE_NDT))


def test_non_finite_guess_rejected(A, ctx, gpu_grid, cfg0):
    """A guess or pose with a NaN or infinite entry is a bad argument for every entry
    point that takes one; the context stays usable."""
    scene, mp, _ = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(800), 11)
    sc = A.Scans(ctx, [scan, scan])
    bm = A.BevMap(ctx, mp)
    for bad in (np.nan, np.inf, -np.inf):
        g = np.stack([guess, guess])
        g[1, 0, 3 if bad != bad else 1] = bad
        calls = [lambda: gpu_grid.align(scan, g[1]), lambda: sc.align(gpu_grid, g),
                 lambda: gpu_grid.align_batch([scan, scan], g), lambda: sc.fitness(gpu_grid, g, 0.0),
                 lambda: bm.features(sc, np.array([0, 1], np.int32), g)]
        for k, call in enumerate(calls):
            with pytest.raises(A.NdtError) as e:
                call()
            assert e.value.code == 1 and "non-finite" in str(e.value), (bad, k)
    r = gpu_grid.align(scan, guess)
    assert r["status"] == 0 and r["iterations"] > 0


def test_degenerate_scans_match_oracle(A, orc, gpu_grid, cfg0):
    """Scans whose Hessian is singular or rank-deficient (one point; points on a
    line; points in one cell) take the same path as the oracle: same iteration count
    and convergence, same pose."""
    scene, mp, og = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(2000), 0)
    line = np.zeros((50, 4), np.float32)
    line[:, 0] = np.linspace(0, 5, 50)
    line[:, 3] = 1
    blob = np.tile(scan[:1], (64, 1))
    for name, s in (("one", scan[:1]), ("two", scan[:2]), ("line", line), ("blob", blob)):
        r = gpu_grid.align(s, guess)
        o = og.align(s, guess)
        dt, dr = orc.pose_error(r["final_f64"], o["final_f64"])
        assert r["iterations"] == o["iterations"] and r["converged"] == o["converged"], name
        assert dt <= 1e-6 and dr <= 1e-7 and np.isfinite(r["final_f64"]).all(), (name, dt, dr)
        assert abs(r["score"] - o["score"]) <= 1e-9 * max(abs(o["score"]), 1.0), name


def test_map_without_valid_cells(A, orc, ctx, cfg0):
    """A map too sparse for any cell to hold the minimum point count builds a grid
    with no valid cell; alignments then find no correspondences and stop at the guess
    (as the oracle)."""
    scene, mp, _ = cfg0
    sparse = mp[:: len(mp) // 3][:3].copy()
    g = A.Grid(ctx, sparse, A.params(A.NDT_MODE_NDT))
    info = g.info()
    assert info["voxels"] >= 1 and info["valid_voxels"] == 0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(600), 3)
    r = g.align(scan, guess)
    o = orc.NdtGrid(sparse).align(scan, guess)
    assert r["iterations"] == o["iterations"] == 0 and r["converged"] == o["converged"]
    assert r["score"] == o["score"] == 0.0 and np.allclose(r["final_f64"], guess, atol=1e-9)


def test_handles_outlive_their_closed_context(A, ctx, gpu_grid, cfg0):
    """Destroying a context before the grids, scan batches and BEV maps made on it is
    safe: each handle keeps the context until it is destroyed itself, and nothing is
    left behind for the next call on another context to report."""
    scene, mp, _ = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(900), 12)
    want = gpu_grid.align(scan, guess)
    c2 = A.Context(0)
    g2 = A.Grid(c2, mp, A.params(A.NDT_MODE_NDT))
    s2 = A.Scans(c2, [scan])
    b2 = A.BevMap(c2, mp)
    got = s2.align(g2, guess[None])
    c2.close()
    g2.close()
    s2.close()
    b2.close()
    r = gpu_grid.align(scan, guess)
    assert r["status"] == 0 and np.array_equal(r["final_f64"], want["final_f64"])
    assert np.array_equal(got[0]["final_f64"], want["final_f64"])


def test_device_wait_timeout_is_reported(A, ctx, gpu_grid, cfg0, monkeypatch):
    """NDT_WAIT_LIMIT_US=1 bounds every device-side wait (grid barrier, work-queue
    claim) to 1 us.  Each scan is one 65536-point chunk, so idle workgroups wait far
    longer than that: the wait expires, the launch drains, and the call returns
    NDT_ERR_TIMEOUT (4) with both schedulers and the captured submission.  With the
    limit lifted the same context and batch give the normal results again."""
    scene, mp, _ = cfg0
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 140 + k) for k, n in enumerate([6000, 3000])]
    scans = [m[0] for m in made]
    guesses = np.stack([m[2] for m in made])
    prm = {s: A.params(A.NDT_MODE_NDT, scheduler=s, points_per_wg=65536)
           for s in (A.NDT_SCHED_QUEUE, A.NDT_SCHED_ROUNDS)}
    want = gpu_grid.align_batch(scans, guesses, prm=prm[A.NDT_SCHED_QUEUE])
    assert all(w["iterations"] >= 2 for w in want)
    sc = A.Scans(ctx, scans)
    try:
        monkeypatch.setenv("NDT_WAIT_LIMIT_US", "1")
        for sched, p in prm.items():
            with pytest.raises(A.NdtError) as e:
                gpu_grid.align_batch(scans, guesses, prm=p)
            assert e.value.code == 4, sched
        with pytest.raises(A.NdtError) as e:
            sc.align(gpu_grid, guesses, prm[A.NDT_SCHED_QUEUE])
        assert e.value.code == 4
    finally:
        monkeypatch.delenv("NDT_WAIT_LIMIT_US", raising=False)
    runs = [sc.align(gpu_grid, guesses, prm[A.NDT_SCHED_QUEUE])]
    runs += [gpu_grid.align_batch(scans, guesses, prm=p) for p in prm.values()]
    for res in runs:
        for r, w in zip(res, want):
            assert r["status"] == 0 and np.array_equal(r["final_f64"], w["final_f64"])


def test_injected_allocation_failure_recovers(A, ctx, gpu_grid, cfg0, monkeypatch):
    """With NDT_DEVICE_MEMORY_LIMIT below what a call needs, grid building, scan
    upload and alignment fail with NDT_ERR_NOMEM (3) and leave nothing half-built;
    once the limit is lifted the same context gives the normal results."""
    scene, mp, _ = cfg0
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 40 + k) for k, n in enumerate([1200, 700])]
    scans = [m[0] for m in made]
    guesses = np.stack([m[2] for m in made])
    want = A.Scans(ctx, scans).align(gpu_grid, guesses)
    fresh = A.Context(0)
    try:
        monkeypatch.setenv("NDT_DEVICE_MEMORY_LIMIT", "0")
        with pytest.raises(A.NdtError) as e:
            A.Grid(fresh, mp, A.params(A.NDT_MODE_NDT))
        assert e.value.code == 3 and "NDT_DEVICE_MEMORY_LIMIT" in str(e.value)
        with pytest.raises(A.NdtError) as e:
            A.Scans(fresh, scans)
        assert e.value.code == 3
        monkeypatch.delenv("NDT_DEVICE_MEMORY_LIMIT")
        grid = A.Grid(fresh, mp, A.params(A.NDT_MODE_NDT))
        sc = A.Scans(fresh, scans)
        monkeypatch.setenv("NDT_DEVICE_MEMORY_LIMIT", "0")
        with pytest.raises(A.NdtError) as e:
            sc.align(grid, guesses)
        assert e.value.code == 3
        monkeypatch.delenv("NDT_DEVICE_MEMORY_LIMIT")
        got = sc.align(grid, guesses)
        for r, w in zip(got, want):
            assert np.array_equal(r["final_f64"], w["final_f64"]) and r["iterations"] == w["iterations"]
    finally:
        monkeypatch.delenv("NDT_DEVICE_MEMORY_LIMIT", raising=False)
        fresh.close()


def test_nan_points_are_ignored(A, orc, gpu_grid, cfg0):
    scene, mp, og = cfg0
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(1500), 5)
    dirty = scan.copy()
    dirty[::97, 1] = np.nan
    prm = A.params(A.NDT_MODE_NDT, point_precision=A.NDT_POINT_FP64)
    r = gpu_grid.align(dirty, guess, prm=prm)
    o = og.align(dirty, guess)
    dt, dr = orc.pose_error(r["final_f64"], o["final_f64"])
    assert dt <= 1e-3 and dr <= 1e-4 and r["iterations"] == o["iterations"]


def test_utm_offset_origin(A, orc, ctx):
    """A map given relative to a UTM-sized origin: poses in world coordinates agree
    with the oracle run in the local frame, shifted by the origin."""
    origin = (4.2e5, 5.6e6, 80.0)
    scene = A.scene_cfg(9, extent=60.0, origin=origin)
    mp = A.make_map(scene, 50000)
    og = orc.NdtGrid(mp)
    g = A.Grid(ctx, mp, A.params(A.NDT_MODE_NDT, origin=origin))
    for s in range(4):
        scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(2000), s)
        r = g.align(scan, guess, prm=A.params(A.NDT_MODE_NDT, origin=origin, point_precision=A.NDT_POINT_FP64))
        local = guess.copy()
        local[:3, 3] -= origin
        o = og.align(scan, local)
        world = o["final_f64"].copy()
        world[:3, 3] += origin
        dt, dr = orc.pose_error(r["final_f64"], world)
        assert dt <= 1e-3 and dr <= 1e-4 and r["iterations"] == o["iterations"]


# ------------------------------------------------------------------ compat mode

def test_compat_grid_matches_oracle(A, orc, ctx, compat_cases):
    for c in compat_cases[::3]:
        g = A.Grid(ctx, c["map"], A.params(A.NDT_MODE_COMPAT))
        v = g.voxels()
        info = g.info()
        assert list(info["dims"]) == list(c["dims"]) and info["dense_cells"] == int(c["ncells"][0])
        assert np.array_equal(np.array(info["min_corner"], np.float32), c["min"][:3])
        assert np.array_equal(v["key"][:, 0], c["cell_idx"]) and np.array_equal(v["n"], c["cell_n"])
        # bit-exact with the reference's float64 sums (same order, no FMA)
        assert np.array_equal(v["mean"], c["cell_mean"]) and np.array_equal(v["icov"], c["cell_icov"])


def test_compat_first_pass(A, orc, ctx, compat_cases):
    """The reference's first derivative pass (p0, score, gradient, Hessian): the GPU
    sums in a tree instead of the reference's serial order and uses the device
    exp, so it agrees to rounding (relative 1e-9), with identical pair counts."""
    for c in compat_cases:
        g = A.Grid(ctx, c["map"], A.params(A.NDT_MODE_COMPAT))
        s, gr, H, npairs = g.derivatives(c["scan"], c["p0"], compat_transform=c["guess"])
        assert abs(s - c["score0"][0]) <= 1e-9 * abs(c["score0"][0])
        assert np.abs(gr - c["grad0"]).max() <= 1e-9 * np.abs(c["grad0"]).max()
        assert np.abs(H - c["hess0"]).max() <= 1e-9 * np.abs(c["hess0"]).max()


def test_compat_callback_matches_reference(A, orc, ctx, compat_cases):
    """partial_points_callback on the GPU against the reference's own outputs: the
    reference's acceptance check holds on every case, iterations (2) and the flag
    are equal, the translation column is the guess's, and the rotation agrees
    element-wise within 1e-4 (the result depends on the last bits of the sums)."""
    exact = 0
    for c in compat_cases:
        r = ctx.compat_callback(c["scan"], c["guess"], c["map"])
        ref = A.CallbackResult(bool(c["converged"][0]), c["final"])
        ok, _ = A.check_outputs([A.CallbackResult(r["converged"], r["final_f32"])], [ref])
        assert ok and r["iterations"] == 2 and r["converged"]
        assert np.array_equal(r["final_f32"][:, 3], c["final"][:, 3])
        assert np.abs(r["final_f32"] - c["final"]).max() <= 1e-4
        exact += np.array_equal(r["final_f32"], c["final"])
    print(f"compat bit-exact with the reference on {exact}/{len(compat_cases)} cases")


def test_compat_callback_rejects_non_finite_maps(A, ctx, compat_cases):
    """A map with a NaN or an infinite coordinate, in the first chunk of the upload
    or a later one, is refused before any kernel runs; the next callback on the
    same context still matches the one before it bit for bit."""
    c = compat_cases[0]
    first = ctx.compat_callback(c["scan"], c["guess"], c["map"])
    big = np.concatenate([c["map"]] * max(1, -(-40000 // len(c["map"]))))[:40000].copy()
    for at, v in ((5, np.nan), (20000, np.inf), (39999, -np.inf)):
        bad = big.copy()
        bad[at, 1] = v
        with pytest.raises(A.NdtError) as e:
            ctx.compat_callback(c["scan"], c["guess"], bad)
        assert e.value.code == 1 and "finite" in str(e.value), at
    again = ctx.compat_callback(c["scan"], c["guess"], c["map"])
    assert np.array_equal(first["final_f32"], again["final_f32"]) and first["iterations"] == again["iterations"]


def test_daphne_harness_end_to_end(tmp_path):
    """The DAPHNE testcase binary (same interface, files and check as the reference)
    on a reference-produced output file: exits 0 with "result ok"."""
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(repo, "daphne-benchmark_amd", "ndt_mapping_hip")
    assert os.path.exists(exe), "build the harness first"
    data = tmp_path / "data"
    run = tmp_path / "a" / "b" / "c"
    data.mkdir()
    run.mkdir(parents=True)
    import gzip
    for f in ("ndt_input.dat", "ndt_output.dat"):
        with gzip.open(os.path.join(repo, "tests", "golden", "daphne_small", f + ".gz")) as src:
            (data / f).write_bytes(src.read())
    # blocks of 1, 2 and 3 testcases; the testcases of a block run concurrently on
    # up to NDT_HARNESS_STREAMS contexts
    for blk, streams in (("1", "4"), ("2", "4"), ("3", "3"), ("3", "1")):
        env = dict(os.environ, NDT_HARNESS_STREAMS=streams)
        p = subprocess.run([exe, "-p", blk], cwd=run, capture_output=True, text=True, timeout=120, env=env)
        assert p.returncode == 0 and "result ok" in p.stdout, (blk, streams, p.stdout + p.stderr)


def test_fitness_score_matches_oracle(A, orc, ctx, gpu_grid, cfg0):
    """getFitnessScore on the GPU (nearest map point by shell search over the
    voxel-sorted map) against the oracle's brute force, at the truth, at the guess
    and with a range limit."""
    scene, mp, _ = cfg0
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 80 + k) for k, n in enumerate([1500, 400, 2500])]
    sc = A.Scans(ctx, [m[0] for m in made])
    for poses, rng in ((np.stack([m[1] for m in made]), 0.0), (np.stack([m[2] for m in made]), 0.0),
                       (np.stack([m[2] for m in made]), 0.3)):
        fit, cnt = sc.fitness(gpu_grid, poses, rng)
        for k, m in enumerate(made):
            want, n = orc.fitness_score(mp, m[0], poses[k], rng if rng > 0 else np.inf)
            assert cnt[k] == n and abs(fit[k] - want) <= 1e-9 * want, (k, rng, fit[k], want, cnt[k], n)


def test_captured_submission_replays(A, ctx, gpu_grid, cfg0):
    """Device-resident batches are submitted as a captured graph, replayed for new
    guesses.  Every replay, with eager work and synchronisations in between, equals
    the eager (uncaptured) alignment bit for bit."""
    scene, mp, _ = cfg0
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 120 + k) for k, n in enumerate([900, 2600, 1700, 3100, 600])]
    pts = [m[0] for m in made]
    g1 = np.stack([m[2] for m in made])
    g2 = A.perturb_poses(np.stack([m[1] for m in made]), np.random.default_rng(3), xy=0.5, yaw_deg=2.0)
    prm = A.params(A.NDT_MODE_NDT)
    eager = {1: gpu_grid.align_batch(pts, g1, prm=prm), 2: gpu_grid.align_batch(pts, g2, prm=prm)}
    sc = A.Scans(ctx, pts)
    for k, which in enumerate([1, 2, 1, 1, 2]):
        res = sc.align(gpu_grid, g1 if which == 1 else g2, prm)
        for r, e in zip(res, eager[which]):
            assert np.array_equal(r["final_f64"], e["final_f64"]) and r["passes"] == e["passes"], (k, which)
        # eager work between replays
        gpu_grid.align_batch(pts[:2], g2[:2], prm=prm)


def test_autotune_chooses_and_caches(A, ctx, gpu_grid, cfg0, tmp_path):
    scene, mp, _ = cfg0
    made = [A.make_scan(scene, mp, A.scan_cfg(n), 140 + k) for k, n in enumerate([2500, 4000, 1200, 3300])]
    sc = A.Scans(ctx, [m[0] for m in made])
    gs = np.stack([m[2] for m in made])
    cache = str(tmp_path / "tune.txt")
    prm, ms = sc.autotune(gpu_grid, gs, A.params(A.NDT_MODE_NDT), cache)
    assert prm.block_size in (128, 256, 512) and prm.points_per_wg in (1024, 2048, 4096) and ms > 0
    again, ms2 = sc.autotune(gpu_grid, gs, A.params(A.NDT_MODE_NDT), cache)
    assert (again.block_size, again.points_per_wg) == (prm.block_size, prm.points_per_wg)
    assert abs(ms2 - ms) < 1e-3 and len(open(cache).read().splitlines()) == 1
    res = sc.align(gpu_grid, gs, prm)
    assert all(r["status"] == 0 for r in res)
