"""The seeded scene generator: the batched, bucketed scan generator reproduces the
single-scan generator bit for bit."""
import numpy as np


def test_batched_scans_equal_single_scans(A):
    scene = A.scene_cfg(4, extent=120.0)
    mp = A.make_map(scene, 150000)
    sizes = [500, 3000, 1200, 7000, 800]
    seeds = [3, 17, 40000, 5, 123456]
    scans, gt, gs = A.make_scans(scene, mp, sizes, seeds, perturb_xyz=2.0, perturb_yaw_deg=7.0, threads=3)
    for k in range(len(sizes)):
        one, g1, u1 = A.make_scan(scene, mp, A.scan_cfg(sizes[k], 2.0, 7.0, 0.5), seeds[k])
        assert np.array_equal(one, scans[k]) and np.array_equal(g1, gt[k]) and np.array_equal(u1, gs[k])


def test_scan_points_lie_within_range(A):
    scene = A.scene_cfg(5, extent=100.0)
    mp = A.make_map(scene, 50000)
    scans, gt, _ = A.make_scans(scene, mp, [2000, 2000], [1, 2])
    for s in scans:
        r = np.linalg.norm(s[:, :3].astype(np.float64), axis=1)
        assert r.max() <= A.scan_cfg(2000).max_range + 0.1
