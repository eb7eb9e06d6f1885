"""GPU parity of the hashed query-cell index.  Maps whose cell box exceeds 2^28 cells
(multi-kilometre maps at fine resolutions) look up their query cells in an open-
addressing hash and evaluate each point's pairs on its own thread (hash_find and
point_accumulate instead of the dense index and the wave-balanced pair lists).  The
environment variable NDT_FORCE_HASHED_INDEX=1 makes a grid of any size use that
path, so it is checked here against the oracle on the cfg0 scene (100 k-point map,
2 k-point scans), to the same bars as the dense path: pair counts exact, float64
passes within 1e-10 relative, final poses within 1e-3 m / 1e-4 rad with the same
iterations, passes and flags."""
import os

import numpy as np
import pytest

from test_gpu_parity import _compare_runs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(A, orc):
    sc = A.scene_cfg(1)
    mp = A.make_map(sc, 100000)
    return sc, mp, orc.NdtGrid(mp)


@pytest.fixture(scope="module")
def hashed_grid(A, ctx, scene):
    old = os.environ.get("NDT_FORCE_HASHED_INDEX")
    os.environ["NDT_FORCE_HASHED_INDEX"] = "1"
    try:
        g = A.Grid(ctx, scene[1], A.params(A.NDT_MODE_NDT))
    finally:
        if old is None:
            del os.environ["NDT_FORCE_HASHED_INDEX"]
        else:
            os.environ["NDT_FORCE_HASHED_INDEX"] = old
    return g


def test_index_kind(A, ctx, scene, hashed_grid):
    dense = A.Grid(ctx, scene[1], A.params(A.NDT_MODE_NDT))
    assert hashed_grid.info()["dense_cells"] == 0
    assert dense.info()["dense_cells"] > 0
    for k in ("voxels", "valid_voxels", "query_cells", "candidates"):
        assert hashed_grid.info()[k] == dense.info()[k]


def test_hashed_derivative_pass(A, orc, scene, hashed_grid):
    sc, mp, og = scene
    for s in range(3):
        scan, gt, guess = A.make_scan(sc, mp, A.scan_cfg(2000), s)
        p = np.r_[guess[:3, 3], orc.ndt_euler(guess)] + np.array([0.02, -0.01, 0.0, 0.01, -0.02, 0.004])
        so, go, Ho, npo = og.derivatives(scan, p)
        sg, gg, Hg, npg = hashed_grid.derivatives(scan, p, prm=A.params(A.NDT_MODE_NDT))
        assert npg == npo
        assert abs(sg - so) <= 1e-10 * abs(so)
        assert np.abs(gg - go).max() <= 1e-10 * np.abs(go).max()
        assert np.abs(Hg - Ho).max() <= 1e-10 * np.abs(Ho).max()


def test_hashed_align_matches_oracle(A, orc, scene, hashed_grid):
    sc, mp, og = scene
    scans = [A.make_scan(sc, mp, A.scan_cfg(2000), s) for s in range(12)]
    prm = A.params(A.NDT_MODE_NDT)
    res = [hashed_grid.align(s[0], s[2], prm=prm) for s in scans]
    ref = og.align_many([s[0] for s in scans], np.stack([s[2] for s in scans]), threads=8)
    assert _compare_runs(res, ref, 1e-3, 1e-4, orc) == []


def test_hashed_batch_matches_oracle(A, orc, scene, hashed_grid):
    """Uneven scans in one work-queue launch through the hashed index."""
    sc, mp, og = scene
    sizes = [600, 3500, 1200, 2500, 900, 4200, 1500, 2100]
    scans = [A.make_scan(sc, mp, A.scan_cfg(n), 80 + k) for k, n in enumerate(sizes)]
    pts, gs = [s[0] for s in scans], np.stack([s[2] for s in scans])
    res = hashed_grid.align_batch(pts, gs, prm=A.params(A.NDT_MODE_NDT))
    ref = og.align_many(pts, gs, threads=8)
    assert _compare_runs(res, ref, 1e-3, 1e-4, orc) == []
