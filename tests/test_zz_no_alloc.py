"""The allocation-free entry point ndt_align_scans_noalloc, called through the C ABI,
gives the same results as ndt_align_scans (same kernels, launch shape and sum
order: equal bit for bit), on two calls in a row, the second with a smaller batch
whose workspace regions start where the first call left data behind.  The device's
free memory does not change across a call, and batches whose workspace exceeds the
fixed array are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.fixture(scope="module")
def scene_grid(A, ctx):
    scene = A.scene_cfg(1)
    mp = A.make_map(scene, 100000)
    return scene, mp, A.Grid(ctx, mp, A.params(A.NDT_MODE_NDT))


def _same(res, ref):
    assert len(res) == len(ref)
    for i, (r, o) in enumerate(zip(res, ref)):
        for key in ("final_f64", "final_f32"):
            assert np.array_equal(r[key], o[key]), (i, key)
        for key in ("score", "trans_probability", "iterations", "converged", "passes", "status"):
            assert r[key] == o[key], (i, key, r[key], o[key])


def test_noalloc_matches_align_scans(A, ctx, scene_grid):
    scene, mp, grid = scene_grid
    batches = [[700, 5200, 1300, 2600, 900, 4100, 1600, 2100, 3000, 450, 3600, 1900],
               [1100, 300, 2400]]
    prm = A.params(A.NDT_MODE_NDT)
    for b, sizes in enumerate(batches):
        made = [A.make_scan(scene, mp, A.scan_cfg(n), 500 + 20 * b + k) for k, n in enumerate(sizes)]
        scans = A.Scans(ctx, [m[0] for m in made])
        guesses = np.stack([m[2] for m in made])
        print(f"noalloc batch {b}: {len(sizes)} scans", flush=True)
        ref = scans.align(grid, guesses, prm=prm)
        free0 = _free_device_bytes()
        res = scans.align_noalloc(grid, guesses, prm=prm)
        free1 = _free_device_bytes()
        _same(res, ref)
        assert free1 == free0
        assert all(r["status"] == 0 for r in res)
        _, launches, pairs = ctx.last_timing()
        assert launches == 1 and pairs > 0
        scans.close()


def test_noalloc_refuses_oversized_batches(A, ctx, scene_grid):
    scene, mp, grid = scene_grid
    made = [A.make_scan(scene, mp, A.scan_cfg(4000), 600 + k) for k in range(4)]
    scans = A.Scans(ctx, [m[0] for m in made])
    guesses = np.stack([m[2] for m in made])
    # 64 partial-sum slots of 256 points, each with room for 1 + 11 * 20000 passes of
    # queue items: about 113 MB of items alone
    big = A.params(A.NDT_MODE_NDT, points_per_wg=256, max_iterations=20000)
    with pytest.raises(A.NdtError) as err:
        scans.align_noalloc(grid, guesses, prm=big)
    assert err.value.code == 5 and "workspace" in str(err.value)
    with pytest.raises(A.NdtError) as err:
        scans.align_noalloc(grid, guesses, prm=A.params(A.NDT_MODE_NDT, scheduler=A.NDT_SCHED_ROUNDS))
    assert err.value.code == 5
    # the refusals leave the entry point usable
    _same(scans.align_noalloc(grid, guesses, prm=A.params(A.NDT_MODE_NDT)),
          scans.align(grid, guesses, prm=A.params(A.NDT_MODE_NDT)))
    scans.close()
