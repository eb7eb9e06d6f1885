"""Precision study, pass level: score, gradient and Hessian of single derivative
passes in every arithmetic mode against the float64 oracle, on scans of a scene
offset to UTM-like coordinates (+4.2e5, +5.6e6, +80 m).

float32 point math works on coordinates relative to each query cell's corner (the
transform and cell lookup are float64), so its error does not grow with the map's
distance from the origin.  Bounds (observed medians about 1e-7 for the gradient
and 2.5e-7 for the Hessian): gradient within 2e-6 and Hessian within 1e-5
relative, median over the scans; every scan within 10x that.  float64 point math:
1e-10 relative (the oracle sums in a different order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORIGIN = (4.2e5, 5.6e6, 80.0)


@pytest.fixture(scope="module")
def utm(A, orc, ctx):
    scene = A.scene_cfg(11, extent=120.0, origin=ORIGIN)
    mp = A.make_map(scene, 200000)
    og = orc.NdtGrid(mp)
    g = A.Grid(ctx, mp, A.params(A.NDT_MODE_NDT, origin=ORIGIN))
    scans = [A.make_scan(scene, mp, A.scan_cfg(n, 1.0, 3.0, 0.5), 20 + k)
             for k, n in enumerate([3000, 8000, 1500, 12000, 5000, 2500, 6000, 4000])]
    return og, g, scans


def _params(M):
    R = M[:3, :3]
    return np.array([M[0, 3] - ORIGIN[0], M[1, 3] - ORIGIN[1], M[2, 3] - ORIGIN[2], np.arctan2(-R[1, 2], R[2, 2]),
                     np.arcsin(np.clip(R[0, 2], -1, 1)), np.arctan2(-R[0, 1], R[0, 0])])


@pytest.mark.parametrize("mode", [("fp64", 1, 2, 1e-10, 1e-10), ("fp32/fp32", 0, 0, 2e-6, 1e-5),
                                  ("fp32/comp32", 0, 1, 2e-6, 1e-5), ("fp32/fp64", 0, 2, 2e-6, 1e-5)],
                         ids=lambda m: m[0])
def test_pass_precision_against_oracle(A, utm, mode):
    name, pp, sp, gtol, htol = mode
    og, g, scans = utm
    prm = A.params(A.NDT_MODE_NDT, point_precision=pp, sum_precision=sp, origin=ORIGIN)
    eg, eh = [], []
    for scan, gt, guess in scans:
        p = _params(guess)
        s1, g1, h1, n1 = g.derivatives(scan, p, prm=prm)
        s0, g0, h0, n0 = og.derivatives(scan, p)
        # float32 range tests may differ from float64 ones for pairs on the sphere
        assert n1 == n0 if pp == 1 else abs(n1 - n0) <= max(2, 1e-3 * n0), (name, n1, n0)
        eg.append(np.linalg.norm(g1 - g0) / np.linalg.norm(g0))
        eh.append(np.linalg.norm(h1 - h0) / np.linalg.norm(h0))
    eg, eh = np.array(eg), np.array(eh)
    assert np.median(eg) <= gtol and eg.max() <= 10 * gtol, (name, eg)
    assert np.median(eh) <= htol and eh.max() <= 10 * htol, (name, eh)
