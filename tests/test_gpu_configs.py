"""GPU parity at the benchmark's own scales, against the float64 oracle started from
the same guesses:

  cfg1  plain single-scan NDT at real scale: a 1 M-point map of a 300 m scene and
        sequential 20 k-point scans (guesses +-0.5 m / +-2 deg), one ndt_align call
        each;
  cfg2  the 512-scan batch (1 k - 60 k points, guesses +-1 m / +-3 deg yaw /
        +-0.5 deg roll-pitch) in one work-queue launch, a fixed 64-scan oracle
        sample;
  cfg3  the cfg2 batch with the scene at UTM-like coordinates (+4.2e5, +5.6e6,
        +80 m) in each of the four arithmetic modes against the all-float64 oracle.

These trajectories are sensitive: the oracle itself lands elsewhere on some scans
when its guess moves by 1e-12 m.  A scan is "reproducible" when the oracle's results
from the guess and from the guess shifted by 1e-12 m in x agree within 1e-4 m,
1e-5 rad, with the same iterations; the strict bars apply to reproducible scans and
the others are reported (printed) and held to the metric's bar only where stated."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAMPLE = 64


def _oracle_runs(orc, og, scans, starts, origin):
    """Oracle results (map-local) from world starts, and per-scan reproducibility."""
    local = np.array(starts, dtype=np.float64).copy()
    local[:, :3, 3] -= np.asarray(origin)
    ref = og.align_many(scans, local, threads=0)
    shifted = local.copy()
    shifted[:, 0, 3] += 1e-12
    ref2 = og.align_many(scans, shifted, threads=0)
    stable = []
    for a, b in zip(ref, ref2):
        dt, dr = orc.pose_error(a["final_f64"], b["final_f64"])
        stable.append(dt <= 1e-4 and dr <= 1e-5 and a["iterations"] == b["iterations"])
    for r in ref:
        r["final_world"] = r["final_f64"].copy()
        r["final_world"][:3, 3] += np.asarray(origin)
    return ref, np.array(stable)


def _errors(orc, res, ref):
    return np.array([orc.pose_error(r["final_f64"], o["final_world"]) for r, o in zip(res, ref)])


@pytest.fixture(scope="module")
def cfg2(A, orc, ctx):
    w = A.make_workload("cfg2", seed=1)
    prm = A.params(A.NDT_MODE_NDT, origin=w.origin, block_size=256, points_per_wg=4096)
    grid = A.Grid(ctx, w.map_xyzw, prm)
    scans = A.Scans(ctx, w.scans)
    idx = np.sort(np.random.default_rng(0).choice(len(w.scans), size=SAMPLE, replace=False))
    og = orc.NdtGrid(w.map_xyzw)
    ref, stable = _oracle_runs(orc, og, [w.scans[i] for i in idx], w.guesses[idx], w.origin)
    return w, prm, grid, scans, idx, ref, stable


@pytest.fixture(scope="module")
def cfg3(A, orc, ctx):
    """The cfg3 workload, its grid, resident scans and the oracle sample, built once
    for the four modes."""
    w = A.make_workload("cfg3", seed=1)
    grid = A.Grid(ctx, w.map_xyzw, A.params(A.NDT_MODE_NDT, origin=w.origin))
    scans = A.Scans(ctx, w.scans)
    idx = np.sort(np.random.default_rng(0).choice(len(w.scans), size=SAMPLE, replace=False))
    og = orc.NdtGrid(w.map_xyzw)
    ref, stable = _oracle_runs(orc, og, [w.scans[i] for i in idx], w.guesses[idx], w.origin)
    return w, idx, ref, stable, grid, scans


def test_cfg1_sequential_scans_match_oracle(A, orc, ctx):
    """16 sequential 20 k-point scans on the 1 M-point map: within 1e-3 m / 1e-4 rad
    with the same iterations, passes and flag on every reproducible scan, within
    1 cm / 0.05 deg on all of them; with the default chunk of 2048 points and with
    the 256-point chunks that minimise a lone scan's latency."""
    w = A.make_workload("cfg1", seed=1, scans=16)
    prm = A.params(A.NDT_MODE_NDT)
    grid = A.Grid(ctx, w.map_xyzw, prm)
    assert grid.info()["map_points"] == 1_000_000
    og = orc.NdtGrid(w.map_xyzw)
    ref, stable = _oracle_runs(orc, og, w.scans, w.guesses, w.origin)
    assert stable.sum() >= 12
    for chunk in (0, 256):
        q = A.params(A.NDT_MODE_NDT, points_per_wg=chunk)
        res = [grid.align(s, g, prm=q) for s, g in zip(w.scans, w.guesses)]
        err = _errors(orc, res, ref)
        print(f"cfg1, chunk {chunk}: {stable.sum()} of {len(stable)} scans reproducible; max error "
              f"{err[:, 0].max():.2e} m {err[:, 1].max():.2e} rad", flush=True)
        for k, (r, o) in enumerate(zip(res, ref)):
            assert r["status"] == 0
            assert err[k, 0] <= 0.01 and err[k, 1] <= np.deg2rad(0.05), (chunk, k, err[k])
            if stable[k]:
                assert err[k, 0] <= 1e-3 and err[k, 1] <= 1e-4, (chunk, k, err[k])
                assert (r["iterations"], r["passes"], r["converged"]) == \
                    (o["iterations"], o["passes"], o["converged"]), (chunk, k)


def test_cfg2_batch_matches_oracle_sample(A, orc, cfg2):
    """The whole 512-scan batch in one launch; on the 64-scan oracle sample every
    reproducible scan within 1 cm / 0.05 deg and with the oracle's converged flag (and
    1e-3 m / 1e-4 rad with the same iterations on most of them).  On the other scans
    the oracle's own flag depends on the last bit of its guess, so their flags are
    only reported."""
    w, prm, grid, scans, idx, ref, stable = cfg2
    res_all = scans.align(grid, w.guesses, prm)
    assert len(res_all) == 512 and all(r["status"] == 0 for r in res_all)
    res = [res_all[i] for i in idx]
    err = _errors(orc, res, ref)
    metric = (err[:, 0] <= 0.01) & (err[:, 1] <= np.deg2rad(0.05))
    strict = (err[:, 0] <= 1e-3) & (err[:, 1] <= 1e-4) & np.array(
        [(r["iterations"], r["converged"]) == (o["iterations"], o["converged"]) for r, o in zip(res, ref)])
    print(f"cfg2: {stable.sum()} of {SAMPLE} sampled scans reproducible; within 1 cm / 0.05 deg "
          f"{metric[stable].sum()} of them, {metric[~stable].sum()} of {(~stable).sum()} others; strict "
          f"{strict[stable].sum()}", flush=True)
    assert stable.sum() >= SAMPLE // 2
    assert metric[stable].all(), np.nonzero(stable & ~metric)[0]
    assert strict[stable].mean() >= 0.9
    assert all(res[k]["converged"] == ref[k]["converged"] for k in np.nonzero(stable)[0])
    print(f"cfg2: converged {sum(r['converged'] for r in res)} of {SAMPLE} sampled scans (oracle "
          f"{sum(o['converged'] for o in ref)}), {np.mean([r['converged'] for r in res_all]):.3f} of the batch",
          flush=True)


@pytest.mark.parametrize("mode", [("fp64", 1, 2), ("fp32/fp32", 0, 0), ("fp32/comp32", 0, 1), ("fp32/fp64", 0, 2)],
                         ids=lambda m: m[0])
def test_cfg3_precision_modes_against_fp64_oracle(A, orc, cfg3, mode):
    """The cfg2 batch at UTM-like coordinates.  float64 point math meets the metric's
    1 cm / 0.05 deg on every reproducible sampled scan.  The float32 modes (per-pair
    float32 on cell-relative coordinates, three kinds of sums) keep the per-pass
    error near 1e-7 relative (test_gpu_precision.py), which these sensitive
    trajectories amplify: they are characterised, not held to the metric: at least
    60 % of the reproducible sampled scans within 1 cm / 0.05 deg and convergence
    within 5 % of the oracle's."""
    name, pp, sp = mode
    w, idx, ref, stable, grid, scans = cfg3
    prm = A.params(A.NDT_MODE_NDT, origin=w.origin, point_precision=pp, sum_precision=sp, block_size=256,
                   points_per_wg=4096)
    res_all = scans.align(grid, w.guesses, prm)
    res = [res_all[i] for i in idx]
    err = _errors(orc, res, ref)
    metric = (err[:, 0] <= 0.01) & (err[:, 1] <= np.deg2rad(0.05))
    conv = np.mean([r["converged"] for r in res])
    conv_o = np.mean([o["converged"] for o in ref])
    print(f"cfg3 {name}: within 1 cm / 0.05 deg of the float64 oracle on {metric[stable].sum()} of "
          f"{stable.sum()} reproducible sampled scans ({metric.sum()} of {SAMPLE} in all); converged "
          f"{conv:.3f} (oracle {conv_o:.3f}); median error {np.median(err[:, 0]):.2e} m", flush=True)
    assert all(r["status"] == 0 for r in res_all)
    if pp == 1:
        assert metric[stable].all(), np.nonzero(stable & ~metric)[0]
    else:
        assert metric[stable].mean() >= 0.6
        assert conv >= conv_o - 0.05
