"""Benchmark: NDT scan alignments per second on one MI355X, batched + learned warm
start against plain single-scan NDT.

Workload (default, the batched configuration): a 1,000,000-point map of a
300 m x 300 m synthetic scene and a batch of 512 scans with point counts
log-uniform in [1,000, 60,000], initial guesses perturbed by +-1 m, +-3 deg yaw,
+-0.5 deg roll/pitch; NDT with 1 m voxels, step 0.1, epsilon 0.01, 30 iterations,
outlier ratio 0.55.  The map's voxel grid and the pose proposer are built before
the timed region (both reported); scans are resident in device memory.

A step aligns the whole batch with the learned warm start (fallback policy, the
refined guess taking the model's translation): the scan grids and the map grids
at the guesses, the bf16 MLP forward pass, then one work-queue launch in which
every scan takes one pass at its given guess and one at the refined guess, runs
NDT from the refined guess where it scores at least as high, and from the given
guess otherwise or where the refined run ended below the given guess's score; the
results are read back.  The scans are resident in device
memory; a second figure includes their upload (packing, Morton order, copy) in
the step.  The same scans are also aligned batched without the proposer and one
at a time ("plain" single-scan GPU NDT, one launch and host sync per scan).

The proposer is trained on the GPU before the timed region on 200,000 (scan,
guess, NDT-converged pose) triples from scans of the same map at other sensor
poses (guesses within +-3 m / +-10 deg yaw, or +-1 m / +-3 deg), and evaluated on
the batch (whose scans it has not seen) and on 2,048 further held-out scans.
Also reported: cfg1 (1,000 sequential 20 k-point scans on the 1 M-point map),
the reference's own harness on cfg0- and cfg1-sized testcases, and the batch with
the Hessian computed only at accepted steps.

Prints one JSON line (the benchmark contract).  The chosen runs are checked
against the float64 CPU oracle started from the same guesses on a sample of
scans, and the oracle's throughput on the host cores is the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "daphne-benchmark_amd"))
import ndt_amd as A  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E

TRAFFIC_PROFILE = "profiles/k_align_wq_cfg2_batch512_fp64.md"


def profiled_traffic(args, prm, kind="plain"):
    """HBM bytes per align launch (kind "plain": the batch without the proposer,
    "learned": the headline step's launch) measured with rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, calibrated) and recorded in the committed profile, when that profile
    covers this workload, precision and launch shape; None otherwise."""
    import re
    try:
        with open(os.path.join(REPO, TRAFFIC_PROFILE)) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch " + kind + r": (\d+) \(workload (\w+), (\d+) scans, (\w+), block (\d+), "
                  r"chunk (\d+), fine_pass (\d+)\)", text)
    if not m:
        return None
    shape = (prm.block_size or 256, prm.points_per_wg or 2048, prm.fine_pass or 24)
    if (m.group(2), int(m.group(3)), m.group(4), (int(m.group(5)), int(m.group(6)), int(m.group(7)))) != (
            args.workload, args.scans, args.precision, shape):
        return None
    return int(m.group(1))


FP64_PEAK_TFLOPS = 78.6        # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3       # MI355X vector FP32
BF16_PEAK_TFLOPS = 2500.0      # MI355X dense bf16 MFMA
MLP_FWD_FLOPS = 2.0 * (65536 * 1024 + 1024 * 256 + 256 * 6)  # per sample, 65536-1024-256-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_work(results, counts, pairs):
    """Work of the derivative passes summed over a launch: bytes = 16 N (scan
    point) + 36 P (voxel mean and inverse covariance per pair), flops = 133 N +
    340 P (transform and point Jacobian per point, score / gradient / Hessian per
    pair); N counts every point of every pass, P every (point, voxel) pair."""
    pts = float(sum(r["passes"] * n for r, n in zip(results, counts)))
    return 16.0 * pts + 36.0 * pairs, 133.0 * pts + 340.0 * pairs


def coarse_register_params():
    """the first registration of training guesses: +-4 m, +-10 deg in 2 deg steps"""
    return A.register_params(radius=8, yaws=11, yaw_step=float(np.radians(2.0)), level=0)


REGISTERED_LOSS_WEIGHTS = [0.1, 0.1, 1.0, 1.0, 1.0, 1.0]


def train_proposer(ctx, w, grid, bm, prm, n_scans, per_scan, epochs, weight_decay, rp=None, seed=5):
    """Training triples from scans of the workload's map at other sensor poses
    (generator seeds disjoint from the batch's), labelled with NDT run from the
    true pose; half the guesses within +-3 m / +-10 deg yaw, half within
    +-1 m / +-3 deg.  With a registration (rp) the guesses are registered first, so
    that the model learns the correction of the registered guess, and the model
    gets the same setting: a coarse search (+-4 m, +-10 deg in 2 deg steps) and
    then rp, so that the wide guesses' residuals are as small as the narrow ones'
    (the model otherwise learns corrections from the wide guesses' failed
    registrations and moves good registered guesses); the x and y terms of the loss
    weigh 0.1, the registration having set x and y."""
    rng = np.random.default_rng(seed)
    ds, t_gen, t_lab = None, 0.0, 0.0
    chunk = 10000
    for c0 in range(0, n_scans, chunk):
        n = min(chunk, n_scans - c0)
        t = time.time()
        tr, tr_gt, _ = A.make_scans(w.scene, w.map_xyzw, A.scan_sizes(1000, 60000, n, 99 + c0),
                                    100000 + c0 + np.arange(n))
        t_gen += time.time() - t
        t = time.time()
        trS = A.Scans(ctx, tr)
        lab = trS.align(grid, tr_gt, prm)
        t_lab += time.time() - t
        labels = np.stack([r["final_f64"] for r in lab])
        ok = np.array([r["converged"] and np.linalg.norm(r["final_f64"][:3, 3] - g[:3, 3]) < 0.2
                       for r, g in zip(lab, tr_gt)])
        idx = np.repeat(np.nonzero(ok)[0], per_scan).astype(np.int32)
        wide = rng.uniform(size=len(idx)) < 0.5
        g1 = A.perturb_poses(tr_gt[idx], rng)
        g2 = A.perturb_poses(tr_gt[idx], rng, xy=1.0, z=1.0, yaw_deg=3.0)
        guesses = np.where(wide[:, None, None], g1, g2)
        if rp is not None:
            guesses = bm.register(trS, idx, bm.register(trS, idx, guesses, coarse_register_params())[0], rp)[0]
        if ds is None:
            ds = A.Dataset(ctx, bm, trS, idx, guesses, labels[idx])
        else:
            ds.add(trS, idx, guesses, labels[idx])
        del trS, tr
    model = A.Model(ctx, seed=1)
    model.set_register(rp)
    t = time.time()
    tp = dict(epochs=epochs, batch=512, weight_decay=weight_decay)
    if rp is not None:
        tp["weights"] = REGISTERED_LOSS_WEIGHTS
    losses = model.train(ds, A.train_params(**tp))
    t_train = time.time() - t
    samples = int(ds.n)
    ds.close()
    return model, dict(samples=samples, scans=n_scans, epochs=epochs,
                       weight_decay=weight_decay, generate_s=t_gen, label_s=t_lab, train_s=t_train,
                       loss_first=float(losses[0]), loss_last=float(losses[-1]))


def cfg4_leg(ctx, w, grid, bm, model, prm, prm_l, policy, n_scans, steps) -> dict:
    """Held-out evaluation of the learned warm start: n_scans scans of the batched
    kind (1 k - 60 k points, log-uniform; guesses +-1 m / +-3 deg yaw / +-0.5 deg
    roll-pitch) of the same map, from generator seeds disjoint from the training
    scans and the benchmark batch, aligned in one batch from the given guesses
    (plain NDT) and with the learned warm start.  Reported: Newton iterations and
    derivative passes per scan (the learned figure counts every pass spent: the
    probe, the refined run and any fallback), alignments/s of both, and the learned
    results' distance from the plain ones and from the generator's truth."""
    ev, ev_gt, ev_g = A.make_scans(w.scene, w.map_xyzw, A.scan_sizes(1000, 60000, n_scans, 4242),
                                   700000 + np.arange(n_scans), perturb_xyz=1.0, perturb_yaw_deg=3.0,
                                   perturb_rollpitch_deg=0.5)
    evS = A.Scans(ctx, ev)
    evS.align(grid, ev_g, prm)
    t = time.perf_counter()
    for _ in range(steps):
        plain = evS.align(grid, ev_g, prm)
    plain_s = (time.perf_counter() - t) / steps
    model.align(grid, bm, evS, ev_g, prm_l, policy)
    t = time.perf_counter()
    for _ in range(steps):
        res, used = model.align(grid, bm, evS, ev_g, prm_l, policy)
    learned_s = (time.perf_counter() - t) / steps
    evS.close()
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    err = np.array([orc.pose_error(a["final_f64"], b["final_f64"]) for a, b in zip(res, plain)])

    def truth_ok(R):
        return float(np.mean([np.linalg.norm(r["final_f64"][:3, 3] - g[:3, 3]) < 0.05 for r, g in zip(R, ev_gt)]))

    it_p = float(np.mean([r["iterations"] for r in plain]))
    it_l = float(np.mean([r["iterations"] for r in res]))
    pa_p = float(np.mean([r["passes"] for r in plain]))
    pa_l = float(np.mean([r["passes"] for r in res]))
    return {"held_out_scans": n_scans, "points": int(sum(e.shape[0] for e in ev)),
            "plain": {"alignments_per_s": n_scans / plain_s, "mean_iterations": it_p, "mean_passes": pa_p,
                      "converged": float(np.mean([r["converged"] for r in plain])), "within_5cm_of_truth": truth_ok(plain)},
            "learned": {"alignments_per_s": n_scans / learned_s, "mean_iterations": it_l, "mean_passes": pa_l,
                        "converged": float(np.mean([r["converged"] for r in res])), "within_5cm_of_truth": truth_ok(res),
                        "refined_start_kept": float(np.mean(used))},
            "iterations_saved_per_scan": it_p - it_l, "passes_saved_per_scan": pa_p - pa_l,
            "learned_vs_plain": {"median_translation_m": float(np.median(err[:, 0])),
                                 "p95_translation_m": float(np.percentile(err[:, 0], 95)),
                                 "median_rotation_rad": float(np.median(err[:, 1])),
                                 "within_1cm_0.05deg": float(np.mean((err[:, 0] <= 0.01) &
                                                                     (err[:, 1] <= np.deg2rad(0.05))))}}


def _harness_time(exe, cwd, args=(), env=None):
    """Runs a DAPHNE testcase binary; returns (seconds per testcase, result ok)."""
    import subprocess
    p = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=600, env=env)
    per = None
    for line in p.stdout.splitlines():
        if line.startswith("elapsed time:") and "):" in line:
            per = float(line.rsplit("):", 1)[1].split()[0])
    return per, (p.returncode == 0 and "result ok" in p.stdout)


def daphne_harness_leg(cases_n: int, extent: float = 100.0, map_points: int = 100_000,
                       scan_points: int = 2000, seed: int = 3) -> dict:
    """The reference's own benchmark harness and the drop-in on one DAPHNE testcase
    file (by default cfg0-sized testcases: 100 k-point map, 2 k-point scan, guess
    within 0.5 m / 2 deg).  The compiled reference (oracle/_ref: its
    single-threaded CPU build and its OpenMP build) and
    daphne-benchmark_amd/ndt_mapping_hip read the same ndt_input.dat and check the
    same ndt_output.dat, written by the reference's datagen build; each prints its
    kernel time per testcase (grid build and alignment, compat semantics)."""
    import shutil
    import subprocess
    import tempfile
    ref = os.path.join(REPO, "oracle", "_ref")
    exe = {k: os.path.join(ref, f) for k, f in (("datagen", "ndt_ref_datagen"), ("cpu", "ndt_ref_cpu"),
                                                 ("omp", "ndt_ref_omp"))}
    hip = os.path.join(REPO, "daphne-benchmark_amd", "ndt_mapping_hip")
    missing = [v for v in list(exe.values()) + [hip] if not os.path.exists(v)]
    if missing:
        return {"skipped": "not built: " + ", ".join(os.path.relpath(m, REPO) for m in missing)}
    scene = A.scene_cfg(seed, extent=extent)
    mp = A.make_map(scene, map_points)
    cases = []
    for k in range(cases_n):
        scan, _gt, guess = A.make_scan(scene, mp, A.scan_cfg(scan_points, 0.5, 2.0), 500 + k)
        cases.append(A.Testcase(guess.astype(np.float32), scan, mp))
    d = tempfile.mkdtemp()
    try:
        data, run = os.path.join(d, "data"), os.path.join(d, "a", "b", "c")
        os.makedirs(data)
        os.makedirs(run)
        A.write_ndt_input(os.path.join(data, "ndt_input.dat"), cases)
        # the datagen build compares against ndt_output.dat while writing dategen.dat
        A.write_ndt_output(os.path.join(data, "ndt_output.dat"),
                           [A.CallbackResult(True, np.eye(4, dtype=np.float32))] * cases_n)
        subprocess.run([exe["datagen"]], cwd=run, capture_output=True, timeout=600)
        res = A.read_ndt_output(os.path.join(run, "dategen.dat"))
        for r in res:
            r.fitness_score = 0.0  # never set by the reference
        A.write_ndt_output(os.path.join(data, "ndt_output.dat"), res)
        threads = os.cpu_count() or 1
        out = {"testcases": cases_n, "map_points": int(mp.shape[0]), "scan_points": scan_points,
               "unit": "seconds per testcase (the harness's own timer)"}
        out["reference_cpu"], ok_cpu = _harness_time(exe["cpu"], run)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        out["reference_openmp"], ok_omp = _harness_time(exe["omp"], run, env=env)
        out["reference_openmp_threads"] = threads
        _harness_time(hip, run)  # warm-up: device initialisation and code loading
        out["hip_p1"], ok_h1 = _harness_time(hip, run)
        out["hip_p8"], ok_h8 = _harness_time(hip, run, ("-p", "8"))
        out["hip_p16"], ok_h16 = _harness_time(hip, run, ("-p", "16"))
        out["result_ok"] = {"reference_cpu": ok_cpu, "reference_openmp": ok_omp, "hip_p1": ok_h1, "hip_p8": ok_h8,
                            "hip_p16": ok_h16}
        if out["reference_openmp"]:
            for k in ("p1", "p8", "p16"):
                if out["hip_" + k]:
                    out[f"speedup_{k}_vs_openmp"] = out["reference_openmp"] / out["hip_" + k]
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cfg1_leg(ctx, n_scans: int, cpu_sample: int, harness_cases: int, chunk: int = 256) -> dict:
    """Plain NDT at real scale: a 1 M-point map of a 300 m scene, its voxel grid
    built on the GPU, and n_scans sequential 20 k-point scans (guesses +-0.5 m /
    +-2 deg), one ndt_align call per scan from host memory (upload, Morton order and
    one launch each, as a caller aligning scans as they arrive would).  Beside it:
    the float64 oracle on a sample of the same scans on the host cores, and the
    reference's own harness (compiled reference, CPU and OpenMP builds, grid rebuilt
    per testcase) on testcases of the same sizes.  `chunk` points per partial sum
    (256 by default: a lone 20 k-point scan then spreads each pass over 79
    workgroups instead of 10)."""
    t0 = time.time()
    w = A.make_workload("cfg1", seed=1, scans=n_scans)
    gen_s = time.time() - t0
    prm = A.params(A.NDT_MODE_NDT, points_per_wg=chunk)
    grid = A.Grid(ctx, w.map_xyzw, prm)
    for i in range(min(3, n_scans)):
        grid.align(w.scans[i], w.guesses[i], prm)
    t = time.perf_counter()
    res = [grid.align(sc, g, prm) for sc, g in zip(w.scans, w.guesses)]
    wall = time.perf_counter() - t
    out = {"workload": f"cfg1: {w.map_xyzw.shape[0]}-point map of a 300 m scene, {n_scans} sequential "
                       f"{w.scans[0].shape[0]}-point scans, guesses +-0.5 m / +-2 deg",
           "value": n_scans / wall, "unit": "alignments/s", "ms_per_scan": wall / n_scans * 1e3,
           "points_per_wg": chunk, "grid_build_ms": grid.info()["build_ms"],
           "mean_iterations": float(np.mean([r["iterations"] for r in res])),
           "mean_passes": float(np.mean([r["passes"] for r in res])),
           "converged": float(np.mean([r["converged"] for r in res])),
           "success_vs_truth": float(np.mean([np.linalg.norm(r["final_f64"][:3, 3] - g[:3, 3]) < 0.05
                                              for r, g in zip(res, w.gt)])),
           "generation_s": gen_s}
    if cpu_sample > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import orc
        og = orc.NdtGrid(w.map_xyzw)
        k = min(cpu_sample, n_scans)
        threads = os.cpu_count() or 1
        tc = time.perf_counter()
        ref = og.align_many(w.scans[:k], w.guesses[:k], threads=threads)
        cpu_s = time.perf_counter() - tc
        err = np.array([orc.pose_error(r["final_f64"], o["final_f64"]) for r, o in zip(res[:k], ref)])
        out["cpu_oracle"] = {"value": k / cpu_s, "unit": "alignments/s", "cores": threads, "scans": k,
                             "kind": "port (float64 oracle, grid built once, one scan per thread)"}
        out["accuracy_vs_fp64_oracle"] = {
            "scans": k, "within_1cm_0.05deg": int(np.sum((err[:, 0] <= 0.01) & (err[:, 1] <= np.deg2rad(0.05)))),
            "within_1mm_1e-4rad": int(np.sum((err[:, 0] <= 1e-3) & (err[:, 1] <= 1e-4))),
            "same_iterations": int(sum(r["iterations"] == o["iterations"] for r, o in zip(res[:k], ref)))}
    if harness_cases > 0:
        try:
            out["daphne_harness"] = daphne_harness_leg(harness_cases, extent=300.0, map_points=1_000_000,
                                                       scan_points=20000, seed=1)
        except Exception as e:  # reported, never fatal to the bench line
            out["daphne_harness"] = {"error": f"{type(e).__name__}: {e}"}
    return out


def precision_study(args) -> dict:
    """The cfg3 batch (the cfg2 batch with the scene at UTM-like coordinates) in the
    four arithmetic modes, each against the all-float64 CPU oracle started from the
    same guesses on a sample of scans: alignments/s, convergence and per-scan
    translation / rotation error.  A sampled scan counts as reproducible when the
    oracle's result moves by less than 1e-4 m when its guess moves by 1e-12 m."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    w = A.make_workload("cfg3", seed=1, scans=args.scans)
    S = len(w.scans)
    ctx = A.Context(0)
    grid = A.Grid(ctx, w.map_xyzw, A.params(A.NDT_MODE_NDT, origin=w.origin))
    scans = A.Scans(ctx, w.scans)
    idx = np.sort(np.random.default_rng(0).choice(S, size=min(args.cpu_sample, S), replace=False))
    og = orc.NdtGrid(w.map_xyzw)
    local = w.guesses[idx].copy()
    local[:, :3, 3] -= np.array(w.origin)
    ref = og.align_many([w.scans[i] for i in idx], local, threads=0)
    local[:, 0, 3] += 1e-12
    ref2 = og.align_many([w.scans[i] for i in idx], local, threads=0)
    stable = np.array([orc.pose_error(a["final_f64"], b["final_f64"])[0] < 1e-4 for a, b in zip(ref, ref2)])
    modes = {"fp64": (A.NDT_POINT_FP64, A.NDT_SUM_FP64), "fp32/fp32": (A.NDT_POINT_FP32, A.NDT_SUM_FP32),
             "fp32/comp32": (A.NDT_POINT_FP32, A.NDT_SUM_COMP32), "fp32/fp64": (A.NDT_POINT_FP32, A.NDT_SUM_FP64)}
    out = {"workload": f"cfg3: {S} scans, origin {list(w.origin)}", "oracle_sample": len(idx),
           "oracle_reproducible": int(stable.sum()), "modes": {}}
    for name, (pp, sp) in modes.items():
        prm = A.params(A.NDT_MODE_NDT, point_precision=pp, sum_precision=sp, origin=w.origin, block_size=256,
                       points_per_wg=4096)
        scans.align(grid, w.guesses, prm)
        t = time.perf_counter()
        for _ in range(args.steps):
            res = scans.align(grid, w.guesses, prm)
        rate = S * args.steps / (time.perf_counter() - t)
        err = []
        for k, i in enumerate(idx):
            o = ref[k]["final_f64"].copy()
            o[:3, 3] += np.array(w.origin)
            err.append(orc.pose_error(res[i]["final_f64"], o))
        err = np.array(err)
        ok = (err[:, 0] <= 0.01) & (err[:, 1] <= np.deg2rad(0.05))
        out["modes"][name] = {
            "alignments_per_s": rate, "kernel_ms": ctx.last_timing()[0],
            "converged": float(np.mean([r["converged"] for r in res])),
            "within_1cm_0.05deg": int(ok.sum()), "within_1cm_0.05deg_reproducible": int(ok[stable].sum()),
            "within_1mm_1e-4rad": int(np.sum((err[:, 0] <= 1e-3) & (err[:, 1] <= 1e-4))),
            "translation_m": {"median": float(np.median(err[:, 0])), "max": float(err[:, 0].max())},
            "rotation_rad": {"median": float(np.median(err[:, 1])), "max": float(err[:, 1].max())}}
        log(f"{name}: {rate:.0f}/s, {ok[stable].sum()} of {stable.sum()} reproducible within 1 cm / 0.05 deg")
    return out


def hbm_copy_calibration(mib: int) -> dict:
    """One device-to-device copy of a known size through torch (mib MiB read, mib
    MiB written): profiling it with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE gives the
    factors between those counters and bytes moved."""
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    return {"copied_mib": mib}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scans", type=int, default=512)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--plain-scans", type=int, default=128, help="scans timed one at a time (subset of the batch)")
    ap.add_argument("--cpu-sample", type=int, default=48, help="scans aligned by the CPU oracle")
    ap.add_argument("--harness-cases", type=int, default=96,
                    help="DAPHNE testcases for the harness comparison with the compiled reference (0: off)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="batched NDT only (no pose proposer)")
    ap.add_argument("--train-scans", type=int, default=25000)
    ap.add_argument("--train-per-scan", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--weight-decay", type=float, default=0.05)
    ap.add_argument("--policy", choices=["race", "fallback"], default="fallback")
    ap.add_argument("--learned-fine-pass", type=int, default=8,
                    help="pass of a run from which the learned step's launch cuts scans into quarter-size chunks")
    ap.add_argument("--no-register", action="store_true",
                    help="no BEV registration in front of the model (it then corrects the given guess)")
    ap.add_argument("--cfg4-scans", type=int, default=2048, help="held-out scans of the learned evaluation (0: off)")
    ap.add_argument("--full-pose", action="store_true",
                    help="refined guesses take the model's rotation too (default: its translation only)")
    ap.add_argument("--no-autotune", action="store_true", help="default block size and chunk instead of the tuned ones")
    ap.add_argument("--block-size", type=int, default=0, help="fixed launch shape (with --points-per-wg): no tuning")
    ap.add_argument("--points-per-wg", type=int, default=0, help="fixed chunk size (with --block-size): no tuning")
    ap.add_argument("--cfg1-scans", type=int, default=1000, help="sequential scans of the cfg1 leg (0: off)")
    ap.add_argument("--cfg1-harness-cases", type=int, default=16,
                    help="cfg1-sized DAPHNE testcases for the reference harness comparison (0: off)")
    ap.add_argument("--precision-study", action="store_true",
                    help="only the cfg3 precision study (four arithmetic modes against the float64 oracle)")
    ap.add_argument("--only-headline", action="store_true",
                    help="only the batched launch and the headline step (for profiling: no plain single-scan, "
                         "upload, held-out, Hessian-at-accept, CPU, harness or cfg1 legs)")
    ap.add_argument("--hbm-copy-calibration", type=int, default=0, metavar="MIB",
                    help="only one device copy of MIB MiB (calibration of rocprofv3 traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); every box here has one GPU", file=sys.stderr)
        sys.exit(2)
    if args.hbm_copy_calibration:
        print(json.dumps(hbm_copy_calibration(args.hbm_copy_calibration)))
        return
    if args.precision_study:
        print(json.dumps(precision_study(args)))
        return
    if args.only_headline:
        args.plain_scans, args.cfg4_scans, args.cfg1_scans, args.no_cpu = 1, 0, 0, True

    t0 = time.time()
    w = A.make_workload(args.workload, seed=1, scans=args.scans)
    gen_s = time.time() - t0
    counts = [s.shape[0] for s in w.scans]
    S = len(w.scans)
    point_prec = A.NDT_POINT_FP64 if args.precision == "fp64" else A.NDT_POINT_FP32
    prm = A.params(A.NDT_MODE_NDT, point_precision=point_prec, origin=w.origin)

    ctx = A.Context(0)
    cus, dev_name = ctx.device_info()
    grid = A.Grid(ctx, w.map_xyzw, prm)
    ginfo = grid.info()
    scans = A.Scans(ctx, w.scans)
    log(f"workload {gen_s:.1f}s, grid {ginfo['build_ms']:.1f} ms")
    tuned_ms = None
    if args.block_size or args.points_per_wg:
        prm.block_size, prm.points_per_wg = args.block_size, args.points_per_wg
        log(f"fixed launch shape: block {prm.block_size or 256}, chunk {prm.points_per_wg or 2048}")
    elif not args.no_autotune:
        prm, tuned_ms = scans.autotune(grid, w.guesses, prm)
        log(f"autotuned: block {prm.block_size}, chunk {prm.points_per_wg} ({tuned_ms:.2f} ms)")

    # ---- batched NDT without the proposer (also the roofline measurement)
    for _ in range(args.warmup):
        scans.align(grid, w.guesses, prm)
    kernel_ms, t_b = [], time.perf_counter()
    for _ in range(args.steps):
        res_b = scans.align(grid, w.guesses, prm)
        ms, launches, pairs_b = ctx.last_timing()
        kernel_ms.append(ms)
    batched_s = (time.perf_counter() - t_b) / args.steps
    profile = ctx.last_profile()
    log(f"batched {S / batched_s:.0f}/s")

    # ---- batched + learned warm start
    learned = None
    if not args.no_learned:
        bm = A.BevMap(ctx, w.map_xyzw, w.origin)
        rp = None if args.no_register else A.register_params()
        model, tinfo = train_proposer(ctx, w, grid, bm, prm, args.train_scans, args.train_per_scan, args.epochs,
                                      args.weight_decay, rp)
        if rp is not None:
            tinfo["registration"] = dict(radius_cells=rp.radius, yaws=rp.yaws, yaw_step_deg=float(np.degrees(rp.yaw_step)),
                                         refine_yaw=bool(rp.refine_yaw), z_search=bool(rp.z_search),
                                         level=bool(rp.level),
                                         training_guesses="registered +-4 m / +-10 deg in 2 deg steps, then as above",
                                         loss_weights=REGISTERED_LOSS_WEIGHTS)
        log(f"proposer trained: {tinfo}")
        policy = A.LEARNED_RACE if args.policy == "race" else A.LEARNED_FALLBACK
        # the learned step's launch shape: the batch's, with fine chunks from pass
        # learned_fine_pass of a run on (its runs start near the optimum and are short)
        prm_l = A.NdtParams.from_buffer_copy(prm)
        prm_l.fine_pass = args.learned_fine_pass
        if not args.full_pose:
            policy |= A.LEARNED_KEEP_ROTATION
        for _ in range(args.warmup):
            model.align(grid, bm, scans, w.guesses, prm_l, policy)
        # the timed step: scan and map grids, the MLP, the alignment launch, results
        align_ms_l, step_ms_l = [], []
        t_l = time.perf_counter()
        for _ in range(args.steps):
            res_l, used = model.align(grid, bm, scans, w.guesses, prm_l, policy)
            ms_l, _, pairs_l = ctx.last_timing()
            align_ms_l.append(ms_l)
            step_ms_l.append(ctx.last_step_ms())
        learned_s = (time.perf_counter() - t_l) / args.steps
        learned_ms = float(np.mean(step_ms_l))
        # the same step with the scans uploaded from host memory inside it (packing,
        # Morton order and the copy of every scan)
        upload_s = float("nan")
        if not args.only_headline:
            t_u = time.perf_counter()
            for _ in range(args.steps):
                fresh = A.Scans(ctx, w.scans)
                model.align(grid, bm, fresh, w.guesses, prm_l, policy)
                fresh.close()
            upload_s = (time.perf_counter() - t_u) / args.steps
        refined = model.predict(bm, scans, w.guesses)
        if not args.full_pose:
            # the translation-only start keeps the rotation of the guess the model
            # corrects: the registered one with a registration
            base = w.guesses if rp is None else bm.register(scans, np.arange(S, dtype=np.int32), w.guesses, rp)[0]
            refined[:, :3, :3] = base[:, :3, :3]
        policy_name = args.policy + ("" if args.full_pose else ", translation from the model") + \
            ("" if args.no_register else ", BEV registration before the model")
        learned = dict(model=model, bm=bm, prm=prm_l, res=res_l, used=used, step_s=learned_s, kernel_ms=learned_ms,
                       align_ms=float(np.mean(align_ms_l)), pairs=pairs_l, refined=refined, train=tinfo,
                       policy=policy_name, upload_s=upload_s, policy_flags=policy)
        log(f"batched+learned {S / learned_s:.0f}/s ({S / upload_s:.0f}/s with the scan upload in the step)")

    # ---- plain single-scan GPU NDT on the same scans: one submission and host sync
    # per scan, with the chunk size that minimises single-scan latency (256 points)
    n_plain = max(1, min(args.plain_scans, S))
    prm_plain = A.params(A.NDT_MODE_NDT, point_precision=point_prec, origin=w.origin, points_per_wg=256)
    plain_scans = [A.Scans(ctx, [w.scans[i]]) for i in range(n_plain)]
    for i in range(n_plain):  # warm-up: each handle captures its submission once
        plain_scans[i].align(grid, w.guesses[i:i + 1], prm_plain)
    plain_res = []
    t1 = time.perf_counter()
    for i in range(n_plain):
        plain_res += plain_scans[i].align(grid, w.guesses[i:i + 1], prm_plain)
    plain_rate = n_plain / (time.perf_counter() - t1)
    plain_vs_batched = np.array([np.abs(a["final_f64"][:3, 3] - b["final_f64"][:3, 3]).max()
                                 for a, b in zip(plain_res, res_b[:n_plain])])

    def success(R):
        return float(np.mean([np.linalg.norm(R[i]["final_f64"][:3, 3] - w.gt[i][:3, 3]) < 0.05 and
                              abs(np.arctan2(-R[i]["final_f64"][0, 1], R[i]["final_f64"][0, 0]) -
                                  np.arctan2(-w.gt[i][0, 1], w.gt[i][0, 0])) < np.deg2rad(0.25)
                              for i in range(len(R))]))

    peak_flops = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    avg_ms = float(np.mean(kernel_ms))

    def roofline(results, pairs, ms, kernel, traffic_kind):
        bytes_alg, flops_alg = algorithmic_work(results, counts, pairs)
        achieved = bytes_alg / (ms * 1e-3) / 1e9
        traffic = profiled_traffic(args, learned["prm"] if traffic_kind == "learned" else prm, traffic_kind)
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": TRAFFIC_PROFILE if traffic else None,
                "algorithmic_bytes": bytes_alg, "kernel_ms": ms,
                "valu_frac": flops_alg / (ms * 1e-3) / (peak_flops * 1e12), "valu_peak_tflops": peak_flops,
                "valu_frac_of_fp32_peak": flops_alg / (ms * 1e-3) / (FP32_PEAK_TFLOPS * 1e12), "kernel": kernel}

    roof_b = roofline(res_b, pairs_b, avg_ms, "ndt::k_align_wq (work-queue align launch of the 512-scan batch, "
                                              "no proposer)", "plain")
    roof_l = roofline(learned["res"], learned["pairs"], learned["align_ms"],
                      "ndt::k_align_wq (the headline step's launch: probe pass, refined or given-guess run per "
                      "scan)", "learned") if learned else None
    final = learned["res"] if learned else res_b
    value = S / (learned["step_s"] if learned else batched_s)
    ws = w.meta

    out = {
        "metric": "scan alignments/sec at <=1 cm/0.05 deg vs FP64: batched+learned vs plain NDT",
        "value": value,
        "unit": "alignments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": S / value * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded scene generator, seed 1); proposer trained on synthetic scans of the same map",
        "config": {
            "workload": f"{args.workload}: {w.map_xyzw.shape[0]}-point map of a {ws['extent']:.0f} m scene, "
                        f"{S} scans of {min(counts)}-{max(counts)} points (log-uniform), guesses "
                        f"+-{ws['perturb_xyz']} m / +-{ws['perturb_yaw_deg']} deg yaw / "
                        f"+-{ws['perturb_rollpitch_deg']} deg roll-pitch; "
                        + (f"batched + learned warm start ({learned['policy']})" if learned else "batched, no proposer"),
            "scans_per_step": S,
            "scan_points_total": int(sum(counts)),
            "map_points": int(w.map_xyzw.shape[0]),
            "voxel": 1.0, "step": 0.1, "epsilon": 0.01, "max_iterations": 30, "outlier_ratio": 0.55,
        },
        "plain_single_scan": {"value": plain_rate, "unit": "alignments/s", "scans": n_plain, "points_per_wg": 256,
                              "within_1mm_of_batched": int(np.sum(plain_vs_batched <= 1e-3)),
                              "converged": float(np.mean([r["converged"] for r in plain_res]))},
        "batched_no_proposer": {"value": S / batched_s, "unit": "alignments/s", "kernel_ms": avg_ms,
                                "launches": launches},
        "vs_plain_single_scan": value / plain_rate,
        "converged_fraction": float(np.mean([r["converged"] for r in final])),
        "converged_fraction_batched_plain": float(np.mean([r["converged"] for r in res_b])),
        "success_vs_truth": {"criterion": "within 5 cm and 0.25 deg yaw of the generator's true pose",
                             "plain": success(res_b), "chosen": success(final)},
        "mean_passes": float(np.mean([r["passes"] for r in final])),
        "mean_passes_batched_plain": float(np.mean([r["passes"] for r in res_b])),
        "grid_build_ms": ginfo["build_ms"],
        "launch_shape": {"block_size": prm.block_size or 256, "points_per_wg": prm.points_per_wg or 2048,
                         "fine_pass": prm.fine_pass, "autotuned": tuned_ms is not None,
                         "learned_fine_pass": learned["prm"].fine_pass if learned else None},
        "profile_us_batched_plain": {k: float(v) for k, v in zip(("lists", "compute", "wait", "control"), profile)},
        "roofline": roof_l or roof_b,
        "roofline_batched_no_proposer": roof_b,
        "device": f"{dev_name}, {cus} CUs",
        "generation_s": gen_s,
    }
    if learned:
        tr = learned["train"]
        # training step = forward + backward (3x the forward flops) per sample, over the
        # whole training call (Adam, map crops and transposes included)
        train_tflops = 3.0 * MLP_FWD_FLOPS * tr["samples"] * tr["epochs"] / tr["train_s"] / 1e12
        out["learned"] = dict(tr, policy=learned["policy"],
                              mlp={"forward_flops_per_sample": MLP_FWD_FLOPS, "train_tflops": train_tflops,
                                   "train_frac_of_bf16_peak": train_tflops / BF16_PEAK_TFLOPS,
                                   "bf16_peak_tflops": BF16_PEAK_TFLOPS},
                              used_refined=float(np.mean(learned["used"])),
                              device_ms_per_step=learned["kernel_ms"], align_kernel_ms=learned["align_ms"],
                              with_scan_upload_alignments_per_s=(None if np.isnan(learned["upload_s"])
                                                                 else S / learned["upload_s"]))
        if args.cfg4_scans > 0:
            out["cfg4_held_out"] = cfg4_leg(ctx, w, grid, learned["bm"], learned["model"], prm, learned["prm"],
                                            learned["policy_flags"], args.cfg4_scans, max(1, min(args.steps, 5)))
            log(f"cfg4: {out['cfg4_held_out']['passes_saved_per_scan']:.2f} passes saved per scan")
    # the reference's order of passes (Hessian only at accepted steps): same poses,
    # more passes
    prm_ha = None if args.only_headline else A.params(
        A.NDT_MODE_NDT, point_precision=point_prec, origin=w.origin, block_size=prm.block_size,
        points_per_wg=prm.points_per_wg, hessian_at_accept=1)
    if prm_ha is not None:
        scans.align(grid, w.guesses, prm_ha)
        t_h = time.perf_counter()
        for _ in range(max(1, min(args.steps, 5))):
            res_h = scans.align(grid, w.guesses, prm_ha)
        dt_h = np.array([np.abs(a["final_f64"][:3, 3] - b["final_f64"][:3, 3]).max() for a, b in zip(res_h, res_b)])
        out["batched_hessian_at_accept"] = {
            "value": S * max(1, min(args.steps, 5)) / (time.perf_counter() - t_h), "unit": "alignments/s",
            "mean_passes": float(np.mean([r["passes"] for r in res_h])),
            # the extra passes move the late-pass chunking (pass 24 on), so sums round
            # differently there
            "same_poses_as_default": int(sum(np.array_equal(a["final_f64"], b["final_f64"])
                                             for a, b in zip(res_h, res_b))),
            "within_1mm_of_default": int(np.sum(dt_h <= 1e-3))}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import orc
        og = orc.NdtGrid(w.map_xyzw)
        rng = np.random.default_rng(0)
        idx = np.sort(rng.choice(S, size=min(args.cpu_sample, S), replace=False))
        # the oracle starts where the chosen GPU run started
        starts = np.stack([learned["refined"][i] if (learned and learned["used"][i]) else w.guesses[i] for i in idx])
        local = starts.copy()
        local[:, :3, 3] -= np.array(w.origin)
        threads = os.cpu_count() or 1
        tc = time.perf_counter()
        ref = og.align_many([w.scans[i] for i in idx], local, threads=threads)
        cpu_s = time.perf_counter() - tc
        # single-thread latency on the first 4 of them
        tc1 = time.perf_counter()
        og.align_many([w.scans[i] for i in idx[:4]], local[:4], threads=1)
        single_s = (time.perf_counter() - tc1) / min(4, len(idx))
        shifted = local.copy()
        shifted[:, 0, 3] += 1e-12
        ref2 = og.align_many([w.scans[i] for i in idx], shifted, threads=threads)
        errs = []
        for k, i in enumerate(idx):
            o = ref[k]["final_f64"].copy()
            o[:3, 3] += np.array(w.origin)
            chosen = final[i]
            dt, dr = orc.pose_error(chosen["final_f64"], o)
            stable = orc.pose_error(ref[k]["final_f64"], ref2[k]["final_f64"])[0] < 1e-4
            errs.append((dt, dr, chosen["converged"] == ref[k]["converged"], stable))
        errs = np.array(errs, dtype=np.float64)
        ok = (errs[:, 0] <= 0.01) & (errs[:, 1] <= np.deg2rad(0.05))
        out["cpu_baseline"] = {
            "value": len(idx) / cpu_s, "unit": "alignments/s", "cores": threads, "kind": "port",
            "sample": f"{len(idx)} scans of the same batch ({int(sum(counts[i] for i in idx))} points) from the "
                      f"same starting guesses, float64 oracle (oracle/ndt_oracle.c), one scan per OpenMP thread",
            "single_thread_ms_per_alignment": single_s * 1e3,
        }
        out["accuracy_vs_fp64_oracle"] = {
            "scans": len(idx),
            "within_1cm_0.05deg": int(ok.sum()),
            "within_1mm_1e-4rad": int(np.sum((errs[:, 0] <= 1e-3) & (errs[:, 1] <= 1e-4))),
            "max_translation_m": float(errs[:, 0].max()),
            "max_rotation_rad": float(errs[:, 1].max()),
            "same_converged": int(errs[:, 2].sum()),
            "oracle_reproducible_scans": int(errs[:, 3].sum()),
            "within_1cm_0.05deg_on_reproducible": int(np.sum(ok & (errs[:, 3] > 0))),
        }
        if args.harness_cases > 0:
            try:
                out["daphne_harness"] = daphne_harness_leg(args.harness_cases)
            except Exception as e:  # reported, never fatal to the bench line
                out["daphne_harness"] = {"error": f"{type(e).__name__}: {e}"}
    if args.cfg1_scans > 0:
        out["cfg1_sequential"] = cfg1_leg(ctx, args.cfg1_scans, 0 if args.no_cpu else 16,
                                          0 if args.no_cpu else args.cfg1_harness_cases)
        log(f"cfg1 sequential {out['cfg1_sequential']['value']:.0f}/s")
    # every device handle closed, newest first, before the interpreter and the HIP
    # runtime shut down
    A.close_all()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
