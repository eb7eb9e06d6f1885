"""Precision study of the NDT engine (the UTM-offset batch).

The batched workload with the whole scene offset to UTM-like coordinates
(+4.2e5, +5.6e6, +80 m) is aligned in four arithmetic modes:

  fp64          float64 point math, float64 sums (the reference run)
  fp32/fp32     float32 point math on cell-relative coordinates, float32 sums
  fp32/comp32   float32 point math, Neumaier-compensated float32 sums
  fp32/fp64     float32 point math, float64 sums

In every mode the transform and cell lookup, the covariance inversion, the 6x6
solve and the line search run in float64.  Two levels are reported against the
fp64 run:

  pass      relative error of score, gradient and Hessian of single derivative
            passes at the given guesses (no trajectory effects);
  align     per-scan translation / rotation difference of the final poses,
            convergence, alignments/s.

Prints one JSON line; a table goes to stderr.  Usage:
    python daphne-benchmark_amd/precision_study.py [--scans 512] [--pass-scans 16]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import ndt_amd as A  # noqa: E402

MODES = {
    "fp64": (A.NDT_POINT_FP64, A.NDT_SUM_FP64),
    "fp32/fp32": (A.NDT_POINT_FP32, A.NDT_SUM_FP32),
    "fp32/comp32": (A.NDT_POINT_FP32, A.NDT_SUM_COMP32),
    "fp32/fp64": (A.NDT_POINT_FP32, A.NDT_SUM_FP64),
}


def pose_error(T, R):
    """(translation m, rotation rad) between two 4x4 poses."""
    dt = float(np.linalg.norm(T[:3, 3] - R[:3, 3]))
    c = (np.trace(T[:3, :3].T @ R[:3, :3]) - 1.0) / 2.0
    return dt, float(np.arccos(np.clip(c, -1.0, 1.0)))


def euler_params(M, origin):
    """(x, y, z, roll, pitch, yaw) of a world pose in the map-local frame."""
    R = M[:3, :3]
    return np.array([M[0, 3] - origin[0], M[1, 3] - origin[1], M[2, 3] - origin[2],
                     np.arctan2(-R[1, 2], R[2, 2]), np.arcsin(np.clip(R[0, 2], -1, 1)), np.arctan2(-R[0, 1], R[0, 0])])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--scans", type=int, default=512)
    ap.add_argument("--pass-scans", type=int, default=16)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    w = A.make_workload(args.workload, seed=1, scans=args.scans)
    ctx = A.Context(0)
    S = len(w.scans)
    out = {"workload": f"{args.workload}: {S} scans, origin {list(w.origin)}", "modes": {}}
    grid = A.Grid(ctx, w.map_xyzw, A.params(A.NDT_MODE_NDT, origin=w.origin))
    scans = A.Scans(ctx, w.scans)
    ref = None
    results = {}
    for name, (pp, sp) in MODES.items():
        prm = A.params(A.NDT_MODE_NDT, point_precision=pp, sum_precision=sp, origin=w.origin)
        scans.align(grid, w.guesses, prm)
        t = time.perf_counter()
        for _ in range(args.repeats):
            res = scans.align(grid, w.guesses, prm)
        rate = S * args.repeats / (time.perf_counter() - t)
        results[name] = res
        if name == "fp64":
            ref = res
        errs = np.array([pose_error(r["final_f64"], q["final_f64"]) for r, q in zip(res, ref)])
        # single derivative passes at the guesses (map-local parameters)
        pe = {"score": [], "gradient": [], "hessian": []}
        for i in range(min(args.pass_scans, S)):
            p = euler_params(w.guesses[i], w.origin)
            s1, g1, h1, _ = grid.derivatives(w.scans[i], p, prm=prm)
            prm64 = A.params(A.NDT_MODE_NDT, origin=w.origin)
            s0, g0, h0, _ = grid.derivatives(w.scans[i], p, prm=prm64)
            pe["score"].append(abs(s1 - s0) / max(abs(s0), 1e-300))
            pe["gradient"].append(np.linalg.norm(g1 - g0) / max(np.linalg.norm(g0), 1e-300))
            pe["hessian"].append(np.linalg.norm(h1 - h0) / max(np.linalg.norm(h0), 1e-300))
        out["modes"][name] = {
            "alignments_per_s": rate,
            "converged": float(np.mean([r["converged"] for r in res])),
            "translation_m": {"median": float(np.median(errs[:, 0])), "max": float(errs[:, 0].max())},
            "rotation_rad": {"median": float(np.median(errs[:, 1])), "max": float(errs[:, 1].max())},
            "within_1cm_0.05deg": float(np.mean((errs[:, 0] <= 0.01) & (errs[:, 1] <= np.deg2rad(0.05)))),
            "within_1mm_1e-4rad": float(np.mean((errs[:, 0] <= 1e-3) & (errs[:, 1] <= 1e-4))),
            "pass_relative_error_median": {k: float(np.median(v)) for k, v in pe.items()},
            "pass_relative_error_max": {k: float(np.max(v)) for k, v in pe.items()},
        }
        m = out["modes"][name]
        print(f"{name:12s} {rate:9.0f}/s conv {m['converged']:.3f} within1cm {m['within_1cm_0.05deg']:.3f} "
              f"within1mm {m['within_1mm_1e-4rad']:.3f} dt med {m['translation_m']['median']:.2e} "
              f"max {m['translation_m']['max']:.2e} | pass grad med {m['pass_relative_error_median']['gradient']:.1e} "
              f"hess med {m['pass_relative_error_median']['hessian']:.1e}", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
