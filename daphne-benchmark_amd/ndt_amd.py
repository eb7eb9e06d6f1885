"""Host-side Python interface of the MI355X NDT scan-matching engine.

Mirrors the reference's testcase interface for the ndt_mapping kernel
(src/CPU/ndt_mapping/kernel.cpp): the DAPHNE data files (read_next_testcases
:143-217, parseResult :178-190), ``partial_points_callback`` (:1405-1414) and the
output check (check_next_outputs :1436-1496).  Compute goes through the C ABI of
``libndthip.so`` (include/ndt_hip.h); there is no Python or CPU fallback: a
missing library raises.

Synthetic inputs come from ``libndt_scene.so`` (include/ndt_scene.h).
"""
from __future__ import annotations

import atexit
import ctypes as C
import gzip
import os
import struct
import weakref
from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------------------
# synthetic scenes (include/ndt_scene.h)
# ---------------------------------------------------------------------------


class SceneCfg(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("extent_x", C.c_double),
        ("extent_y", C.c_double),
        ("n_boxes", C.c_int32),
        ("n_cylinders", C.c_int32),
        ("box_side_min", C.c_double),
        ("box_side_max", C.c_double),
        ("box_h_min", C.c_double),
        ("box_h_max", C.c_double),
        ("cyl_radius", C.c_double),
        ("cyl_height", C.c_double),
        ("ground_sigma", C.c_double),
        ("origin", C.c_double * 3),
    ]


class ScanCfg(C.Structure):
    _fields_ = [
        ("n_points", C.c_int64),
        ("max_range", C.c_double),
        ("noise_sigma", C.c_double),
        ("sensor_height", C.c_double),
        ("sensor_margin", C.c_double),
        ("perturb_xyz", C.c_double * 3),
        ("perturb_yaw_deg", C.c_double),
        ("perturb_rollpitch_deg", C.c_double),
    ]


_scene_lib = None


def scene_lib():
    global _scene_lib
    if _scene_lib is None:
        path = os.path.join(PKG_DIR, "libndt_scene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        lib.ndt_scene_cfg_default.argtypes = [C.POINTER(SceneCfg), C.c_uint64]
        lib.ndt_scan_cfg_default.argtypes = [C.POINTER(ScanCfg)]
        lib.ndt_scene_map.argtypes = [C.POINTER(SceneCfg), C.c_int64, C.c_void_p]
        lib.ndt_scene_map.restype = C.c_int
        lib.ndt_scene_scan.argtypes = [C.POINTER(SceneCfg), C.c_void_p, C.c_int64, C.POINTER(ScanCfg),
                                       C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ndt_scene_scan.restype = C.c_int
        lib.ndt_scene_pose.argtypes = [C.c_double] * 6 + [C.c_void_p]
        lib.ndt_scene_scans.argtypes = [C.POINTER(SceneCfg), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        _scene_lib = lib
    return _scene_lib


def scene_cfg(seed: int, extent: float = 100.0, n_boxes: int | None = None, n_cylinders: int | None = None,
              origin=(0.0, 0.0, 0.0)) -> SceneCfg:
    """Scene of ``extent`` x ``extent`` metres.  Object counts default to the
    100 m x 100 m densities (40 boxes, 60 cylinders) scaled by area."""
    cfg = SceneCfg()
    scene_lib().ndt_scene_cfg_default(C.byref(cfg), seed)
    scale = (extent * extent) / (100.0 * 100.0)
    cfg.extent_x = cfg.extent_y = extent
    cfg.n_boxes = int(round(40 * scale)) if n_boxes is None else n_boxes
    cfg.n_cylinders = int(round(60 * scale)) if n_cylinders is None else n_cylinders
    for i in range(3):
        cfg.origin[i] = origin[i]
    return cfg


def scan_cfg(n_points: int = 2000, perturb_xyz: float = 0.5, perturb_yaw_deg: float = 2.0,
             perturb_rollpitch_deg: float = 0.0, max_range: float = 40.0, noise: float = 0.01) -> ScanCfg:
    cfg = ScanCfg()
    scene_lib().ndt_scan_cfg_default(C.byref(cfg))
    cfg.n_points = n_points
    cfg.max_range = max_range
    cfg.noise_sigma = noise
    for i in range(3):
        cfg.perturb_xyz[i] = perturb_xyz
    cfg.perturb_yaw_deg = perturb_yaw_deg
    cfg.perturb_rollpitch_deg = perturb_rollpitch_deg
    return cfg


def make_map(cfg: SceneCfg, m: int) -> np.ndarray:
    out = np.empty((m, 4), dtype=np.float32)
    rc = scene_lib().ndt_scene_map(C.byref(cfg), m, out.ctypes.data)
    if rc:
        raise ValueError(f"ndt_scene_map failed ({rc})")
    return out


def make_scan(scene: SceneCfg, map_xyzw: np.ndarray, cfg: ScanCfg, scan_seed: int):
    """Returns (scan float32 (n,4), gt_pose float64 (4,4), guess_pose float64 (4,4))."""
    map_xyzw = np.ascontiguousarray(map_xyzw, dtype=np.float32)
    scan = np.empty((cfg.n_points, 4), dtype=np.float32)
    gt = np.empty((4, 4), dtype=np.float64)
    guess = np.empty((4, 4), dtype=np.float64)
    rc = scene_lib().ndt_scene_scan(C.byref(scene), map_xyzw.ctypes.data, map_xyzw.shape[0], C.byref(cfg),
                                    scan_seed, scan.ctypes.data, gt.ctypes.data, guess.ctypes.data)
    if rc:
        raise ValueError(f"ndt_scene_scan failed ({rc})")
    return scan, gt, guess


def pose(x, y, z, roll, pitch, yaw) -> np.ndarray:
    out = np.empty((4, 4), dtype=np.float64)
    scene_lib().ndt_scene_pose(x, y, z, roll, pitch, yaw, out.ctypes.data)
    return out


# Workloads named in the benchmark configuration (sizes, perturbations, seeds).
WORKLOADS = {
    # plain single-scan NDT, CPU-runnable: 100 m scene, 100 k map, 2 k scan
    "cfg0": dict(extent=100.0, map_points=100_000, scans=1, scan_points=(2000, 2000),
                 perturb_xyz=0.5, perturb_yaw_deg=2.0, perturb_rollpitch_deg=0.0),
    # plain NDT at real scale: 300 m scene, 1 M map, 1000 sequential 20 k scans
    "cfg1": dict(extent=300.0, map_points=1_000_000, scans=1000, scan_points=(20000, 20000),
                 perturb_xyz=0.5, perturb_yaw_deg=2.0, perturb_rollpitch_deg=0.0),
    # batched: 512 scans, point counts log-uniform in [1 k, 60 k]
    "cfg2": dict(extent=300.0, map_points=1_000_000, scans=512, scan_points=(1000, 60000),
                 perturb_xyz=1.0, perturb_yaw_deg=3.0, perturb_rollpitch_deg=0.5),
    # the batch of cfg2 with UTM-sized world offsets
    "cfg3": dict(extent=300.0, map_points=1_000_000, scans=512, scan_points=(1000, 60000),
                 perturb_xyz=1.0, perturb_yaw_deg=3.0, perturb_rollpitch_deg=0.5,
                 origin=(4.2e5, 5.6e6, 80.0)),
}


@dataclass
class Workload:
    name: str
    map_xyzw: np.ndarray
    scans: list
    gt: np.ndarray        # (S, 4, 4) float64
    guesses: np.ndarray   # (S, 4, 4) float64
    origin: tuple = (0.0, 0.0, 0.0)
    scene: SceneCfg | None = None
    meta: dict = field(default_factory=dict)


def scan_sizes(lo: int, hi: int, count: int, seed: int) -> np.ndarray:
    if lo == hi:
        return np.full(count, lo, dtype=np.int64)
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def make_workload(name: str, seed: int = 1, scans: int | None = None, map_points: int | None = None,
                  scan_points=None) -> Workload:
    """Builds a named workload (optionally shrunk: fewer scans / points) from seeds."""
    w = dict(WORKLOADS[name])
    if scans is not None:
        w["scans"] = scans
    if map_points is not None:
        w["map_points"] = map_points
    if scan_points is not None:
        w["scan_points"] = scan_points
    origin = w.get("origin", (0.0, 0.0, 0.0))
    scene = scene_cfg(seed, extent=w["extent"], origin=origin)
    mp = make_map(scene, w["map_points"])
    sizes = scan_sizes(w["scan_points"][0], w["scan_points"][1], w["scans"], seed + 7)
    out_scans, gts, guesses = [], [], []
    for s in range(w["scans"]):
        cfg = scan_cfg(int(sizes[s]), w["perturb_xyz"], w["perturb_yaw_deg"], w["perturb_rollpitch_deg"])
        sc, gt, gu = make_scan(scene, mp, cfg, s)
        out_scans.append(sc)
        gts.append(gt)
        guesses.append(gu)
    return Workload(name, mp, out_scans, np.stack(gts), np.stack(guesses), tuple(origin), scene,
                    dict(seed=seed, **w))


def make_scans(scene: SceneCfg, map_xyzw: np.ndarray, sizes, seeds, perturb_xyz=1.0, perturb_yaw_deg=3.0,
               perturb_rollpitch_deg=0.5, threads: int | None = None):
    """Scans of the given sizes and seeds, drawn on host threads; each equals
    make_scan with the same arguments.  Returns (scans, gt (S,4,4), guesses (S,4,4))."""
    n = len(sizes)
    cfgs = (ScanCfg * n)()
    for k in range(n):
        cfgs[k] = scan_cfg(int(sizes[k]), perturb_xyz, perturb_yaw_deg, perturb_rollpitch_deg)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    counts = np.asarray(sizes, dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    out = np.zeros((int(counts.sum()), 4), np.float32)
    gt = np.zeros((n, 4, 4))
    gs = np.zeros((n, 4, 4))
    status = np.zeros(n, np.int32)
    m = np.ascontiguousarray(map_xyzw, dtype=np.float32)
    rc = scene_lib().ndt_scene_scans(C.byref(scene), m.ctypes.data, m.shape[0], cfgs, seeds.ctypes.data, n,
                                     offsets.ctypes.data, out.ctypes.data, gt.ctypes.data, gs.ctypes.data,
                                     threads or 0, status.ctypes.data)
    if rc != 0 or np.any(status != 0):
        raise RuntimeError(f"ndt_scene_scans failed ({rc}, statuses {sorted(set(status.tolist()))})")
    scans = [out[o:o + c] for o, c in zip(offsets, counts)]
    return scans, gt, gs


def perturb_poses(gt, rng, xy=3.0, z=0.3, yaw_deg=10.0, rollpitch_deg=0.5):
    """Guesses around ground-truth poses: world translation U[-xy, xy] in x and y,
    U[-z, z] in z, yaw U[-yaw, yaw] added to the pose's yaw, roll and pitch
    U[-rollpitch, rollpitch] (the generator's convention, R = Rx Ry Rz)."""
    gt = np.asarray(gt, np.float64).reshape(-1, 4, 4)
    out = np.empty_like(gt)
    d = np.pi / 180.0
    for i, G in enumerate(gt):
        yaw = np.arctan2(-G[0, 1], G[0, 0])
        e = rng.uniform(-1, 1, 6)
        out[i] = pose(G[0, 3] + xy * e[0], G[1, 3] + xy * e[1], G[2, 3] + z * e[2], rollpitch_deg * d * e[3],
                      rollpitch_deg * d * e[4], yaw + yaw_deg * d * e[5])
    return out


# ---------------------------------------------------------------------------
# DAPHNE data files (src/CPU/ndt_mapping/kernel.cpp:143-217, :178-190)
# ---------------------------------------------------------------------------


@dataclass
class Testcase:
    guess: np.ndarray   # (4, 4) float32
    scan: np.ndarray    # (n, 4) float32
    map: np.ndarray     # (m, 4) float32


@dataclass
class CallbackResult:
    """Same fields as the reference's CallbackResult (datatypes.h:51-55)."""
    converged: bool
    final_transformation: np.ndarray  # (4, 4) float32
    fitness_score: float = 0.0


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def write_ndt_input(path, cases) -> None:
    with _open(path, "wb") as f:
        f.write(struct.pack("<i", len(cases)))
        for c in cases:
            f.write(np.ascontiguousarray(c.guess, dtype="<f4").tobytes())
            for cloud in (c.scan, c.map):
                cloud = np.ascontiguousarray(cloud, dtype="<f4").reshape(-1, 4)
                f.write(struct.pack("<i", cloud.shape[0]))
                f.write(cloud.tobytes())


class DataFormatError(ValueError):
    """A DAPHNE data file that ends inside a record (or holds a negative point
    count).  The message is the reference's for that record
    (src/CPU/ndt_mapping/kernel.cpp:157-226)."""


def _unpack(fmt, data, pos, what):
    if pos + struct.calcsize(fmt) > len(data):
        raise DataFormatError(what)
    return struct.unpack_from(fmt, data, pos)


def _floats(data, count, pos, what):
    if count < 0 or pos + 4 * count > len(data):
        raise DataFormatError(what)
    return np.frombuffer(data, "<f4", count, pos).copy()


def read_ndt_input(path) -> list:
    with _open(path, "rb") as f:
        data = f.read()
    (t,) = _unpack("<i", data, 0, "Error reading number of test cases")
    pos, cases = 4, []
    for _ in range(t):
        guess = _floats(data, 16, pos, "Error reading initial guess").reshape(4, 4)
        pos += 64
        clouds = []
        for _k in range(2):
            (n,) = _unpack("<i", data, pos, "Error reading filtered scan")
            pos += 4
            clouds.append(_floats(data, 4 * n, pos, "Error reading filtered scan").reshape(n, 4))
            pos += 16 * n
        cases.append(Testcase(guess, clouds[0], clouds[1]))
    return cases


def write_ndt_output(path, results) -> None:
    with _open(path, "wb") as f:
        for r in results:
            f.write(np.ascontiguousarray(r.final_transformation, dtype="<f4").tobytes())
            f.write(struct.pack("<d", r.fitness_score))
            f.write(struct.pack("<B", 1 if r.converged else 0))


def read_ndt_output(path) -> list:
    with _open(path, "rb") as f:
        data = f.read()
    if len(data) % 73:
        raise DataFormatError("Error reading reference result")
    out = []
    for i in range(len(data) // 73):
        rec = data[73 * i: 73 * (i + 1)]
        m = np.frombuffer(rec, "<f4", 16, 0).reshape(4, 4).copy()
        (fit,) = struct.unpack_from("<d", rec, 64)
        out.append(CallbackResult(bool(rec[72]), m, fit))
    return out


MAX_TRANSLATION_EPS = 0.001   # kernel.cpp:18
MAX_EPS = 2.0                 # kernel.cpp:20


def check_outputs(results, references):
    """The reference's acceptance test (check_next_outputs, kernel.cpp:1436-1496).
    Returns (ok, max_delta)."""
    error = False
    max_delta = 0.0
    origin = np.array([0.724, 0.447, 0.525, 1.0], dtype=np.float32)
    for res, ref in zip(results, references):
        if bool(res.converged) != bool(ref.converged):
            error = True
        a = np.asarray(res.final_transformation, dtype=np.float32)
        b = np.asarray(ref.final_transformation, dtype=np.float32)
        if np.any(np.isnan(a) != np.isnan(b)):
            error = True
        for h in range(4):
            delta = float(np.float32(abs(np.float32(a[h, 3]) - np.float32(b[h, 3]))))
            if delta > max_delta:
                max_delta = delta
                if delta > MAX_TRANSLATION_EPS:
                    error = True
        rp = np.zeros(4, dtype=np.float32)
        fp = np.zeros(4, dtype=np.float32)
        for h in range(4):
            for w in range(4):
                rp[h] = np.float32(rp[h] + np.float32(a[h, w] * origin[w]))
                fp[h] = np.float32(fp[h] + np.float32(b[h, w] * origin[w]))
        for w in range(4):
            delta = float(abs(rp[w] - fp[w]))
            if delta > max_delta:
                max_delta = delta
                if delta > MAX_EPS:
                    error = True
    return (not error), max_delta


# ---------------------------------------------------------------------------
# C ABI of libndthip.so (include/ndt_hip.h)
# ---------------------------------------------------------------------------

NDT_MODE_COMPAT, NDT_MODE_NDT = 0, 1
NDT_POINT_FP32, NDT_POINT_FP64 = 0, 1
NDT_SUM_FP32, NDT_SUM_COMP32, NDT_SUM_FP64 = 0, 1, 2
NDT_SCHED_DEFAULT, NDT_SCHED_ROUNDS, NDT_SCHED_QUEUE = 0, 1, 2


class NdtParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("resolution", C.c_float), ("step_size", C.c_double), ("trans_eps", C.c_double),
                ("max_iterations", C.c_int32), ("outlier_ratio", C.c_double), ("min_points", C.c_int32),
                ("min_eig_ratio", C.c_double), ("point_precision", C.c_int32), ("sum_precision", C.c_int32),
                ("origin", C.c_double * 3), ("block_size", C.c_int32), ("points_per_wg", C.c_int32),
                ("scheduler", C.c_int32), ("hessian_at_accept", C.c_int32), ("fine_pass", C.c_int32)]


class NdtResult(C.Structure):
    _fields_ = [("final_transformation", C.c_float * 16), ("final_transformation_f64", C.c_double * 16),
                ("fitness_score", C.c_double), ("score", C.c_double), ("trans_probability", C.c_double),
                ("iterations", C.c_int32), ("converged", C.c_int32), ("passes", C.c_int32), ("status", C.c_int32)]


class NdtGridInfo(C.Structure):
    _fields_ = [("map_points", C.c_int64), ("voxels", C.c_int64), ("valid_voxels", C.c_int64),
                ("query_cells", C.c_int64), ("candidates", C.c_int64), ("hash_capacity", C.c_int64),
                ("dense_cells", C.c_int64), ("dims", C.c_int32 * 3), ("min_corner", C.c_float * 3),
                ("max_corner", C.c_float * 3), ("build_ms", C.c_double)]


class TrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64), ("weights", C.c_float * 6),
                ("weight_decay", C.c_float), ("cosine", C.c_int32)]


class RegisterParams(C.Structure):
    _fields_ = [("radius", C.c_int32), ("yaws", C.c_int32), ("yaw_step", C.c_double), ("refine_yaw", C.c_int32),
                ("z_search", C.c_int32), ("level", C.c_int32), ("pad", C.c_int32)]


class NdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ndt error {code}: {msg}")
        self.code = code


_hip_lib = None
HIP_LIB_PATH = os.path.join(PKG_DIR, "libndthip.so")


def hip_lib():
    """Loads libndthip.so; raises when it is missing (no fallback path exists)."""
    global _hip_lib
    if _hip_lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError(f"{HIP_LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(HIP_LIB_PATH)
        P = C.c_void_p
        L.ndt_params_default.argtypes = [C.POINTER(NdtParams), C.c_int32]
        L.ndt_last_error.restype = C.c_char_p
        L.ndt_version.restype = C.c_char_p
        L.ndt_ctx_create.argtypes = [C.c_int32, C.POINTER(P)]
        L.ndt_ctx_destroy.argtypes = [P]
        L.ndt_ctx_device_info.argtypes = [P, C.POINTER(C.c_int32), C.c_char_p, C.c_int32]
        L.ndt_grid_build.argtypes = [P, P, C.c_int64, C.POINTER(NdtParams), C.POINTER(P)]
        L.ndt_grid_info_get.argtypes = [P, C.POINTER(NdtGridInfo)]
        L.ndt_grid_destroy.argtypes = [P]
        L.ndt_grid_voxels.argtypes = [P, C.c_int64, P, P, P, P, P]
        L.ndt_align.argtypes = [P, P, P, C.c_int64, P, C.POINTER(NdtParams), C.POINTER(NdtResult)]
        L.ndt_align_batch.argtypes = [P, P, P, P, C.c_int32, P, C.POINTER(NdtParams), P]
        L.ndt_scans_upload.argtypes = [P, P, P, C.c_int32, C.POINTER(P)]
        L.ndt_scans_destroy.argtypes = [P]
        L.ndt_align_scans.argtypes = [P, P, P, P, C.POINTER(NdtParams), P]
        L.ndt_align_scans_noalloc.argtypes = [P, P, P, P, C.POINTER(NdtParams), P]
        L.ndt_compat_callback.argtypes = [P, P, C.c_int64, P, P, C.c_int64, C.POINTER(NdtResult)]
        L.ndt_derivatives.argtypes = [P, P, P, C.c_int64, P, P, C.POINTER(NdtParams), P, P, P, P]
        L.ndt_ctx_last_timing.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.ndt_ctx_last_profile.argtypes = [P, P]
        L.ndt_ctx_last_step_time.argtypes = [P, C.POINTER(C.c_double)]
        L.ndt_fitness_score.argtypes = [P, P, P, P, C.c_double, P, P]
        L.ndt_autotune.argtypes = [P, P, P, P, C.POINTER(NdtParams), C.c_char_p, C.POINTER(C.c_double)]
        # learned warm start (ndt_learn.h)
        L.ndt_bevmap_build.argtypes = [P, P, C.c_int64, P, C.POINTER(P)]
        L.ndt_bevmap_destroy.argtypes = [P]
        L.ndt_features.argtypes = [P, P, P, P, P, C.c_int32, P]
        L.ndt_model_create.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(P)]
        L.ndt_model_destroy.argtypes = [P]
        L.ndt_model_dims.argtypes = [P, P]
        L.ndt_model_save.argtypes = [P, C.c_char_p]
        L.ndt_model_load.argtypes = [P, C.c_char_p, C.POINTER(P)]
        L.ndt_model_get.argtypes = [P, C.c_int32, P, C.c_int64]
        L.ndt_model_set.argtypes = [P, C.c_int32, P, C.c_int64]
        L.ndt_model_forward.argtypes = [P, P, P, C.c_int32, P]
        L.ndt_model_backward.argtypes = [P, P, P, P, P, C.c_int32, C.POINTER(C.c_float)]
        L.ndt_model_grad.argtypes = [P, C.c_int32, P, C.c_int64]
        L.ndt_model_adam_step.argtypes = [P, P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
        L.ndt_train_params_default.argtypes = [C.POINTER(TrainParams)]
        L.ndt_train_params_default.restype = None
        L.ndt_dataset_create.argtypes = [P, P, P, P, P, P, C.c_int64, C.POINTER(P)]
        L.ndt_dataset_destroy.argtypes = [P]
        L.ndt_dataset_add.argtypes = [P, P, P, P, P, C.c_int64]
        L.ndt_dataset_size.argtypes = [P, P, P]
        L.ndt_model_train.argtypes = [P, P, P, C.POINTER(TrainParams), P]
        L.ndt_model_predict.argtypes = [P, P, P, P, P, P]
        L.ndt_align_scans_learned.argtypes = [P, P, P, P, P, P, C.POINTER(NdtParams), C.c_int32, P, P]
        L.ndt_register_params_default.argtypes = [C.POINTER(RegisterParams)]
        L.ndt_register_params_default.restype = None
        L.ndt_bev_register.argtypes = [P, P, P, P, P, C.c_int32, C.POINTER(RegisterParams), P, P]
        L.ndt_model_set_register.argtypes = [P, C.POINTER(RegisterParams)]
        L.ndt_model_get_register.argtypes = [P, C.POINTER(RegisterParams)]
        _hip_lib = L
    return _hip_lib


EXPORTED_SYMBOLS = ("ndt_params_default", "ndt_last_error", "ndt_version", "ndt_ctx_create", "ndt_ctx_destroy",
                    "ndt_ctx_device_info", "ndt_grid_build", "ndt_grid_info_get", "ndt_grid_destroy",
                    "ndt_grid_voxels", "ndt_align", "ndt_align_batch", "ndt_scans_upload", "ndt_scans_destroy",
                    "ndt_align_scans", "ndt_compat_callback", "ndt_derivatives", "ndt_ctx_last_timing",
                    "ndt_ctx_last_profile", "ndt_fitness_score", "ndt_autotune", "ndt_align_scans_noalloc")
LEARN_SYMBOLS = ("ndt_bevmap_build", "ndt_bevmap_destroy", "ndt_features", "ndt_model_create", "ndt_model_destroy",
                 "ndt_model_dims", "ndt_model_save", "ndt_model_load", "ndt_model_get", "ndt_model_set",
                 "ndt_model_forward", "ndt_model_backward", "ndt_model_grad", "ndt_model_adam_step",
                 "ndt_train_params_default", "ndt_dataset_create", "ndt_dataset_destroy", "ndt_dataset_add",
                 "ndt_dataset_size", "ndt_model_train",
                 "ndt_model_predict", "ndt_align_scans_learned", "ndt_ctx_last_step_time",
                 "ndt_register_params_default", "ndt_bev_register", "ndt_model_set_register",
                 "ndt_model_get_register")


def _check(rc):
    if rc != 0:
        raise NdtError(rc, hip_lib().ndt_last_error().decode())


def params(mode=NDT_MODE_NDT, **kw) -> NdtParams:
    p = NdtParams()
    hip_lib().ndt_params_default(C.byref(p), mode)
    for k, v in kw.items():
        if k == "origin":
            for i in range(3):
                p.origin[i] = v[i]
        else:
            setattr(p, k, v)
    return p


_RESULT_DTYPE = np.dtype({"names": ["final_transformation", "final_transformation_f64", "fitness_score", "score",
                                     "trans_probability", "iterations", "converged", "passes", "status"],
                           "formats": [("<f4", (16,)), ("<f8", (16,)), "<f8", "<f8", "<f8", "<i4", "<i4", "<i4", "<i4"],
                           "offsets": [NdtResult.final_transformation.offset, NdtResult.final_transformation_f64.offset,
                                       NdtResult.fitness_score.offset, NdtResult.score.offset,
                                       NdtResult.trans_probability.offset, NdtResult.iterations.offset,
                                       NdtResult.converged.offset, NdtResult.passes.offset, NdtResult.status.offset],
                           "itemsize": C.sizeof(NdtResult)})


class Results(Sequence):
    """Alignment results of a batch, one dict per scan (final_f32, final_f64,
    score, trans_probability, iterations, converged, passes, status), made from
    the returned records (in one vectorised pass) when first read: a caller that
    reads none of them does not pay for the conversion."""

    __slots__ = ("_a", "_d")

    def __init__(self, out, n: int):
        self._a = np.frombuffer(out, dtype=_RESULT_DTYPE, count=n).copy()
        self._d = None

    def _dicts(self) -> list:
        if self._d is None:
            a, n = self._a, len(self._a)
            f32 = a["final_transformation"].reshape(n, 4, 4).copy()
            f64 = a["final_transformation_f64"].reshape(n, 4, 4).copy()
            score, tp = a["score"].tolist(), a["trans_probability"].tolist()
            it, cv, ps, st = (a["iterations"].tolist(), a["converged"].tolist(), a["passes"].tolist(),
                              a["status"].tolist())
            self._d = [dict(final_f32=f32[i], final_f64=f64[i], score=score[i], trans_probability=tp[i],
                            iterations=it[i], converged=bool(cv[i]), passes=ps[i], status=st[i]) for i in range(n)]
        return self._d

    def __len__(self) -> int:
        return len(self._a)

    def __getitem__(self, i):
        return self._dicts()[i]

    def __iter__(self):
        return iter(self._dicts())

    def field(self, name: str) -> np.ndarray:
        """One record field for every scan as an array (for example "passes")."""
        return self._a[name].copy()


def _results(out, n: int) -> Results:
    """Results of the first n entries of a ctypes NdtResult array."""
    return Results(out, n)


def _result(r: NdtResult) -> dict:
    return dict(final_f32=np.ctypeslib.as_array(r.final_transformation).reshape(4, 4).copy(),
                final_f64=np.ctypeslib.as_array(r.final_transformation_f64).reshape(4, 4).copy(),
                score=r.score, trans_probability=r.trans_probability, iterations=r.iterations,
                converged=bool(r.converged), passes=r.passes, status=r.status)


# Handles still open, in creation order (weak references).  close_all() closes them
# newest first; it runs at interpreter exit, so that no device memory, stream or
# captured graph is left for the HIP runtime's own teardown (or a profiler attached
# to it) to release after the library's users are gone.
_live: list = []


def _track(obj) -> None:
    if len(_live) >= 4096:
        _live[:] = [r for r in _live if r() is not None]
    _live.append(weakref.ref(obj))


@atexit.register
def close_all() -> None:
    """Closes every open handle of this module, newest first (handles first, then
    the contexts they were made on)."""
    while _live:
        obj = _live.pop()()
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


class Context:
    """One HIP device context (one stream)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        _check(hip_lib().ndt_ctx_create(device, C.byref(self.h)))
        _track(self)

    def close(self):
        if self.h:
            hip_lib().ndt_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        cus = C.c_int32()
        name = C.create_string_buffer(256)
        _check(hip_lib().ndt_ctx_device_info(self.h, C.byref(cus), name, 256))
        return cus.value, name.value.decode()

    def last_timing(self):
        ms, launches, pairs = C.c_double(), C.c_int32(), C.c_int64()
        _check(hip_lib().ndt_ctx_last_timing(self.h, C.byref(ms), C.byref(launches), C.byref(pairs)))
        return ms.value, launches.value, pairs.value

    def last_step_ms(self) -> float:
        """Device milliseconds of the whole last learned call (features, model, NDT)."""
        ms = C.c_double()
        _check(hip_lib().ndt_ctx_last_step_time(self.h, C.byref(ms)))
        return ms.value

    def last_profile(self):
        """(prologue, compute, barrier, control) microseconds of the last ndt alignment."""
        out = np.zeros(4)
        _check(hip_lib().ndt_ctx_last_profile(self.h, out.ctypes.data))
        return out

    def compat_callback(self, scan, guess, map_xyzw) -> dict:
        """partial_points_callback with the reference's semantics."""
        scan = np.ascontiguousarray(scan, dtype=np.float32)
        map_xyzw = np.ascontiguousarray(map_xyzw, dtype=np.float32)
        guess = np.ascontiguousarray(guess, dtype=np.float32)
        r = NdtResult()
        _check(hip_lib().ndt_compat_callback(self.h, scan.ctypes.data, scan.shape[0], guess.ctypes.data,
                                             map_xyzw.ctypes.data, map_xyzw.shape[0], C.byref(r)))
        return _result(r)


class Grid:
    def __init__(self, ctx: Context, map_xyzw, prm: NdtParams):
        self.ctx = ctx
        self.params = prm
        m = np.ascontiguousarray(map_xyzw, dtype=np.float32)
        self.h = C.c_void_p()
        _check(hip_lib().ndt_grid_build(ctx.h, m.ctypes.data, m.shape[0], C.byref(prm), C.byref(self.h)))
        _track(self)

    def close(self):
        if self.h:
            hip_lib().ndt_grid_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = NdtGridInfo()
        _check(hip_lib().ndt_grid_info_get(self.h, C.byref(i)))
        return {k: (list(getattr(i, k)) if k in ("dims", "min_corner", "max_corner") else getattr(i, k))
                for k, _ in NdtGridInfo._fields_}

    def voxels(self) -> dict:
        k = self.info()["voxels"]
        keys = np.zeros((k, 3), np.int64)
        n = np.zeros(k, np.int32)
        mean = np.zeros((k, 3))
        icov = np.zeros((k, 9))
        valid = np.zeros(k, np.int32)
        _check(hip_lib().ndt_grid_voxels(self.h, k, keys.ctypes.data, n.ctypes.data, mean.ctypes.data,
                                         icov.ctypes.data, valid.ctypes.data))
        return dict(key=keys, n=n, mean=mean, icov=icov, valid=valid)

    def derivatives(self, scan, p, compat_transform=None, prm: NdtParams | None = None):
        prm = prm or self.params
        scan = np.ascontiguousarray(scan, dtype=np.float32)
        p = np.ascontiguousarray(p, dtype=np.float64)
        M = None if compat_transform is None else np.ascontiguousarray(compat_transform, dtype=np.float32)
        score = C.c_double()
        g = np.zeros(6)
        H = np.zeros(36)
        pairs = C.c_int64()
        _check(hip_lib().ndt_derivatives(self.ctx.h, self.h, scan.ctypes.data, scan.shape[0], p.ctypes.data,
                                         None if M is None else M.ctypes.data, C.byref(prm), C.byref(score),
                                         g.ctypes.data, H.ctypes.data, C.byref(pairs)))
        return score.value, g, H.reshape(6, 6), pairs.value

    def align(self, scan, guess, prm: NdtParams | None = None) -> dict:
        prm = prm or self.params
        scan = np.ascontiguousarray(scan, dtype=np.float32)
        guess = np.ascontiguousarray(guess, dtype=np.float64)
        r = NdtResult()
        _check(hip_lib().ndt_align(self.ctx.h, self.h, scan.ctypes.data, scan.shape[0], guess.ctypes.data,
                                   C.byref(prm), C.byref(r)))
        return _result(r)

    def align_batch(self, scans, guesses, prm: NdtParams | None = None) -> list:
        prm = prm or self.params
        scans = [np.ascontiguousarray(s, dtype=np.float32) for s in scans]
        S = len(scans)
        ptrs = (C.c_void_p * S)(*[s.ctypes.data for s in scans])
        counts = np.array([s.shape[0] for s in scans], dtype=np.int64)
        guesses = np.ascontiguousarray(guesses, dtype=np.float64).reshape(S, 16)
        out = (NdtResult * S)()
        _check(hip_lib().ndt_align_batch(self.ctx.h, self.h, ptrs, counts.ctypes.data, S, guesses.ctypes.data,
                                         C.byref(prm), out))
        return _results(out, S)


class Scans:
    """Scan batch resident in device memory."""

    def __init__(self, ctx: Context, scans):
        self.ctx = ctx
        self.arrays = [np.ascontiguousarray(s, dtype=np.float32) for s in scans]
        S = len(self.arrays)
        ptrs = (C.c_void_p * S)(*[s.ctypes.data for s in self.arrays])
        self.counts = np.array([s.shape[0] for s in self.arrays], dtype=np.int64)
        self.h = C.c_void_p()
        _check(hip_lib().ndt_scans_upload(ctx.h, ptrs, self.counts.ctypes.data, S, C.byref(self.h)))
        self.S = S
        _track(self)

    def close(self):
        if self.h:
            hip_lib().ndt_scans_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def align(self, grid: Grid, guesses, prm: NdtParams | None = None) -> list:
        prm = prm or grid.params
        guesses = np.ascontiguousarray(guesses, dtype=np.float64).reshape(self.S, 16)
        out = (NdtResult * self.S)()
        _check(hip_lib().ndt_align_scans(self.ctx.h, grid.h, self.h, guesses.ctypes.data, C.byref(prm), out))
        return _results(out, self.S)

    def align_noalloc(self, grid: Grid, guesses, prm: NdtParams | None = None) -> list:
        """align() through the allocation-free entry point (fixed device workspace)."""
        prm = prm or grid.params
        guesses = np.ascontiguousarray(guesses, dtype=np.float64).reshape(self.S, 16)
        out = (NdtResult * self.S)()
        _check(hip_lib().ndt_align_scans_noalloc(self.ctx.h, grid.h, self.h, guesses.ctypes.data, C.byref(prm),
                                                 out))
        return _results(out, self.S)

    def autotune(self, grid: Grid, guesses, prm: NdtParams | None = None, cache_path: str | None = None):
        """Fastest (block size, chunk) for this batch; returns (params, kernel ms)."""
        import copy
        prm = copy.copy(prm or grid.params)
        guesses = np.ascontiguousarray(guesses, dtype=np.float64).reshape(self.S, 16)
        ms = C.c_double()
        _check(hip_lib().ndt_autotune(self.ctx.h, grid.h, self.h, guesses.ctypes.data, C.byref(prm),
                                      None if cache_path is None else os.fsencode(cache_path), C.byref(ms)))
        return prm, ms.value

    def fitness(self, grid: Grid, poses, max_range: float = 0.0):
        """(fitness (S,), counts (S,)) at world poses (getFitnessScore)."""
        poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(self.S, 16)
        fit = np.zeros(self.S)
        cnt = np.zeros(self.S, np.int64)
        _check(hip_lib().ndt_fitness_score(self.ctx.h, grid.h, self.h, poses.ctypes.data, max_range,
                                           fit.ctypes.data, cnt.ctypes.data))
        return fit, cnt


def partial_points_callback(ctx: Context, input_cloud, init_guess, target_cloud) -> CallbackResult:
    """The reference's per-testcase callback (kernel.cpp:1405-1414), on the GPU."""
    r = ctx.compat_callback(input_cloud, init_guess, target_cloud)
    return CallbackResult(r["converged"], r["final_f32"], 0.0)


# ---------------------------------------------------------------- learned warm start

BEV_CELLS = 32768
FEATURES = 65536
W1, B1, W2, B2, W3, B3 = range(6)
CORR_SCALE = np.array([3.0, 3.0, 1.0, 0.05, 0.05, 0.2])
LEARNED_RACE, LEARNED_FALLBACK, LEARNED_KEEP_ROTATION = 0, 1, 0x10


def train_params(**kw) -> TrainParams:
    tp = TrainParams()
    hip_lib().ndt_train_params_default(C.byref(tp))
    for k, v in kw.items():
        if k == "weights":
            for i in range(6):
                tp.weights[i] = v[i]
        else:
            setattr(tp, k, v)
    return tp


def register_params(**kw) -> RegisterParams:
    """BEV registration settings (include/ndt_learn.h): the defaults, overridden by kw."""
    rp = RegisterParams()
    hip_lib().ndt_register_params_default(C.byref(rp))
    for k, v in kw.items():
        setattr(rp, k, v)
    return rp


REGISTER_OFF = dict(radius=0)


class _Handle:
    _destroy = None

    def close(self):
        if getattr(self, "h", None):
            getattr(hip_lib(), self._destroy)(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BevMap(_Handle):
    """Map points bucketed in 4 m tiles for the map grids (origin: world offset)."""
    _destroy = "ndt_bevmap_destroy"

    def __init__(self, ctx: Context, map_xyzw, origin=(0.0, 0.0, 0.0)):
        self.ctx = ctx
        m = np.ascontiguousarray(map_xyzw, dtype=np.float32)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        self.h = C.c_void_p()
        _check(hip_lib().ndt_bevmap_build(ctx.h, m.ctypes.data, m.shape[0], o.ctypes.data, C.byref(self.h)))
        _track(self)

    def features(self, scans: "Scans", scan_index, guesses) -> np.ndarray:
        """(B, 65536) uint16 bf16 bit patterns: [scan grid | map grid at the guess]."""
        idx = np.ascontiguousarray(scan_index, dtype=np.int32)
        B = idx.shape[0]
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(B, 16)
        out = np.zeros((B, FEATURES), np.uint16)
        _check(hip_lib().ndt_features(self.ctx.h, self.h, scans.h, idx.ctypes.data, g.ctypes.data, B,
                                      out.ctypes.data))
        return out

    def register(self, scans: "Scans", scan_index, guesses, rp: RegisterParams | None = None):
        """Registered guesses (B, 4, 4) and best scores (B,) of (scan, guess) pairs."""
        rp = rp or register_params()
        idx = np.ascontiguousarray(scan_index, dtype=np.int32)
        B = idx.shape[0]
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(B, 16)
        out = np.zeros_like(g)
        score = np.zeros(B, np.int32)
        _check(hip_lib().ndt_bev_register(self.ctx.h, self.h, scans.h, idx.ctypes.data, g.ctypes.data, B,
                                          C.byref(rp), out.ctypes.data, score.ctypes.data))
        return out.reshape(B, 4, 4), score


class Model(_Handle):
    """bf16 MLP f0 -> f1 -> f2 -> 6 on the GPU (weights float32 masters)."""
    _destroy = "ndt_model_destroy"

    def __init__(self, ctx: Context, f0=FEATURES, f1=1024, f2=256, seed=1, path=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        if path is not None:
            _check(hip_lib().ndt_model_load(ctx.h, os.fsencode(path), C.byref(self.h)))
        else:
            _check(hip_lib().ndt_model_create(ctx.h, f0, f1, f2, seed, C.byref(self.h)))
        _track(self)
        d = np.zeros(3, np.int32)
        _check(hip_lib().ndt_model_dims(self.h, d.ctypes.data))
        self.dims = tuple(int(v) for v in d)

    def shape(self, which):
        f0, f1, f2 = self.dims
        return [(f1, f0), (f1,), (f2, f1), (f2,), (6, f2), (6,)][which]

    def get(self, which) -> np.ndarray:
        out = np.zeros(self.shape(which), np.float32)
        _check(hip_lib().ndt_model_get(self.h, which, out.ctypes.data, out.size))
        return out

    def set(self, which, value):
        v = np.ascontiguousarray(value, dtype=np.float32).reshape(self.shape(which))
        _check(hip_lib().ndt_model_set(self.h, which, v.ctypes.data, v.size))

    def grad(self, which) -> np.ndarray:
        out = np.zeros(self.shape(which), np.float32)
        _check(hip_lib().ndt_model_grad(self.h, which, out.ctypes.data, out.size))
        return out

    def save(self, path):
        _check(hip_lib().ndt_model_save(self.h, os.fsencode(path)))

    def set_register(self, rp: RegisterParams | None):
        """The BEV registration in front of the model (None: none)."""
        rp = rp if rp is not None else register_params(**REGISTER_OFF)
        _check(hip_lib().ndt_model_set_register(self.h, C.byref(rp)))

    def get_register(self) -> RegisterParams:
        rp = RegisterParams()
        _check(hip_lib().ndt_model_get_register(self.h, C.byref(rp)))
        return rp

    def forward(self, x_bf16) -> np.ndarray:
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        B = x.shape[0]
        y = np.zeros((B, 6), np.float32)
        _check(hip_lib().ndt_model_forward(self.ctx.h, self.h, x.ctypes.data, B, y.ctypes.data))
        return y

    def backward(self, x_bf16, target, weights=None) -> float:
        x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        loss = C.c_float()
        _check(hip_lib().ndt_model_backward(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data,
                                            None if w is None else w.ctypes.data, x.shape[0], C.byref(loss)))
        return loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        _check(hip_lib().ndt_model_adam_step(self.ctx.h, self.h, lr, beta1, beta2, eps, weight_decay))

    def train(self, dataset: "Dataset", tp: TrainParams | None = None) -> np.ndarray:
        tp = tp or train_params()
        losses = np.zeros(tp.epochs, np.float32)
        _check(hip_lib().ndt_model_train(self.ctx.h, self.h, dataset.h, C.byref(tp), losses.ctypes.data))
        return losses

    def predict(self, bevmap: BevMap, scans: "Scans", guesses) -> np.ndarray:
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(scans.S, 16)
        out = np.zeros_like(g)
        _check(hip_lib().ndt_model_predict(self.ctx.h, self.h, bevmap.h, scans.h, g.ctypes.data, out.ctypes.data))
        return out.reshape(scans.S, 4, 4)

    def align(self, grid: Grid, bevmap: BevMap, scans: "Scans", guesses, prm: NdtParams | None = None,
              policy: int = LEARNED_RACE):
        """Batched NDT with the learned warm start; returns (results, used_learned)."""
        prm = prm or grid.params
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(scans.S, 16)
        out = (NdtResult * scans.S)()
        used = np.zeros(scans.S, np.int32)
        _check(hip_lib().ndt_align_scans_learned(self.ctx.h, grid.h, bevmap.h, self.h, scans.h, g.ctypes.data,
                                                 C.byref(prm), policy, out, used.ctypes.data))
        return _results(out, scans.S), used


class Dataset(_Handle):
    """Training triples (scan, guess, target pose) resident on the device."""
    _destroy = "ndt_dataset_destroy"

    def __init__(self, ctx: Context, bevmap: BevMap, scans: "Scans", scan_index, guesses, targets):
        self.ctx = ctx
        idx = np.ascontiguousarray(scan_index, dtype=np.int32)
        n = idx.shape[0]
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(n, 16)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(n, 16)
        self.h = C.c_void_p()
        _check(hip_lib().ndt_dataset_create(ctx.h, bevmap.h, scans.h, idx.ctypes.data, g.ctypes.data, t.ctypes.data,
                                            n, C.byref(self.h)))
        _track(self)

    def add(self, scans: "Scans", scan_index, guesses, targets):
        idx = np.ascontiguousarray(scan_index, dtype=np.int32)
        n = idx.shape[0]
        g = np.ascontiguousarray(guesses, dtype=np.float64).reshape(n, 16)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(n, 16)
        _check(hip_lib().ndt_dataset_add(self.h, scans.h, idx.ctypes.data, g.ctypes.data, t.ctypes.data, n))

    @property
    def n(self):
        a, b = C.c_int64(), C.c_int64()
        _check(hip_lib().ndt_dataset_size(self.h, C.byref(a), C.byref(b)))
        return a.value
