"""ctypes bindings of the CPU oracle (liborc.so, ndt_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU baseline,
never the thing measured on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liborc.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

D = C.c_double
P = C.c_void_p


class OrcResult(C.Structure):
    _fields_ = [("final_f32", C.c_float * 16), ("final_f64", C.c_double * 16), ("score", D),
                ("trans_probability", D), ("iterations", C.c_int32), ("converged", C.c_int32),
                ("passes", C.c_int32), ("pad", C.c_int32)]


class OrcVoxel(C.Structure):
    _fields_ = [("icov", D * 9), ("mean", D * 3), ("n", C.c_int32)]


class OrcCompatGrid(C.Structure):
    _fields_ = [("minv", C.c_float * 4), ("maxv", C.c_float * 4), ("dims", C.c_int32 * 3),
                ("ncells", C.c_int64), ("cells", C.POINTER(OrcVoxel))]


class OrcNdtParams(C.Structure):
    _fields_ = [("resolution", D), ("step_size", D), ("trans_eps", D), ("outlier_ratio", D),
                ("max_iterations", C.c_int32), ("min_points", C.c_int32), ("min_eig_ratio", D),
                ("hessian_at_accept", C.c_int32), ("repeat_passes", C.c_int32)]


class OrcMtOut(C.Structure):
    _fields_ = [("a", D), ("sign", C.c_int32), ("evals", C.c_int32), ("loop_iterations", C.c_int32),
                ("exit_reason", C.c_int32)]


LINE_FN = C.CFUNCTYPE(D, C.c_void_p, D, C.c_int, C.POINTER(D))
MT_ACCEPTED, MT_INTERVAL, MT_MAX_ITER, MT_NO_DESCENT = range(4)
MT_DEAD_LOOP, MT_HESSIAN_AT_ACCEPT = 1, 2


class OrcNdtVoxel(C.Structure):
    _fields_ = [("key", C.c_int64 * 3), ("mean", D * 3), ("cov", D * 9), ("icov", D * 9),
                ("n", C.c_int32), ("valid", C.c_int32)]


class OrcNdtGrid(C.Structure):
    _fields_ = [("resolution", D), ("nvox", C.c_int64), ("vox", C.POINTER(OrcNdtVoxel)),
                ("cap", C.c_int64), ("slot", C.POINTER(C.c_int64))]


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_gauss_consts.argtypes = [D, D, P, P]
        L.orc_angle_tables.argtypes = [P, P, P]
        L.orc_ndt_angle_tables.argtypes = [P, P, P]
        L.orc_point_derivs.argtypes = [P, P, P, P, P]
        L.orc_solve6.argtypes = [P, P, P]
        L.orc_mt_update_interval.argtypes = [P, P, P, P, P, P, D, D, D]
        L.orc_mt_update_interval.restype = C.c_int
        L.orc_mt_trial_value.argtypes = [D] * 9
        L.orc_mt_trial_value.restype = D
        L.orc_mt_search.argtypes = [LINE_FN, C.c_void_p, D, D, D, D, D, C.c_int32, C.POINTER(OrcMtOut)]
        L.orc_mt_search.restype = C.c_int
        L.orc_compat_grid_build.argtypes = [P, C.c_int64, C.c_float, C.POINTER(OrcCompatGrid)]
        L.orc_compat_grid_free.argtypes = [C.POINTER(OrcCompatGrid)]
        L.orc_compat_radius_search.argtypes = [C.POINTER(OrcCompatGrid), C.c_float, P, D, P, P]
        L.orc_compat_update.argtypes = [P, P, P, P, P, P, D, D, C.c_int]
        L.orc_compat_update.restype = D
        L.orc_compat_derivatives.argtypes = [C.POINTER(OrcCompatGrid), C.c_float, P, P, C.c_int64, P, D, D,
                                             P, P, P]
        L.orc_compat_derivatives.restype = D
        L.orc_compat_transform.argtypes = [P, P, C.c_int64, P]
        L.orc_compat_build_matrix.argtypes = [P, P]
        L.orc_compat_euler.argtypes = [P, P]
        for fn in (L.orc_compat_align, L.orc_compat_align_omp):
            fn.argtypes = [P, C.c_int64, P, C.c_int64, P, C.POINTER(OrcResult)]
            fn.restype = C.c_int
        L.orc_ndt_params_default.argtypes = [C.POINTER(OrcNdtParams)]
        L.orc_ndt_grid_build.argtypes = [P, C.c_int64, C.POINTER(OrcNdtParams), C.POINTER(OrcNdtGrid)]
        L.orc_ndt_grid_build.restype = C.c_int
        L.orc_ndt_grid_free.argtypes = [C.POINTER(OrcNdtGrid)]
        L.orc_ndt_grid_find.argtypes = [C.POINTER(OrcNdtGrid), C.c_int64, C.c_int64, C.c_int64]
        L.orc_ndt_grid_find.restype = C.c_int64
        L.orc_sym3_eig.argtypes = [P, P, P]
        L.orc_ndt_pose_matrix.argtypes = [P, P]
        L.orc_ndt_euler.argtypes = [P, P]
        L.orc_ndt_derivatives.argtypes = [C.POINTER(OrcNdtGrid), C.POINTER(OrcNdtParams), P, C.c_int64, P, P, P,
                                          C.c_int, P]
        L.orc_ndt_derivatives.restype = D
        L.orc_ndt_score.argtypes = [C.POINTER(OrcNdtGrid), C.POINTER(OrcNdtParams), P, C.c_int64, P]
        L.orc_ndt_score.restype = D
        L.orc_ndt_align.argtypes = [C.POINTER(OrcNdtGrid), C.POINTER(OrcNdtParams), P, C.c_int64, P,
                                    C.POINTER(OrcResult)]
        L.orc_ndt_align.restype = C.c_int
        L.orc_ndt_align_many.argtypes = [C.POINTER(OrcNdtGrid), C.POINTER(OrcNdtParams), P, P, P, C.c_int32,
                                         C.c_int32, P]
        L.orc_ndt_align_many.restype = C.c_int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def result_dict(r: OrcResult) -> dict:
    return dict(final_f32=np.ctypeslib.as_array(r.final_f32).reshape(4, 4).copy(),
                final_f64=np.ctypeslib.as_array(r.final_f64).reshape(4, 4).copy(),
                score=r.score, trans_probability=r.trans_probability, iterations=r.iterations,
                converged=bool(r.converged), passes=r.passes)


# ----------------------------- compat -----------------------------

def compat_align(map4, scan4, guess, omp: bool = False) -> dict:
    map4, scan4, guess = _f32(map4), _f32(scan4), _f32(guess)
    r = OrcResult()
    fn = lib().orc_compat_align_omp if omp else lib().orc_compat_align
    if fn(map4.ctypes.data, map4.shape[0], scan4.ctypes.data, scan4.shape[0], guess.ctypes.data, C.byref(r)):
        raise ValueError("orc_compat_align failed")
    return result_dict(r)


class CompatGrid:
    def __init__(self, map4, resolution: float = 1.0):
        self.map4 = _f32(map4)
        self.res = resolution
        self.g = OrcCompatGrid()
        if lib().orc_compat_grid_build(self.map4.ctypes.data, self.map4.shape[0], resolution, C.byref(self.g)):
            raise ValueError("orc_compat_grid_build failed")

    def __del__(self):
        if getattr(self, "g", None) is not None and _lib is not None:
            _lib.orc_compat_grid_free(C.byref(self.g))

    @property
    def dims(self):
        return tuple(self.g.dims)

    def occupied(self):
        """(idx, n, mean(k,3), icov(k,9)) of cells with points, in index order."""
        cells = np.ctypeslib.as_array(C.cast(self.g.cells, C.POINTER(C.c_uint8)),
                                      shape=(self.g.ncells * C.sizeof(OrcVoxel),))
        rec = np.frombuffer(cells.tobytes(), dtype=np.dtype([("icov", "<f8", 9), ("mean", "<f8", 3),
                                                              ("n", "<i4"), ("pad", "<i4")]))
        occ = np.nonzero(rec["n"] > 0)[0]
        return occ, rec["n"][occ], rec["mean"][occ], rec["icov"][occ]

    def radius_search(self, pt) -> tuple:
        pt = _f32(pt)
        cells = np.zeros(27, dtype=np.int64)
        dist = np.zeros(27, dtype=np.float32)
        k = lib().orc_compat_radius_search(C.byref(self.g), self.res, pt.ctypes.data, float(self.res),
                                           cells.ctypes.data, dist.ctypes.data)
        return cells[:min(k, 27)], dist[:min(k, 27)]

    def derivatives(self, scan4, trans4, p, d1, d2):
        scan4, trans4, p = _f32(scan4), _f32(trans4), _f64(p)
        g = np.zeros(6)
        H = np.zeros(36)
        npairs = C.c_int64(0)
        s = lib().orc_compat_derivatives(C.byref(self.g), self.res, scan4.ctypes.data, trans4.ctypes.data,
                                         scan4.shape[0], p.ctypes.data, d1, d2, g.ctypes.data, H.ctypes.data,
                                         C.byref(npairs))
        return s, g, H.reshape(6, 6), npairs.value


def gauss_consts(resolution=1.0, outlier=0.55):
    d1, d2 = C.c_double(), C.c_double()
    lib().orc_gauss_consts(resolution, outlier, C.byref(d1), C.byref(d2))
    return d1.value, d2.value


def angle_tables(p, exact: bool = False):
    """The reference's angle tables; exact=True gives the textbook mode's tables."""
    p = _f64(p)
    j = np.zeros((8, 3))
    h = np.zeros((15, 3))
    fn = lib().orc_ndt_angle_tables if exact else lib().orc_angle_tables
    fn(p.ctypes.data, j.ctypes.data, h.ctypes.data)
    return j, h


def point_derivs(x, p, exact: bool = False):
    j, h = angle_tables(p, exact)
    x = _f64(x)
    pg = np.zeros((3, 6))
    ph = np.zeros((18, 6))
    lib().orc_point_derivs(x.ctypes.data, j.ctypes.data, h.ctypes.data, pg.ctypes.data, ph.ctypes.data)
    return pg, ph


def solve6(A, b):
    A, b = _f64(A), _f64(b)
    x = np.zeros(6)
    lib().orc_solve6(x.ctypes.data, A.ctypes.data, b.ctypes.data)
    return x


def mt_trial(*args):
    return lib().orc_mt_trial_value(*[float(a) for a in args])


def mt_search(phi, phi_0, d_phi_0, step_init, step_max, step_min, flags=0):
    """The oracle's More-Thuente search on a Python scalar function phi(t) -> (value,
    derivative), t the signed step along the direction.  Returns (a, sign, evals,
    loop iterations, exit reason, [(t, with_hessian), ...] of every evaluation)."""
    calls = []

    def cb(_ctx, t, with_hessian, dphi):
        v, d = phi(t)
        calls.append((t, with_hessian))
        dphi[0] = d
        return v

    fn = LINE_FN(cb)
    out = OrcMtOut()
    lib().orc_mt_search(fn, None, phi_0, d_phi_0, step_init, step_max, step_min, flags, C.byref(out))
    return out.a, out.sign, out.evals, out.loop_iterations, out.exit_reason, calls


def mt_update(a_l, f_l, g_l, a_u, f_u, g_u, a_t, f_t, g_t):
    v = [C.c_double(x) for x in (a_l, f_l, g_l, a_u, f_u, g_u)]
    ret = lib().orc_mt_update_interval(*[C.byref(x) for x in v], a_t, f_t, g_t)
    return ret, [x.value for x in v]


def compat_update(xt, cinv, x, p, d1, d2):
    pg, ph = point_derivs(x, p)
    g = np.zeros(6)
    H = np.zeros(36)
    xt, cinv = _f64(xt), _f64(cinv)
    inc = lib().orc_compat_update(g.ctypes.data, H.ctypes.data, xt.ctypes.data, cinv.ctypes.data, pg.ctypes.data,
                                  ph.ctypes.data, d1, d2, 1)
    return inc, g, H.reshape(6, 6)


def compat_build_matrix(M, p):
    M = _f32(M).copy()
    p = _f64(p)
    lib().orc_compat_build_matrix(M.ctypes.data, p.ctypes.data)
    return M


def compat_euler(M):
    M = _f32(M)
    out = np.zeros(3)
    lib().orc_compat_euler(M.ctypes.data, out.ctypes.data)
    return out


def compat_transform(pts, M):
    pts, M = _f32(pts), _f32(M)
    out = np.zeros_like(pts)
    lib().orc_compat_transform(pts.ctypes.data, out.ctypes.data, pts.shape[0], M.ctypes.data)
    return out


# ----------------------------- ndt (float64 textbook) -----------------------------

def ndt_params(**kw) -> OrcNdtParams:
    p = OrcNdtParams()
    lib().orc_ndt_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class NdtGrid:
    def __init__(self, map4, params: OrcNdtParams | None = None):
        self.map4 = _f32(map4)
        self.params = params or ndt_params()
        self.g = OrcNdtGrid()
        if lib().orc_ndt_grid_build(self.map4.ctypes.data, self.map4.shape[0], C.byref(self.params),
                                    C.byref(self.g)):
            raise ValueError("orc_ndt_grid_build failed")

    def __del__(self):
        if getattr(self, "g", None) is not None and _lib is not None:
            _lib.orc_ndt_grid_free(C.byref(self.g))

    def voxels(self) -> dict:
        n = self.g.nvox
        raw = np.ctypeslib.as_array(C.cast(self.g.vox, C.POINTER(C.c_uint8)), shape=(n * C.sizeof(OrcNdtVoxel),))
        rec = np.frombuffer(raw.tobytes(), dtype=np.dtype([("key", "<i8", 3), ("mean", "<f8", 3), ("cov", "<f8", 9),
                                                           ("icov", "<f8", 9), ("n", "<i4"), ("valid", "<i4")]))
        return {k: rec[k].copy() for k in ("key", "mean", "cov", "icov", "n", "valid")}

    def derivatives(self, scan4, p, hessian=True):
        scan4, p = _f32(scan4), _f64(p)
        g = np.zeros(6)
        H = np.zeros(36)
        npairs = C.c_int64(0)
        s = lib().orc_ndt_derivatives(C.byref(self.g), C.byref(self.params), scan4.ctypes.data, scan4.shape[0],
                                      p.ctypes.data, g.ctypes.data, H.ctypes.data, 1 if hessian else 0,
                                      C.byref(npairs))
        return s, g, H.reshape(6, 6), npairs.value

    def score(self, scan4, p):
        scan4, p = _f32(scan4), _f64(p)
        return lib().orc_ndt_score(C.byref(self.g), C.byref(self.params), scan4.ctypes.data, scan4.shape[0],
                                   p.ctypes.data)

    def align(self, scan4, guess) -> dict:
        scan4, guess = _f32(scan4), _f64(guess)
        r = OrcResult()
        if lib().orc_ndt_align(C.byref(self.g), C.byref(self.params), scan4.ctypes.data, scan4.shape[0],
                               guess.ctypes.data, C.byref(r)):
            raise ValueError("orc_ndt_align failed")
        return result_dict(r)

    def align_many(self, scans, guesses, threads: int = 0) -> list:
        scans = [_f32(s) for s in scans]
        S = len(scans)
        ptrs = (C.c_void_p * S)(*[s.ctypes.data for s in scans])
        counts = np.array([s.shape[0] for s in scans], dtype=np.int64)
        guesses = _f64(guesses).reshape(S, 16)
        out = (OrcResult * S)()
        if lib().orc_ndt_align_many(C.byref(self.g), C.byref(self.params), ptrs, counts.ctypes.data,
                                    guesses.ctypes.data, S, threads, out):
            raise ValueError("orc_ndt_align_many failed")
        return [result_dict(out[i]) for i in range(S)]


def pose_matrix(p):
    p = _f64(p)
    M = np.zeros(16)
    lib().orc_ndt_pose_matrix(p.ctypes.data, M.ctypes.data)
    return M.reshape(4, 4)


def ndt_euler(M):
    M = _f64(M)
    out = np.zeros(3)
    lib().orc_ndt_euler(M.ctypes.data, out.ctypes.data)
    return out


def sym3_eig(A):
    A = _f64(A)
    ev = np.zeros(3)
    V = np.zeros(9)
    lib().orc_sym3_eig(A.ctypes.data, ev.ctypes.data, V.ctypes.data)
    return ev, V.reshape(3, 3)


def pose_error(T, T_ref):
    """(translation error in metres, rotation angle error in radians) of two 4x4 poses."""
    T = np.asarray(T, dtype=np.float64)
    T_ref = np.asarray(T_ref, dtype=np.float64)
    dt = float(np.linalg.norm(T[:3, 3] - T_ref[:3, 3]))
    R = T[:3, :3].T @ T_ref[:3, :3]
    c = (np.trace(R) - 1.0) / 2.0
    c = min(1.0, max(-1.0, c))
    return dt, float(np.arccos(c))


def fitness_score(map_xyzw, scan_xyzw, pose, max_range=np.inf):
    """Fitness score (PCL Registration::getFitnessScore): mean squared distance from
    the transformed scan points (float64 transform of the float32 points) to their
    nearest map points, over points whose nearest map point lies within max_range;
    the largest double when none does.  Brute force in float64, in blocks.
    Returns (fitness, count).  The reference declares ndt_getFitnessScore
    (src/Cuda/ndt_mapping/kernel.cu:124) without implementing it."""
    m = np.asarray(map_xyzw, np.float32)[:, :3].astype(np.float64)
    m = m[np.all(np.isfinite(m), axis=1)]
    s = np.asarray(scan_xyzw, np.float32)[:, :3].astype(np.float64)
    T = np.asarray(pose, np.float64).reshape(4, 4)
    x = s @ T[:3, :3].T + T[:3, 3]
    x = x[np.all(np.isfinite(x), axis=1)]
    mm = np.sum(m * m, axis=1)
    best = np.empty(len(x))
    for b in range(0, len(x), 256):
        xb = x[b:b + 256]
        # exact differences for the candidate found by the expanded form
        d2 = np.sum(xb * xb, axis=1)[:, None] - 2.0 * xb @ m.T + mm[None, :]
        j = np.argsort(d2, axis=1)[:, :8]
        exact = np.sum((xb[:, None, :] - m[j]) ** 2, axis=2)
        best[b:b + 256] = exact.min(axis=1)
    ok = best <= max_range * max_range if np.isfinite(max_range) else np.ones(len(best), bool)
    n = int(ok.sum())
    return (float(best[ok].sum() / n) if n else 1.7976931348623157e308), n
