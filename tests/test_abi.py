"""The C ABI library loads and exports every entry point its header declares (no
GPU needed), parameter defaults follow the reference, and the product path fails
loudly without a device."""
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    return set(re.findall(r"^\w[\w\s\*]*?\b(ndt_\w+)\s*\(", text, re.M))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_hip_library_exports_header(A):
    A.hip_lib()
    names = declared("ndt_hip.h")
    assert len(names) >= 18
    missing = names - exported(A.HIP_LIB_PATH)
    assert not missing, missing
    assert set(A.EXPORTED_SYMBOLS) <= names


def test_learn_entry_points_exported(A):
    names = declared("ndt_learn.h")
    assert len(names) >= 18
    missing = names - exported(A.HIP_LIB_PATH)
    assert not missing, missing
    assert set(A.LEARN_SYMBOLS) == names


def test_scene_library_exports_header(A):
    A.scene_lib()
    names = declared("ndt_scene.h")
    missing = names - exported(os.path.join(A.PKG_DIR, "libndt_scene.so"))
    assert not missing, missing


def test_param_defaults(A):
    c = A.params(A.NDT_MODE_COMPAT)
    assert (c.resolution, c.step_size, c.trans_eps, c.max_iterations, c.outlier_ratio) == (1.0, 0.1, 0.1, 0, 0.55)
    n = A.params(A.NDT_MODE_NDT)
    assert (n.trans_eps, n.max_iterations, n.min_points, n.min_eig_ratio) == (0.01, 30, 6, 0.01)


def test_no_device_fails_loudly(A):
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(A.NdtError):
        A.Context(0)


def test_result_conversion_matches_fields(A):
    """The structured-view conversion of result arrays gives the same dicts as
    reading the ctypes fields one by one."""
    import ctypes as C
    out = (A.NdtResult * 7)()
    for i in range(7):
        out[i].iterations, out[i].converged, out[i].passes, out[i].status = i, i % 2, 3 * i, i % 3
        out[i].score, out[i].trans_probability = -1.5 * i, 0.25 * i
        for k in range(16):
            out[i].final_transformation[k] = i + 0.5 * k
            out[i].final_transformation_f64[k] = i * 1e7 + k / 3.0
    assert A._RESULT_DTYPE.itemsize == C.sizeof(A.NdtResult)
    for x, y in zip([A._result(out[i]) for i in range(7)], A._results(out, 7)):
        assert x.keys() == y.keys()
        for k in x:
            assert np.array_equal(x[k], y[k]) and type(x[k]) is type(y[k]), k
