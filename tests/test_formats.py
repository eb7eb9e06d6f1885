"""DAPHNE data files (ndt_input.dat / ndt_output.dat) and the reference's golden
output, plus the acceptance check."""
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_roundtrip(A, tmp_path):
    scene = A.scene_cfg(2, extent=30.0)
    mp = A.make_map(scene, 3000)
    scan, gt, guess = A.make_scan(scene, mp, A.scan_cfg(200), 0)
    cases = [A.Testcase(guess.astype(np.float32), scan, mp)] * 2
    A.write_ndt_input(tmp_path / "in.dat", cases)
    back = A.read_ndt_input(tmp_path / "in.dat")
    assert len(back) == 2 and np.array_equal(back[1].map, mp) and np.array_equal(back[0].guess, cases[0].guess)
    res = [A.CallbackResult(True, np.eye(4, dtype=np.float32), 0.0)]
    A.write_ndt_output(tmp_path / "out.dat", res)
    assert os.path.getsize(tmp_path / "out.dat") == 73


def _quirky_rotation(a):
    """Closed form of buildTransformationMatrix (kernel.cpp:949-1013) in float64:
    the half-angle quaternions q1 (roll), q2 (pitch), q3 (yaw) are multiplied with the
    components of q1*q2 relabelled (its w stored as x, x as y, y as z, z as w) and
    the result converted with the standard quaternion-to-matrix formula."""
    h = np.asarray(a) / 2
    c1, s1, c2, s2, c3, s3 = np.cos(h[0]), np.sin(h[0]), np.cos(h[1]), np.sin(h[1]), np.cos(h[2]), np.sin(h[2])
    W, X, Y, Z = c1 * c2, s1 * c2, c1 * s2, s1 * s2
    x, y, z, w = Z * c3 - Y * s3, W * c3 + X * s3, X * c3 - W * s3, Z * s3 + Y * c3
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_reference_golden_output_file(A, orc):
    """The reference's data/ndt_output.dat.gz (copied as a fixture): 5 records,
    converged, zero translation (the reference never writes the translation column,
    so it stays at the guess's), and every rotation block in the image of the
    oracle's compat buildTransformationMatrix: fitted angles reproduce it."""
    from scipy.optimize import least_squares
    recs = A.read_ndt_output(os.path.join(REPO, "tests", "golden", "reference_ndt_output.dat.gz"))
    assert len(recs) == 5
    rng = np.random.default_rng(0)
    p = rng.uniform(-3, 3, 3)
    eye = np.eye(4, dtype=np.float32)
    assert np.abs(_quirky_rotation(p) - orc.compat_build_matrix(eye, np.r_[0, 0, 0, p])[:3, :3]).max() < 1e-6
    for r in recs:
        M = r.final_transformation
        assert r.converged and np.all(M[:3, 3] == 0) and np.array_equal(M[3], [0, 0, 0, 1])
        G = M[:3, :3].astype(np.float64)
        assert abs(G[0, 1]) < 1e-8  # structural zero of the relabelled quaternion product
        fit = min((least_squares(lambda a: (_quirky_rotation(a) - G).ravel(), rng.uniform(-np.pi, np.pi, 3))
                   for _ in range(30)), key=lambda s: s.cost)
        assert np.abs(_quirky_rotation(fit.x) - G).max() < 1e-6
        assert np.abs(orc.compat_build_matrix(eye, np.r_[0, 0, 0, fit.x])[:3, :3] - G).max() < 1e-3


def test_small_daphne_fixture(A):
    cases = A.read_ndt_input(os.path.join(REPO, "tests", "golden", "daphne_small", "ndt_input.dat.gz"))
    outs = A.read_ndt_output(os.path.join(REPO, "tests", "golden", "daphne_small", "ndt_output.dat.gz"))
    assert len(cases) == len(outs) == 3
    ok, md = A.check_outputs(outs, outs)
    assert ok and md == 0.0
    bad = [A.CallbackResult(o.converged, o.final_transformation + np.float32(0.01)) for o in outs]
    assert not A.check_outputs(bad, outs)[0]


def _small_input_bytes():
    import gzip
    with gzip.open(os.path.join(REPO, "tests", "golden", "daphne_small", "ndt_input.dat.gz")) as f:
        return f.read()


def test_truncated_input_raises_reference_messages(A, tmp_path):
    """A file cut inside a record raises DataFormatError with the reference's message
    for that record (kernel.cpp:157-226); so does a negative point count."""
    import struct
    import pytest
    data = _small_input_bytes()
    (n_scan,) = struct.unpack_from("<i", data, 4 + 64)
    cuts = {2: "Error reading number of test cases", 4 + 30: "Error reading initial guess",
            4 + 64 + 2: "Error reading filtered scan", 4 + 64 + 4 + 16 * n_scan - 8: "Error reading filtered scan",
            4 + 64 + 4 + 16 * n_scan + 4 + 40: "Error reading filtered scan", len(data) - 1: "Error reading filtered scan"}
    for cut, msg in cuts.items():
        p = tmp_path / f"cut{cut}.dat"
        p.write_bytes(data[:cut])
        with pytest.raises(A.DataFormatError, match=msg):
            A.read_ndt_input(p)
    neg = bytearray(data)
    struct.pack_into("<i", neg, 4 + 64, -5)
    (tmp_path / "neg.dat").write_bytes(bytes(neg))
    with pytest.raises(A.DataFormatError, match="filtered scan"):
        A.read_ndt_input(tmp_path / "neg.dat")
    # an empty testcase list and an empty scan are well-formed
    (tmp_path / "zero.dat").write_bytes(struct.pack("<i", 0))
    assert A.read_ndt_input(tmp_path / "zero.dat") == []
    empty = A.Testcase(np.eye(4, dtype=np.float32), np.zeros((0, 4), np.float32), np.ones((3, 4), np.float32))
    A.write_ndt_input(tmp_path / "empty.dat", [empty])
    back = A.read_ndt_input(tmp_path / "empty.dat")
    assert back[0].scan.shape == (0, 4) and back[0].map.shape == (3, 4)


def test_truncated_output_raises(A, tmp_path):
    import pytest
    A.write_ndt_output(tmp_path / "o.dat", [A.CallbackResult(True, np.eye(4, dtype=np.float32), 0.5)] * 2)
    data = (tmp_path / "o.dat").read_bytes()
    (tmp_path / "t.dat").write_bytes(data[:-10])
    with pytest.raises(A.DataFormatError, match="reference result"):
        A.read_ndt_output(tmp_path / "t.dat")
    assert len(A.read_ndt_output(tmp_path / "o.dat")) == 2


def test_harness_rejects_missing_or_truncated_files(tmp_path):
    """The testcase binary exits with the reference's status -3 (kernel.cpp:211-214,
    :345-366) and message when a data file is missing or its header is cut, before it
    touches a GPU."""
    import subprocess
    import pytest
    exe = os.path.join(REPO, "daphne-benchmark_amd", "ndt_mapping_hip")
    if not os.path.exists(exe):
        pytest.skip("harness not built")
    data = tmp_path / "data"
    run = tmp_path / "a" / "b" / "c"
    data.mkdir()
    run.mkdir(parents=True)
    env = dict(os.environ, DAPHNE_DATA_DIR=str(data))
    p = subprocess.run([exe], cwd=run, capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 253 and "Error opening the testcase file" in p.stderr, p.stderr
    (data / "ndt_input.dat").write_bytes(b"\x01\x00")
    (data / "ndt_output.dat").write_bytes(b"")
    p = subprocess.run([exe], cwd=run, capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 253 and "Error reading number of test cases" in p.stderr, p.stderr
    p = subprocess.run([exe, "-p", "0"], cwd=run, capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 4 and "Usage" in p.stdout
