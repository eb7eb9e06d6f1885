"""Shared fixtures of the test suite.

Markers: `gpu` marks tests that need an MI355X (they call libndthip.so through its
C ABI); everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "daphne-benchmark_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP engine)")


@pytest.fixture(scope="session")
def A():
    import ndt_amd
    return ndt_amd


@pytest.fixture(scope="session")
def orc():
    import orc as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def compat_cases():
    d = np.load(os.path.join(GOLDEN, "compat_cases.npz"))
    out = []
    for k in range(int(d["ncases"][0])):
        c = {key[len(f"c{k}_"):]: d[key] for key in d.files if key.startswith(f"c{k}_")}
        j = int(c["scene"][0])
        c.update({key[len(f"s{j}_"):]: d[key] for key in d.files if key.startswith(f"s{j}_")})
        out.append(c)
    return out


@pytest.fixture(scope="session")
def compat_math():
    return dict(np.load(os.path.join(GOLDEN, "compat_math.npz")))


@pytest.fixture(scope="session")
def ctx(A):
    c = A.Context(0)
    yield c
    c.close()
