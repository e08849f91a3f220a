import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    # Build what is missing (nothing is rebuilt when a built copy exists): the oracle is plain
    # C, the HIP library cross-compiles for gfx950 without a GPU.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "cedar_amd", "lib", "libcedar_hip.so")):
        subprocess.run([sys.executable, "-m", "cedar_amd.build"], cwd=ROOT, check=True)


@pytest.fixture(scope="session")
def golden2d():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "bmg2_kernels.npz"))
