import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def pytest_sessionstart(session):
    # An incremental build (the Makefiles' dependencies: nothing to do when the libraries
    # are newer than their sources), so a test never runs a library older than its sources.
    # Where no compiler is present the libraries shipped beside the sources are used.
    import shutil
    import subprocess
    for d, tool in ((os.path.join(ROOT, "voxelg_amd", "csrc"), "/opt/rocm/bin/hipcc"),
                    (os.path.join(ROOT, "oracle"), "gcc")):
        if shutil.which("make") and (os.path.exists(tool) or shutil.which(tool)):
            subprocess.run(["make", "-s", "-C", d, "-j", str(min(16, os.cpu_count() or 1))],
                           check=True)


@pytest.fixture(scope="session")
def small_world():
    from voxelg_amd import workloads
    return workloads.SMALL.world()


@pytest.fixture(scope="session")
def ctx():
    from voxelg_amd.device import Context, device_count
    assert device_count() > 0, "GPU tests need a HIP device"
    return Context(0)


@pytest.fixture(scope="session")
def small(ctx, small_world):
    return ctx.world(small_world.dims).upload(small_world)
