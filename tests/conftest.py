import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def pytest_sessionstart(session):
    # Tests load what is built and build only when nothing is.
    lib = os.path.join(ROOT, "voxelg_amd", "libvoxelg_amd.so")
    olib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not (os.path.exists(lib) and os.path.exists(olib)):
        import __graft_entry__
        __graft_entry__.build()


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
