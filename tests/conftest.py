import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def _stale(lib, sources):
    """The library is missing or older than one of its sources."""
    if not os.path.exists(lib):
        return True
    t = os.path.getmtime(lib)
    return any(os.path.getmtime(f) > t for f in sources if os.path.exists(f))


def pytest_sessionstart(session):
    # Rebuild a library that is missing or older than its sources, so a test never runs a
    # library older than the code it checks; an up-to-date library is loaded as it is (its
    # object files need not be present).
    import glob
    csrc = os.path.join(ROOT, "voxelg_amd", "csrc")
    engine = glob.glob(os.path.join(csrc, "*.hip")) + glob.glob(os.path.join(csrc, "*.h")) + \
        glob.glob(os.path.join(csrc, "*.cpp")) + [os.path.join(csrc, "Makefile"),
                                                  os.path.join(ROOT, "include", "vg.h")]
    odir = os.path.join(ROOT, "oracle")
    checker = glob.glob(os.path.join(odir, "*.c")) + glob.glob(os.path.join(odir, "*.h")) + \
        [os.path.join(odir, "Makefile")]
    import subprocess
    jobs = str(min(16, os.cpu_count() or 1))
    if _stale(os.path.join(ROOT, "voxelg_amd", "libvoxelg_amd.so"), engine):
        subprocess.run(["make", "-s", "-C", csrc, "-j", jobs], check=True)
    if _stale(os.path.join(odir, "liboracle.so"), checker):
        subprocess.run(["make", "-s", "-C", odir], check=True)


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
