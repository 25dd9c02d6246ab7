"""Shared fixtures.  `gpu` marks tests that need an MI355X; everything else
runs on the CPU (oracle, tracer, host logic, C-ABI loading)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")


def _ensure_built():
    lib = ROOT / "openmoc_amd" / "libopenmoc_amd.so"
    if not lib.exists():
        from openmoc_amd import build
        build.build()
    orc = ROOT / "oracle" / "liboracle.so"
    if not orc.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return dict(np.load(GOLDEN / f"{name}_ref.npz"))
    return load


@pytest.fixture(scope="session")
def tracks():
    """Cache of generated TrackGenerators by workload name."""
    import openmoc_amd as om
    cache = {}

    def get(name, **kw):
        key = (name, tuple(sorted(kw.items())))
        if key not in cache:
            tg = om.TrackGenerator(om.get_workload(name, **kw))
            tg.generateTracks()
            cache[key] = tg
        return cache[key]
    return get
