import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")


@pytest.fixture(scope="session")
def pack():
    """pack(config, **options) -> Packed, built by the JS host modules and cached per session."""
    from raymond_amd import scenefile
    cache = {}

    def get(config, **options):
        key = (config, repr(sorted(options.items())))
        if key not in cache:
            cache[key] = scenefile.pack(config, options)
        return cache[key]
    return get


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests."""
    from raymond_amd import capi
    c = capi.Context(0)
    yield c
    c.close()
