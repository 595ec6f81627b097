import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libatrc_gpu.so")


def gpu_available() -> bool:
    try:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context for the whole GPU session (tests print what they run)."""
    from atrc_amd import capi

    if not gpu_available():
        pytest.skip("no GPU visible")
    ctx = capi.Context(0)
    yield ctx
    ctx.close()


def log(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()
