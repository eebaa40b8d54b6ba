"""Shared test set-up: import paths for the product binding and the oracle,
the `gpu` marker, and a session-wide device context for GPU tests."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tatsy-pppm_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import pppm_amd
    ctx = pppm_amd.Context(0)
    info = ctx.device_info()
    print(f"[device] {info}", flush=True)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(GOLDEN, "reference_vectors.npz"), allow_pickle=False)
