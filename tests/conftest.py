"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU.

The oracle (oracle/) is the checker only: tests compare the HIP library's
results against it and against the committed golden vectors, never the other
way round."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_trmc():
    return _load("noaa_gpu_kernels_amd", os.path.join(ROOT, "noaa-gpu-kernels_amd", "__init__.py"))


def load_oracle():
    return _load("trmc_oracle", os.path.join(ROOT, "oracle", "oracle.py"))


@pytest.fixture(scope="session")
def trmc():
    return load_trmc()


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


@pytest.fixture(scope="session")
def gpu(trmc):
    """Skips unless a gfx950 device is visible (tests using it are marked gpu)."""
    if trmc.device_count() < 1:
        pytest.skip("no gfx950 device")
    return True


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def tiny_cases():
    z = golden("wavefront_tiny.npz")
    out = {}
    for name in z["names"]:
        pre = f"{name}__"
        out[str(name)] = {k[len(pre):]: z[k] for k in z.files if k.startswith(pre)}
    return out


class TopoView:
    """Attribute view of a topology dict (golden fixtures) for the bindings."""

    def __init__(self, d):
        trmc = load_trmc()
        self.t = trmc.Topology(
            int(d["n_reaches"]), int(d["n_segments"]), int(d["n_levels"]), int(d["max_upstream"]),
            np.asarray(d["reach_seg_start"]), np.asarray(d["reach_seg_len"]),
            np.asarray(d["reach_level"]), np.asarray(d["reach_n_up"]), np.asarray(d["reach_up_ptr"]),
            np.asarray(d["reach_up_idx"]), np.asarray(d["level_ptr"]), np.asarray(d["level_reach"]))


def assert_close(got, want, rtol=1e-12, atol=1e-15, what=""):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    err = np.abs(got - want)
    bad = err > rtol * np.abs(want) + atol
    if bad.any():
        i = int(np.argmax(err / (np.abs(want) + atol)))
        raise AssertionError(
            f"{what}: {int(bad.sum())}/{bad.size} entries off; worst index {i}: got {got.flat[i]!r} "
            f"want {want.flat[i]!r} (rel {err.flat[i] / max(abs(want.flat[i]), 1e-300):.3e})")


def max_rel(got, want, floor=1e-3):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), floor))) if got.size else 0.0
