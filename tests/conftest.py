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


def subnetwork(topo, roots):
    """The upstream closure of reaches `roots`: a topology dict (reaches keep
    their relative order, so ids stay topological and every upstream list its
    order) plus seg_map / reach_map, the original segment and reach ids.  A
    reach's outflows depend only on its upstream closure, so routing the
    closure alone gives the same values as the whole network."""
    up_ptr, up_idx = np.asarray(topo.reach_up_ptr), np.asarray(topo.reach_up_idx)
    keep = np.zeros(topo.n_reaches, bool)
    stack = [int(r) for r in roots]
    while stack:
        r = stack.pop()
        if keep[r]:
            continue
        keep[r] = True
        stack.extend(int(u) for u in up_idx[up_ptr[r]:up_ptr[r + 1]])
    reach_map = np.flatnonzero(keep)
    new_id = -np.ones(topo.n_reaches, np.int64)
    new_id[reach_map] = np.arange(len(reach_map))
    start, length = np.asarray(topo.reach_seg_start), np.asarray(topo.reach_seg_len)
    seg_len = length[reach_map].astype(np.int32)
    seg_start = np.concatenate([[0], np.cumsum(seg_len)[:-1]]).astype(np.int32)
    seg_map = np.concatenate([np.arange(start[r], start[r] + length[r]) for r in reach_map]).astype(np.int64) \
        if len(reach_map) else np.zeros(0, np.int64)
    ups = [new_id[up_idx[up_ptr[r]:up_ptr[r + 1]]] for r in reach_map]
    n_up = np.array([len(u) for u in ups], np.int32)
    level = np.asarray(topo.reach_level)[reach_map].astype(np.int32)
    L = int(level.max()) + 1 if len(level) else 0
    order = np.argsort(level, kind="stable").astype(np.int32)
    d = dict(n_reaches=len(reach_map), n_segments=int(seg_len.sum()), n_levels=L,
             max_upstream=int(n_up.max()) if len(n_up) else 0, reach_seg_start=seg_start, reach_seg_len=seg_len,
             reach_level=level, reach_n_up=n_up,
             reach_up_ptr=np.concatenate([[0], np.cumsum(n_up)]).astype(np.int32),
             reach_up_idx=(np.concatenate(ups) if len(ups) else np.zeros(0)).astype(np.int32),
             level_ptr=np.concatenate([[0], np.cumsum(np.bincount(level, minlength=L))]).astype(np.int32),
             level_reach=order)
    return d, seg_map, reach_map


def upstream_segments(topo):
    """Segments in each reach's upstream closure (itself included); the
    networks are trees, so the counts add up along the upstream lists."""
    up_ptr, up_idx = np.asarray(topo.reach_up_ptr), np.asarray(topo.reach_up_idx)
    total = np.asarray(topo.reach_seg_len).astype(np.int64).copy()
    for r in range(topo.n_reaches):                    # ids are topological
        a, b = up_ptr[r], up_ptr[r + 1]
        if b > a:
            total[r] += total[up_idx[a:b]].sum()
    return total


def sample_roots(topo, max_segments, total_segments, seed):
    """Random reaches whose upstream closures hold at most max_segments
    segments, until the closures hold about total_segments."""
    size = upstream_segments(topo)
    cand = np.flatnonzero(size <= max_segments)
    rng = np.random.default_rng(seed)
    rng.shuffle(cand)
    out, acc = [], 0
    for r in cand:
        out.append(int(r))
        acc += int(size[r])
        if acc >= total_segments:
            break
    return out


def flow_scale_error(got_hist, want_hist, reach_last0, topo, reach_ids=None):
    """Per reach and step: |got - want| against the flow scale through the
    reach, S = |Q_r(t)| + |Q_r(t-1)| + sum over upstream reaches u of
    |Q_u(t)| + |Q_u(t-1)| (the magnitudes the routing combines), from the
    `want` history [T, R] (columns reach_ids of the whole network's reaches
    when given).  Returns (|diff| / S, plain relative |diff| / max(|want|, 1e-3))."""
    want = np.asarray(want_hist, np.float64)
    got = np.asarray(got_hist, np.float64)
    prev = np.vstack([np.asarray(reach_last0, np.float64)[None], want[:-1]])
    S = np.abs(want) + np.abs(prev)
    up_ptr, up_idx = np.asarray(topo["reach_up_ptr"]), np.asarray(topo["reach_up_idx"])
    for r in range(want.shape[1]):
        u = up_idx[up_ptr[r]:up_ptr[r + 1]]
        if len(u):
            S[:, r] += np.abs(want[:, u]).sum(1) + np.abs(prev[:, u]).sum(1)
    d = np.abs(got - want)
    return d / np.maximum(S, 1e-300), d / np.maximum(np.abs(want), 1e-3)
