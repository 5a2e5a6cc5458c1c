"""The sampling helpers of the full-size GPU tests, on the CPU: an upstream
closure routed alone gives the whole network's outflows bit for bit, and
the flow-scale measure is the difference over the routed magnitudes."""
import numpy as np

from conftest import flow_scale_error, load_oracle, load_trmc, sample_roots, subnetwork, upstream_segments

trmc = load_trmc()
orc = load_oracle()


def test_upstream_closure_routes_like_the_whole_network():
    T = 6
    w = trmc.make_workload(2, 20000, T, n_reaches=6000, seed=4)
    full = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T, history=True)
    roots = sample_roots(w.topo, 400, 3000, seed=1)
    size = upstream_segments(w.topo)
    assert all(size[r] <= 400 for r in roots)
    sub, seg_map, reach_map = subnetwork(w.topo, roots)
    assert sub["n_segments"] == len(seg_map) and sub["n_reaches"] == len(reach_map)
    assert np.all(np.diff(reach_map) > 0) and set(roots) <= set(reach_map.tolist())
    part = orc.wavefront_f64(sub, w.params[seg_map], w.qlat[0][:, seg_map], w.qdp0[seg_map], w.dp0[seg_map],
                             w.reach_last0[reach_map], T, history=True)
    assert np.array_equal(part["reach_hist"].view(np.uint64), full["reach_hist"][:, reach_map].view(np.uint64))
    assert np.array_equal(part["qdc"].view(np.uint64), full["qdc"][seg_map].view(np.uint64))
    # the whole network is the closure of its outlets
    outlets = np.setdiff1d(np.arange(w.topo.n_reaches), w.topo.reach_up_idx)
    whole, sm, rm = subnetwork(w.topo, outlets)
    assert np.array_equal(sm, np.arange(w.topo.n_segments)) and np.array_equal(rm, np.arange(w.topo.n_reaches))


def test_flow_scale_error_definition():
    topo = {"reach_up_ptr": np.array([0, 0, 0, 2]), "reach_up_idx": np.array([0, 1])}
    want = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 1e-4]])
    got = want.copy()
    got[1, 2] += 1e-9
    fs, rel = flow_scale_error(got, want, np.array([1.0, 1.0, 2.0]), topo)
    # reach 2 at step 1: S = |1e-4| + |3| + (|2| + |4|) + (|1| + |2|) = 12.0001
    d = got[1, 2] - want[1, 2]
    assert fs[1, 2] == d / 12.0001 and rel[1, 2] == d / 1e-3 and abs(d - 1e-9) < 1e-19
    assert fs[0].max() == 0.0 and fs[1, :2].max() == 0.0
