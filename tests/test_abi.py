"""C-ABI library on the CPU: loads, exports every declared symbol, validates
topologies, plans schedules, generates workloads, reads/writes the formats.
No routing calls here (those need a GPU)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, TopoView, golden, load_oracle, load_trmc

trmc = load_trmc()


def declared_functions():
    text = open(os.path.join(ROOT, "include", "trmc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(trmc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = trmc.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", trmc.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (trmc_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)
    assert lib.trmc_version() >= 10000


def test_no_device_is_an_error_not_a_fallback():
    if trmc.device_count() > 0:
        pytest.skip("a GPU is present")
    topo = trmc.generate_network(1, 200, 0, 1)
    with pytest.raises(trmc.TrmcError) as e:
        trmc.Network(topo)
    assert e.value.code == trmc.ERR_NO_DEVICE


def _tiny():
    return {"reach_seg_start": [0, 2, 5], "reach_seg_len": [2, 3, 4], "reach_up_ptr": [0, 0, 0, 2],
            "reach_up_idx": [0, 1], "level_ptr": [0, 2, 3], "level_reach": [0, 1, 2],
            "n_reaches": 3, "n_levels": 2, "n_segments": 9}


def _topo(d):
    return trmc.Topology(d["n_reaches"], d["n_segments"], d["n_levels"], 2,
                         np.array(d["reach_seg_start"]), np.array(d["reach_seg_len"]),
                         np.zeros(d["n_reaches"]), np.zeros(d["n_reaches"]), np.array(d["reach_up_ptr"]),
                         np.array(d["reach_up_idx"]), np.array(d["level_ptr"]), np.array(d["level_reach"]))


@pytest.mark.parametrize("mutate,code", [
    (lambda d: d.update(level_ptr=[0, 2, 2]), "TOPOLOGY"),                    # reach 2 never swept
    (lambda d: d.update(level_reach=[0, 0, 2]), "TOPOLOGY"),                  # reach listed twice
    (lambda d: d.update(level_ptr=[0, 3, 3], n_levels=2), "TOPOLOGY"),        # upstream in same level
    (lambda d: d.update(reach_seg_start=[0, 1, 5]), "TOPOLOGY"),              # overlapping segments
    (lambda d: d.update(reach_seg_len=[2, 3, 5]), "TOPOLOGY"),                # out of range
    (lambda d: d.update(reach_up_idx=[0, 7]), "TOPOLOGY"),                    # bad upstream id
])
def test_plan_rejects_invalid_topology(mutate, code):
    d = _tiny()
    mutate(d)
    with pytest.raises(trmc.TrmcError) as e:
        trmc.plan(_topo(d))
    assert e.value.code == getattr(trmc, "ERR_" + code)


def check_plan(topo, p, C):
    U, L, N = p["n_units"], p["n_levels"], topo.n_segments
    seg = p["unit_seg"]
    lens = np.diff(seg)
    assert lens.max() <= C
    # every covered segment appears exactly once in device order
    covered = np.concatenate([np.arange(s, s + n) for s, n in zip(topo.reach_seg_start, topo.reach_seg_len)])
    assert np.array_equal(np.sort(p["seg_of_dev"]), np.sort(covered))
    # level-sorted, contiguous levels
    assert np.all(np.diff(p["unit_level"]) >= 0)
    assert p["level_ptr"][0] == 0 and p["level_ptr"][-1] == U and len(p["level_ptr"]) == L + 1
    for l in range(L):
        assert np.all(p["unit_level"][p["level_ptr"][l]:p["level_ptr"][l + 1]] == l)
    # every edge goes strictly up a level; ring covers the widest span
    spans = []
    for u in range(U):
        for k in range(p["up_ptr"][u], p["up_ptr"][u + 1]):
            v = p["up_idx"][k]
            spans.append(p["unit_level"][u] - p["unit_level"][v])
    spans = np.array(spans) if spans else np.array([1])
    assert spans.min() >= 1 and p["ring"] == spans.max() + 2 == p["max_span"] + 2
    # internal units have exactly one upstream: the previous chunk of their reach
    internal = p["unit_flags"] & 1
    n_up = np.diff(p["up_ptr"])
    assert np.all(n_up[internal == 1] == 1)
    # ALAP: every unit with a downstream sits right below its earliest one
    down_min = np.full(U, L, np.int64)
    for u in range(U):
        for k in range(p["up_ptr"][u], p["up_ptr"][u + 1]):
            v = p["up_idx"][k]
            down_min[v] = min(down_min[v], p["unit_level"][u])
    has_down = down_min < L
    assert np.all(p["unit_level"][has_down] == down_min[has_down] - 1)
    assert np.all(p["unit_level"][~has_down] == L - 1)
    # packed records restate the arrays; bit 1 marks the units without slack
    # (longest path from a headwater, ASAP, equal to the ALAP level)
    rec = p["unit_rec"].reshape(-1, 8)
    assert rec.shape[0] == U
    assert np.array_equal(rec[:, 0], p["unit_level"])
    assert np.array_equal(rec[:, 1], seg[:-1]) and np.array_equal(rec[:, 2], seg[1:])
    assert np.array_equal(rec[:, 3] >> 8, n_up) and np.array_equal(rec[:, 3] & 1, internal)
    asap = np.zeros(U, np.int64)
    for u in range(U):                                # level-sorted: upstreams come first
        ups = p["up_idx"][p["up_ptr"][u]:p["up_ptr"][u + 1]]
        if len(ups):
            asap[u] = asap[ups].max() + 1
        assert np.array_equal(rec[u, 4:4 + len(ups)], ups)
    crit = (rec[:, 3] >> 1) & 1
    assert np.array_equal(crit == 1, asap == p["unit_level"])
    # a longest chain runs through every level
    assert set(p["unit_level"][crit == 1].tolist()) == set(range(L))


@pytest.mark.parametrize("kind,n,R", [(1, 1000, 0), (2, 30000, 9000), (3, 1, 6000)])
@pytest.mark.parametrize("C", [1, 2, 5, 1000])
def test_plan_invariants(kind, n, R, C):
    topo = trmc.generate_network(kind, n, R, 42)
    p = trmc.plan(topo, C, 24)
    check_plan(topo, p, C)
    if C >= 40:
        assert p["n_units"] == topo.n_reaches and p["n_levels"] == topo.n_levels
    assert p["super_levels"] == p["n_levels"] + 23


def test_plan_of_tree_needs_ring_of_three():
    p = trmc.plan(trmc.generate_network(2, 30000, 9000, 3), 1, 10)
    assert p["max_span"] == 1 and p["ring"] == 3


def test_generator_is_deterministic_and_well_formed():
    a = trmc.generate_network(2, 50000, 15000, 42)
    b = trmc.generate_network(2, 50000, 15000, 42)
    c = trmc.generate_network(2, 50000, 15000, 43)
    assert np.array_equal(a.reach_up_idx, b.reach_up_idx) and np.array_equal(a.reach_seg_len, b.reach_seg_len)
    assert not np.array_equal(a.reach_seg_len, c.reach_seg_len)
    assert a.n_segments == 50000 and a.n_reaches == 15000
    assert a.reach_seg_len.min() >= 1 and a.reach_seg_len.max() <= 40
    assert np.array_equal(a.reach_seg_start, np.concatenate([[0], np.cumsum(a.reach_seg_len)[:-1]]))
    for r in range(a.n_reaches):                     # topological ids, levels = 1 + max(upstream)
        ups = a.reach_up_idx[a.reach_up_ptr[r]:a.reach_up_ptr[r + 1]]
        assert np.all(ups < r)
        want = 0 if len(ups) == 0 else 1 + a.reach_level[ups].max()
        assert a.reach_level[r] == want
    assert a.max_upstream == 2
    p1 = trmc.generate_params(1000, 5)
    assert np.array_equal(p1, trmc.generate_params(1000, 5))
    assert np.all(p1 == p1.astype(np.float32).astype(np.float64))      # float-representable
    assert p1[:, 0].min() >= 500 and p1[:, 0].max() <= 5000
    assert p1[:, 4].min() >= 0.03 and p1[:, 4].max() <= 0.08
    assert p1[:, 7].min() >= float(np.float32(1e-4)) and p1[:, 7].max() <= float(np.float32(1e-2))  # clip, then float
    q = trmc.generate_forcing(20000, 12, 2, 9)
    assert q.shape == (2, 12, 20000) and q.min() > 0
    assert abs(q.mean() - 0.1) < 0.01                 # lognormal with mean 0.1
    assert not np.array_equal(q[0], q[1])             # members have their own streams


def test_merge_tree_generator_has_the_reference_generator_shape():
    """kind 3 follows the reference's gen_network.py construction: leaves =
    n_reaches / 6 merged two or three at a time, Poisson(2) chain reaches per
    link, max(1, Poisson(12)) segments per reach; at 1M target reaches the
    reference's own run (results/run_1m_4.txt:2) has 872,622 reaches,
    10,466,683 segments, 69 levels and confluences of at most 3."""
    a = trmc.generate_network(3, 1, 60000, 7)
    b = trmc.generate_network(3, 1, 60000, 7)
    assert np.array_equal(a.reach_up_idx, b.reach_up_idx) and np.array_equal(a.reach_seg_len, b.reach_seg_len)
    leaves = round(60000 / 6)
    assert int(np.sum(a.reach_n_up == 0)) == leaves
    assert a.max_upstream == 3 and set(np.unique(a.reach_n_up).tolist()) == {0, 1, 2, 3}
    n2, n3 = int(np.sum(a.reach_n_up == 2)), int(np.sum(a.reach_n_up == 3))
    assert n2 + 2 * n3 == leaves - 1                  # a tree: every merge removes k - 1 nodes
    assert 0.2 < n3 / (n2 + n3) < 0.35                 # three-way merges: 0.34 of the draws
    one = int(np.sum(a.reach_n_up == 1))              # chain reaches: Poisson(2) per link
    links = leaves + n2 + n3 - 1
    assert abs(one / links - 2.0) < 0.1
    assert abs(a.reach_seg_len.mean() - 12.0) < 0.2 and a.reach_seg_len.min() >= 1
    for r in range(a.n_reaches):
        ups = a.reach_up_idx[a.reach_up_ptr[r]:a.reach_up_ptr[r + 1]]
        assert np.all(ups < r) and np.all(np.diff(ups) > 0)
        assert a.reach_level[r] == (0 if len(ups) == 0 else 1 + a.reach_level[ups].max())
    assert a.n_levels < 60


def test_binary_formats_roundtrip(tmp_path):
    topo = trmc.generate_network(1, 1000, 0, 42)
    trmc.save_topo(str(tmp_path / "topo.bin"), topo)
    t2 = trmc.load_topo(str(tmp_path / "topo.bin"))
    for k in ("reach_seg_start", "reach_seg_len", "reach_level", "reach_n_up", "reach_up_ptr",
              "reach_up_idx", "level_ptr", "level_reach"):
        assert np.array_equal(getattr(topo, k), getattr(t2, k)), k
    # gen_network.py:292-310 layout: header of 4 int32, then the arrays
    raw = np.fromfile(tmp_path / "topo.bin", np.int32)
    assert raw[:4].tolist() == [topo.n_reaches, topo.n_segments, topo.n_levels, topo.max_upstream]
    p = trmc.generate_params(topo.n_segments, 1).astype(np.float32)
    trmc.save_params(str(tmp_path / "params.bin"), p)
    assert np.array_equal(trmc.load_params(str(tmp_path / "params.bin"), topo.n_segments), p)
    assert os.path.getsize(tmp_path / "params.bin") == topo.n_segments * 32
    q = trmc.generate_forcing(topo.n_segments, 5, 1, 2)[0].astype(np.float32)
    qup0 = np.arange(topo.n_reaches, dtype=np.float32)
    qdp0 = np.ones(topo.n_segments, np.float32)
    dp0 = np.full(topo.n_segments, 0.5, np.float32)
    trmc.save_forcings(str(tmp_path / "forcings.bin"), q, qup0, qdp0, dp0)
    a, b, c, d = trmc.load_forcings(str(tmp_path / "forcings.bin"), topo.n_reaches, topo.n_segments)
    assert np.array_equal(a, q) and np.array_equal(b, qup0) and np.array_equal(c, qdp0) and np.array_equal(d, dp0)
    with pytest.raises(trmc.TrmcError) as e:
        trmc.load_topo(str(tmp_path / "missing.bin"))
    assert e.value.code == trmc.ERR_IO


def test_compute_stats_matches_oracle_definition():
    orc = load_oracle()
    rng = np.random.default_rng(4)
    truth = rng.lognormal(0, 2, 5000)
    pred = truth * (1 + rng.normal(0, 0.03, truth.size))
    pred[::97] = np.inf
    assert trmc.compute_stats(pred, truth) == orc.compute_stats(pred, truth)


def test_golden_topology_plans():
    z = golden("wavefront_c1.npz")
    tv = TopoView({k: z[k] for k in z.files})
    for C in (1, 3, 64):
        check_plan(tv.t, trmc.plan(tv.t, C, 288), C)


def test_network_rejects_unknown_guess_model():
    trmc = load_trmc()
    t = trmc.generate_network(1, 50, 0, 1)
    with pytest.raises(trmc.TrmcError) as e:
        trmc.Network(t, mode=trmc.MODE_LEARNED, guess_model=2)
    assert e.value.code == trmc.ERR_ARG


def test_wrappers_refuse_short_arrays_before_device_work():
    """Arrays the C side would read past their end are argument errors raised
    by the wrappers (no device is needed to see it)."""
    trmc = load_trmc()
    w = trmc.make_workload(2, 2000, 4, n_reaches=600, seed=3)
    N, R = w.topo.n_segments, w.topo.n_reaches
    p32 = w.params.astype(np.float32)
    q = w.qlat[0]
    bad_calls = [
        lambda: trmc.wavefront_fp64(p32, w.topo, q[:3], w.qdp0, w.dp0, w.reach_last0, 4),      # T rows short
        lambda: trmc.wavefront_fp64(p32, w.topo, q, w.qdp0[:-1], w.dp0, w.reach_last0, 4),
        lambda: trmc.wavefront_f64(w.params, w.topo, q, w.qdp0, w.dp0, w.reach_last0[:-1], 4),
        lambda: trmc.linear_mc([np.zeros(N, np.float32)] * 3 + [np.zeros(N - 1, np.float32)], w.topo, q, w.qdp0, 4),
        lambda: trmc.linear_mc([np.zeros(N, np.float32)] * 4, w.topo, q[:2], w.qdp0, 4),
        lambda: trmc.matrix_build(w.topo.reach_up_ptr, w.topo.reach_up_idx, [np.zeros(R, np.float32)] * 3 +
                                  [np.zeros(R - 5, np.float32)]),
        lambda: trmc.matrix_route((np.array([0, 1, 2]), np.array([0, 1]), np.ones(2, np.float32)),
                                  (np.array([0, 1, 2]), np.array([0, 1]), np.ones(2, np.float32)),
                                  np.zeros(5, np.float32), np.zeros(2, np.float32), 3),           # qlat short
        lambda: trmc.matrix_route((np.array([0, 1, 2]), np.array([0, 1]), np.ones(2, np.float32)),
                                  (np.array([0, 1, 2]), np.array([0, 1]), np.ones(2, np.float32)),
                                  np.zeros(6, np.float32), np.zeros(1, np.float32), 3),           # Q0 short
        lambda: trmc.reach_parallel(w.params[:10], np.zeros(9), [0], [10], np.zeros(10), np.zeros(10),
                                    np.zeros(10)),
        lambda: trmc.reach_parallel(w.params[:10], np.zeros(10), [0], [11], np.zeros(10), np.zeros(10),
                                    np.zeros(10)),
        lambda: trmc.mc_solve(w.params[:3], np.zeros((4, 5))),
    ]
    for i, call in enumerate(bad_calls):
        with pytest.raises(trmc.TrmcError) as e:
            call()
        assert e.value.code == trmc.ERR_ARG, i
