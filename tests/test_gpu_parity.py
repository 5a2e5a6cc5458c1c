"""GPU parity of the HIP routing path against the reference's FP64 host
routine: golden vectors written by the reference itself, the oracle on the
same seeded inputs, and size-independent properties at the full workload.

Tolerance (REFERENCE stopping rule): |gpu - cpu| <= 1e-12 |cpu| + 1e-15 per
segment and, with recorded history, per reach outflow per timestep.  The only
arithmetic difference between the two sides is the last-ulp behaviour of pow
(ocml on the GPU, the host libm on the CPU)."""
import numpy as np
import pytest

from conftest import TopoView, assert_close, golden, load_oracle, load_trmc, max_rel, tiny_cases

pytestmark = pytest.mark.gpu
trmc = load_trmc()
orc = load_oracle()
RTOL = 1e-12


def _log(msg):
    print(msg, flush=True)


def test_mc_solve_batch_matches_reference_golden(gpu):
    z = golden("mc_solve_cases.npz")
    out, its = trmc.mc_solve(z["params"], z["inputs"], float(z["dt"]))
    assert_close(out, z["expected"], RTOL, 1e-15, "mc_solve")
    _, its_cpu = orc.mc_solve_many(z["params"], z["inputs"], float(z["dt"]))
    assert np.mean(its == its_cpu) > 0.999          # a secant branch flip would show here


def test_reach_parallel_matches_reference_golden(gpu):
    z = golden("reach_parallel.npz")
    qo, vo, dpo = trmc.reach_parallel(z["params"], z["ql"], z["rs"], z["rl"], z["qu"], z["qd"], z["dp"],
                                      float(z["uq"]), float(z["dt"]))
    assert_close(qo, z["qo"], RTOL, 1e-15, "qo")
    assert_close(vo, z["vo"], RTOL, 1e-15, "vo")
    assert_close(dpo, z["dpo"], RTOL, 1e-15, "dpo")


@pytest.mark.parametrize("T,qk,rk", [(8, "qdc_out_8", "reach_last_out_8"), (288, "qdc_out", "reach_last_out")])
def test_dropin_f32in_matches_reference_golden_c1(gpu, T, qk, rk):
    z = golden("wavefront_c1.npz")
    d = {k: z[k] for k in z.files}
    topo = TopoView(d).t
    _log(f"drop-in c1 T={T}")
    q, r = trmc.wavefront_fp64(d["params"], topo, d["qlat"][:T], d["qdp0"], d["dp0"], d["reach_last0"], T, 300.0)
    assert_close(q, d[qk], RTOL, 1e-15, "qdc")
    assert_close(r, d[rk], RTOL, 1e-15, "reach_last")


@pytest.mark.parametrize("name", sorted(tiny_cases()))
def test_dropin_tiny_networks(gpu, name):
    d = tiny_cases()[name]
    topo = TopoView(d).t
    T = int(d["n_timesteps"])
    q, r = trmc.wavefront_fp64(d["params"], topo, d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], T)
    assert_close(q, d["qdc_out"], RTOL, 1e-15, "qdc")
    assert_close(r, d["reach_last_out"], RTOL, 1e-15, "reach_last")


@pytest.fixture(scope="module")
def shreve():
    T = 48
    w = trmc.make_workload(2, 40000, T, n_reaches=12000, seed=21)
    ref = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T, history=True)
    return w, T, ref


@pytest.mark.parametrize("schedule,units", [(trmc.SCHED_SKEWED, 1), (trmc.SCHED_SKEWED, 2),
                                            (trmc.SCHED_SKEWED, 7), (trmc.SCHED_SKEWED, 64),
                                            (trmc.SCHED_LEVEL, 1), (trmc.SCHED_LEVEL, 64)])
def test_every_timestep_matches_oracle(gpu, shreve, schedule, units):
    w, T, ref = shreve
    _log(f"shreve 40k x {T} schedule={schedule} units={units}")
    with trmc.Network(w.topo, schedule=schedule, unit_segments=units, record_history=1) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        st = net.run(T)
        q, d, v, rl = net.state()
        hist = net.history(T)
    assert_close(hist[0], ref["reach_hist"], RTOL, 1e-15, "reach outflow per step")
    assert_close(q[0], ref["qdc"], RTOL, 1e-15, "qdc")
    assert_close(d[0], ref["depth"], RTOL, 1e-15, "depth")
    assert_close(v[0], ref["vel"], RTOL, 1e-15, "velocity")
    assert_close(rl[0], ref["reach_last"], RTOL, 1e-15, "reach_last")
    assert st["secant_iters"] > 0 and st["seg_steps"] == 40000 * T
    assert abs(st["secant_iters"] - ref["iterations"]) <= 1e-4 * ref["iterations"]


def test_results_do_not_depend_on_schedule_bitwise(gpu, shreve):
    w, T, _ = shreve
    outs = []
    for sched, units, block in ((trmc.SCHED_SKEWED, 1, 256), (trmc.SCHED_SKEWED, 5, 64),
                                (trmc.SCHED_LEVEL, 3, 128)):
        with trmc.Network(w.topo, schedule=sched, unit_segments=units, block_size=block) as net:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            net.set_forcing(w.qlat)
            net.run(T)
            outs.append(net.state())
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_chained_runs_and_reset(gpu, shreve):
    w, T, ref = shreve
    with trmc.Network(w.topo) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        net.run(17)
        net.run(T - 17)
        a = net.state()
        net.reset()
        net.run(T)
        b = net.state()
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    assert_close(a[0][0], ref["qdc"], RTOL, 1e-15, "chained qdc")


def test_ensemble_members_route_independently(gpu):
    T, M = 24, 4
    w = trmc.make_workload(2, 20000, T, n_reaches=6000, n_members=M, seed=5)
    with trmc.Network(w.topo, n_members=M) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        st = net.run(T)
        q, d, v, rl = net.state()
    assert st["seg_steps"] == 20000 * T * M
    for m in range(M):
        ref = orc.wavefront_f64(w.topo, w.params, w.qlat[m], w.qdp0, w.dp0, w.reach_last0, T)
        assert_close(q[m], ref["qdc"], RTOL, 1e-15, f"member {m} qdc")
        assert_close(rl[m], ref["reach_last"], RTOL, 1e-15, f"member {m} reach_last")


def test_f64_dropin_matches_oracle(gpu):
    T = 30
    w = trmc.make_workload(1, 1000, T, seed=77)
    q, r = trmc.wavefront_f64(w.params, w.topo, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T)
    ref = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T)
    assert_close(q, ref["qdc"], RTOL, 1e-15, "qdc")
    assert_close(r, ref["reach_last"], RTOL, 1e-15, "reach_last")


def test_edge_cases(gpu):
    """Zero timesteps return the initial state; an empty network is a no-op;
    invalid topology is an error, not a crash."""
    d = tiny_cases()["y"]
    topo = TopoView(d).t
    q, r = trmc.wavefront_fp64(d["params"], topo, d["qlat"][:0], d["qdp0"], d["dp0"], d["reach_last0"], 0)
    assert np.array_equal(q, d["qdp0"].astype(np.float64))
    assert np.array_equal(r, d["reach_last0"].astype(np.float64))
    empty = trmc.Topology(0, 0, 0, 0, *[np.zeros(0, np.int32)] * 4, np.zeros(1, np.int32),
                          np.zeros(0, np.int32), np.zeros(1, np.int32), np.zeros(0, np.int32))
    q, r = trmc.wavefront_fp64(np.zeros((0, 8), np.float32), empty, np.zeros((3, 0), np.float32),
                               np.zeros(0), np.zeros(0), np.zeros(0), 3)
    assert q.size == 0 and r.size == 0
    bad = TopoView(d).t
    bad.level_ptr = np.array([0, 3, 3], np.int32)      # upstream and downstream in one level
    with pytest.raises(trmc.TrmcError) as e:
        trmc.wavefront_fp64(d["params"], bad, d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], 2)
    assert e.value.code == trmc.ERR_TOPOLOGY


def test_full_size_network_first_steps_match_oracle(gpu):
    """1,000,000-segment Shreve tree (the benchmark workload): the first 3 steps
    against the oracle, all 288 steps size-independently (bitwise equal across
    unit sizes, finite, non-negative, deterministic)."""
    T = 288
    w = trmc.make_workload(2, 1_000_000, T, n_reaches=300_000, seed=42)
    _log(f"full-size network: levels {w.topo.n_levels}")
    ref = orc.wavefront_f64(w.topo, w.params, w.qlat[0][:3], w.qdp0, w.dp0, w.reach_last0, 3)
    states = []
    with trmc.Network(w.topo, unit_segments=1) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        net.run(3)
        q3 = net.state()
        assert_close(q3[0][0], ref["qdc"], RTOL, 1e-15, "qdc after 3 steps")
        assert_close(q3[3][0], ref["reach_last"], RTOL, 1e-15, "reach_last after 3 steps")
        net.reset()
        st = net.run(T)
        states.append(net.state())
        net.reset()
        net.run(T)
        states.append(net.state())
    _log(f"full-size 288 steps: {st['kernel_ms']:.1f} ms")
    with trmc.Network(w.topo, unit_segments=4) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        net.run(T)
        states.append(net.state())
    for s in states[1:]:
        for a, b in zip(states[0], s):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    q = states[0][0][0]
    assert np.all(np.isfinite(q)) and q.min() >= 0.0
