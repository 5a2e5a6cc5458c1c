"""GPU parity of the HIP routing path against the reference's FP64 host
routine, on the same seeded inputs.

Two CPU sides, one algorithm:
  * oracle "cr"  -- the reference's arithmetic with a correctly rounded pow
                    (oracle/liborc_cr.so).  This is exactly what the GPU
                    implements, so the bar is |gpu - cpu| <= 1e-12 |cpu| + 1e-15
                    per segment and per reach outflow per timestep, and at
                    least 99.99 % of the values bitwise equal.
  * reference    -- cpu_wavefront_fp64 / mc_solve_d / cpu_mc themselves (golden
                    vectors written by them, and oracle "libm", which is bitwise
                    the same).  They call the host libm's pow, which is not
                    correctly rounded (about 1 in 1000 calls is one ulp off),
                    and the loose-tolerance secant can amplify a one-ulp change
                    into a different iterate.  The bar here: every difference
                    between the GPU and the reference is the difference the
                    reference itself shows when only its pow is made correctly
                    rounded (elementwise, to 1e-12).
"""
import numpy as np
import pytest

from conftest import TopoView, assert_close, golden, load_oracle, load_trmc, max_rel, tiny_cases

pytestmark = pytest.mark.gpu
trmc = load_trmc()
orc = load_oracle()
RTOL = 1e-12


def _log(msg):
    print(msg, flush=True)


def check(got, want_cr, what, want_ref=None):
    got = np.asarray(got, np.float64)
    want_cr = np.asarray(want_cr, np.float64)
    assert_close(got, want_cr, RTOL, 1e-15, what + " vs oracle (correctly rounded pow)")
    same = float(np.mean(got.view(np.uint64) == want_cr.view(np.uint64))) if got.size else 1.0
    assert same >= 0.9999, f"{what}: only {same:.6f} bitwise equal"
    if want_ref is not None:
        want_ref = np.asarray(want_ref, np.float64)
        d_gpu = np.abs(got - want_ref)
        d_cr = np.abs(want_cr - want_ref)
        assert np.all(d_gpu <= d_cr + RTOL * np.abs(want_ref) + 1e-15), \
            f"{what}: GPU departs from the reference beyond the reference's own pow sensitivity"
        _log(f"{what}: bitwise {same:.6f}; vs reference max rel {max_rel(got, want_ref):.3e} "
             f"(reference with correctly rounded pow: {max_rel(want_cr, want_ref):.3e}), "
             f"within 1e-12 of reference: {float(np.mean(d_gpu <= RTOL * np.abs(want_ref) + 1e-15)):.6f}")


def test_mc_solve_batch_matches_reference(gpu):
    z = golden("mc_solve_cases.npz")
    out, its = trmc.mc_solve(z["params"], z["inputs"], float(z["dt"]))
    want, its_cr = orc.mc_solve_many(z["params"], z["inputs"], float(z["dt"]), variant="cr")
    check(out, want, "mc_solve", z["expected"])
    assert np.array_equal(its, its_cr)


def test_reach_parallel_matches_reference(gpu):
    z = golden("reach_parallel.npz")
    args = (z["params"], z["ql"], z["rs"], z["rl"], z["qu"], z["qd"], z["dp"], float(z["uq"]), float(z["dt"]))
    qo, vo, dpo = trmc.reach_parallel(*args)
    cq, cv, cd = orc.reach_parallel(*args, variant="cr")
    check(qo, cq, "reach_parallel qo", z["qo"])
    check(vo, cv, "reach_parallel vo", z["vo"])
    check(dpo, cd, "reach_parallel dpo", z["dpo"])


@pytest.mark.parametrize("T,qk,rk", [(8, "qdc_out_8", "reach_last_out_8"), (288, "qdc_out", "reach_last_out")])
def test_dropin_f32in_matches_reference_c1(gpu, T, qk, rk):
    z = golden("wavefront_c1.npz")
    d = {k: z[k] for k in z.files}
    topo = TopoView(d).t
    _log(f"drop-in c1 T={T}")
    q, r = trmc.wavefront_fp64(d["params"], topo, d["qlat"][:T], d["qdp0"], d["dp0"], d["reach_last0"], T, 300.0)
    cq, cr, _ = orc.wavefront_f32in(d, d["params"], d["qlat"][:T], d["qdp0"], d["dp0"], d["reach_last0"], T,
                                    variant="cr")
    check(q, cq, f"c1 T={T} qdc", d[qk])
    check(r, cr, f"c1 T={T} reach_last", d[rk])


@pytest.mark.parametrize("name", sorted(tiny_cases()))
def test_dropin_tiny_networks(gpu, name):
    d = tiny_cases()[name]
    topo = TopoView(d).t
    T = int(d["n_timesteps"])
    q, r = trmc.wavefront_fp64(d["params"], topo, d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], T)
    cq, cr, _ = orc.wavefront_f32in(d, d["params"], d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], T,
                                    variant="cr")
    check(q, cq, f"{name} qdc", d["qdc_out"])
    check(r, cr, f"{name} reach_last", d["reach_last_out"])


@pytest.fixture(scope="module")
def shreve():
    T = 48
    w = trmc.make_workload(2, 40000, T, n_reaches=12000, seed=21)
    ref = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T, history=True)
    cr = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T, history=True,
                           variant="cr")
    return w, T, ref, cr


@pytest.mark.parametrize("schedule,units,variant", [
    (trmc.SCHED_DATAFLOW, 1, 0), (trmc.SCHED_DATAFLOW, 3, 0), (trmc.SCHED_DATAFLOW, 64, 0),
    (trmc.SCHED_DATAFLOW, 1, 1), (trmc.SCHED_DATAFLOW, 4, 1),
    (trmc.SCHED_DATAFLOW, 2, 0), (trmc.SCHED_DATAFLOW, 4, 0),
    (trmc.SCHED_SKEWED, 1, 0), (trmc.SCHED_SKEWED, 2, 0), (trmc.SCHED_SKEWED, 7, 0),
    (trmc.SCHED_SKEWED, 64, 0), (trmc.SCHED_LEVEL, 1, 0), (trmc.SCHED_LEVEL, 64, 0)])
def test_every_timestep_matches_reference(gpu, shreve, schedule, units, variant):
    w, T, ref, cr = shreve
    _log(f"shreve 40k x {T} schedule={schedule} units={units} variant={variant}")
    with trmc.Network(w.topo, schedule=schedule, unit_segments=units, record_history=1,
                      dataflow_variant=variant) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        st = net.run(T)
        q, d, v, rl = net.state()
        hist = net.history(T)
    check(hist[0], cr["reach_hist"], "reach outflow per step", ref["reach_hist"])
    check(q[0], cr["qdc"], "qdc", ref["qdc"])
    check(d[0], cr["depth"], "depth", ref["depth"])
    check(v[0], cr["vel"], "velocity", ref["vel"])
    check(rl[0], cr["reach_last"], "reach_last", ref["reach_last"])
    assert st["seg_steps"] == 40000 * T
    assert st["secant_iters"] == cr["iterations"]


def test_results_do_not_depend_on_schedule_bitwise(gpu, shreve):
    w, T, _, _ = shreve
    outs = []
    for sched, units, block, var in ((trmc.SCHED_DATAFLOW, 1, 256, 0), (trmc.SCHED_SKEWED, 5, 64, 0),
                                     (trmc.SCHED_LEVEL, 3, 128, 0), (trmc.SCHED_DATAFLOW, 2, 64, 1),
                                     (trmc.SCHED_DATAFLOW, 6, 128, 0), (trmc.SCHED_DATAFLOW, 2, 256, 0)):
        with trmc.Network(w.topo, schedule=sched, unit_segments=units, block_size=block,
                          dataflow_variant=var) as net:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            net.set_forcing(w.qlat)
            net.run(T)
            outs.append(net.state())
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("variant", [0, 1])
def test_chained_runs_and_reset(gpu, shreve, variant):
    w, T, ref, cr = shreve
    with trmc.Network(w.topo, dataflow_variant=variant) as net:
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
    check(a[0][0], cr["qdc"], "chained qdc")


@pytest.mark.parametrize("schedule,variant", [(trmc.SCHED_DATAFLOW, 0), (trmc.SCHED_SKEWED, 0),
                                              (trmc.SCHED_DATAFLOW, 1)])
def test_ensemble_members_route_independently(gpu, schedule, variant):
    T, M = 24, 3
    w = trmc.make_workload(2, 12000, T, n_reaches=3600, n_members=M, seed=5)
    with trmc.Network(w.topo, n_members=M, schedule=schedule, dataflow_variant=variant) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        st = net.run(T)
        q, d, v, rl = net.state()
    assert st["seg_steps"] == 12000 * T * M
    for m in range(M):
        cr = orc.wavefront_f64(w.topo, w.params, w.qlat[m], w.qdp0, w.dp0, w.reach_last0, T, variant="cr")
        check(q[m], cr["qdc"], f"member {m} qdc")
        check(rl[m], cr["reach_last"], f"member {m} reach_last")


def test_f64_dropin_matches_reference(gpu):
    T = 30
    w = trmc.make_workload(1, 1000, T, seed=77)
    q, r = trmc.wavefront_f64(w.params, w.topo, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T)
    ref = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T)
    cr = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T, variant="cr")
    check(q, cr["qdc"], "f64 qdc", ref["qdc"])
    check(r, cr["reach_last"], "f64 reach_last", ref["reach_last"])


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


def test_full_size_network(gpu):
    """1,000,000-segment Shreve tree (the benchmark workload): the first 2 steps
    against the oracle, all 288 steps size-independently (bitwise equal across
    unit sizes and repeated runs, finite, non-negative)."""
    T = 288
    w = trmc.make_workload(2, 1_000_000, T, n_reaches=300_000, seed=42)
    _log(f"full-size network: levels {w.topo.n_levels}")
    cr = orc.wavefront_f64(w.topo, w.params, w.qlat[0][:2], w.qdp0, w.dp0, w.reach_last0, 2, variant="cr")
    states = []
    with trmc.Network(w.topo, unit_segments=1) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        net.run(2)
        s2 = net.state()
        check(s2[0][0], cr["qdc"], "1M qdc after 2 steps")
        check(s2[3][0], cr["reach_last"], "1M reach_last after 2 steps")
        net.reset()
        st = net.run(T)
        states.append(net.state())
        net.reset()
        net.run(T)
        states.append(net.state())
    _log(f"full-size 288 steps: {st['kernel_ms']:.1f} ms")
    with trmc.Network(w.topo, unit_segments=4, schedule=trmc.SCHED_SKEWED) as net:
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


@pytest.mark.parametrize("mode,extra,schedule", [(trmc.MODE_REFERENCE, {}, trmc.SCHED_DATAFLOW),
                                                 (trmc.MODE_REFERENCE, {"dataflow_variant": 1}, trmc.SCHED_DATAFLOW),
                                                 (trmc.MODE_REFERENCE, {}, trmc.SCHED_SKEWED)])
def test_mapped_forcing_equals_resident_forcing(gpu, shreve, mode, extra, schedule):
    """The forcing read in place from page-locked host memory while the
    routing runs gives the same bits as the forcing resident in HBM."""
    w, T, _, _ = shreve
    _log(f"mapped forcing mode={mode} schedule={schedule}")
    outs = []
    for mapped in (False, True):
        with trmc.Network(w.topo, mode=mode, record_history=1, schedule=schedule, **extra) as net:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            if mapped:
                net.set_forcing_mapped(net.order_forcing(w.qlat))
            else:
                net.set_forcing(w.qlat)
            st = net.run(T)
            outs.append((net.state(), net.history(T), st["secant_iters"]))
    (a, ha, ia), (b, hb, ib) = outs
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    assert np.array_equal(ha.view(np.uint64), hb.view(np.uint64)) and ia == ib


def test_mapped_forcing_members_chained_and_errors(gpu):
    T, M = 12, 3
    w = trmc.make_workload(2, 6000, T, n_reaches=1800, n_members=M, seed=9)
    with trmc.Network(w.topo, n_members=M) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        qd = net.order_forcing(w.qlat)
        assert qd.array().shape == (T, net.info(T)["n_dev_segments"], M)
        net.set_forcing_mapped(qd)
        del qd                                   # the network keeps the buffer alive
        net.run(5)
        net.run(T - 5)
        q, d, v, rl = net.state()
        with pytest.raises(trmc.TrmcError) as e:
            net.set_forcing_mapped(np.zeros((T, 10, M)))
        assert e.value.code == trmc.ERR_ARG
    with trmc.Network(w.topo, n_members=M, mode=trmc.MODE_CONVERGED) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing_mapped(net.order_forcing(w.qlat))
        with pytest.raises(trmc.TrmcError) as e:      # refused for the converged rules
            net.run(T)
        assert e.value.code == trmc.ERR_ARG
    for m in range(M):
        cr = orc.wavefront_f64(w.topo, w.params, w.qlat[m], w.qdp0, w.dp0, w.reach_last0, T, variant="cr")
        check(q[m], cr["qdc"], f"mapped member {m} qdc")
        check(rl[m], cr["reach_last"], f"mapped member {m} reach_last")
