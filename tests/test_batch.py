"""Many independent networks routed as one: the disjoint union
(trmc_topo_concat) and ensemble members, against each network routed alone
by the oracle; device-generated forcing against the host generator."""
import numpy as np
import pytest

from conftest import load_oracle, load_trmc

trmc = load_trmc()
orc = load_oracle()


def _nets(sizes, seed=7):
    return [trmc.generate_network(2, int(n), int(0.3 * n), seed + i) for i, n in enumerate(sizes)]


def test_concat_renumbers_and_keeps_levels():
    nets = _nets([300, 50, 1200, 7])
    u, so, ro = trmc.concat_topologies(nets)
    assert u.n_segments == sum(t.n_segments for t in nets) and so[-1] == u.n_segments
    assert u.n_reaches == sum(t.n_reaches for t in nets) and ro[-1] == u.n_reaches
    assert u.n_levels == max(t.n_levels for t in nets)
    for i, t in enumerate(nets):
        r = slice(ro[i], ro[i + 1])
        assert np.array_equal(u.reach_seg_start[r], t.reach_seg_start + so[i])
        assert np.array_equal(u.reach_seg_len[r], t.reach_seg_len)
        b, e = u.reach_up_ptr[ro[i]], u.reach_up_ptr[ro[i + 1]]
        assert np.array_equal(u.reach_up_idx[b:e], t.reach_up_idx + ro[i])
    # every reach once, upstream reaches in earlier levels: the plan accepts it
    assert sorted(u.level_reach.tolist()) == list(range(u.n_reaches))
    info = trmc.plan(u, unit_segments=4, n_timesteps=8)
    assert info["n_units"] > 0


def test_concat_empty_and_single():
    one = _nets([100])
    u, so, ro = trmc.concat_topologies(one)
    for f in ("reach_seg_start", "reach_seg_len", "reach_up_ptr", "reach_up_idx", "level_ptr", "level_reach"):
        assert np.array_equal(getattr(u, f), getattr(one[0], f)), f
    u, so, ro = trmc.concat_topologies([])
    assert u.n_segments == 0 and u.n_reaches == 0


def test_union_routes_like_separate_networks_on_the_oracle():
    """The oracle on the union equals the oracle on each network (no coupling,
    same level order within each network)."""
    sizes = [400, 90, 700]
    nets = _nets(sizes, seed=3)
    u, so, ro = trmc.concat_topologies(nets)
    T = 6
    params = trmc.generate_params(u.n_segments, 11)
    qdp0, dp0 = trmc.generate_state(params)
    qlat = trmc.generate_forcing(u.n_segments, T, 1, 12)[0]
    rl0 = trmc.reach_last0_from(u, qdp0)
    whole = orc.wavefront_f64(u, params, qlat, qdp0, dp0, rl0, T)
    for i, t in enumerate(nets):
        s, r = slice(so[i], so[i + 1]), slice(ro[i], ro[i + 1])
        part = orc.wavefront_f64(t, params[s], qlat[:, s], qdp0[s], dp0[s], rl0[r], T)
        assert np.array_equal(part["qdc"], whole["qdc"][s])
        assert np.array_equal(part["reach_last"], whole["reach_last"][r])


def test_batch_sizes_are_log_uniform_and_seeded():
    a = trmc.batch_sizes(64, 10_000, 200_000, 5)
    assert np.array_equal(a, trmc.batch_sizes(64, 10_000, 200_000, 5))
    assert a.min() >= 10_000 and a.max() <= 200_000
    assert abs(np.mean(np.log(a)) - 0.5 * (np.log(1e4) + np.log(2e5))) < 0.3


@pytest.mark.gpu
def test_device_forcing_equals_host_forcing(gpu):
    T, M = 5, 3
    w = trmc.make_workload(2, 3000, T, n_reaches=900, n_members=M, seed=4)
    states = []
    for dev in (False, True):
        with trmc.Network(w.topo, n_members=M) as net:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            if dev:
                net.generate_forcing(T, 4 ^ 0xA1A1)
            else:
                net.set_forcing(w.qlat)
            net.run(T)
            states.append(net.state())
    for a, b in zip(*states):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_batched_networks_and_members_match_oracle(gpu, mode):
    """7 networks of uneven size x 3 members in one launch against each
    (network, member) routed alone by the oracle."""
    T, M = 8, 3
    sizes = [2500, 300, 6000, 40, 1200, 800, 4000]
    w, so, ro = trmc.make_batch_workload(sizes, T, M, seed=9)
    with trmc.Network(w.topo, n_members=M, mode=mode) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        st = net.run(T)
        q, d, v, rl = net.state()
    assert st["seg_steps"] == w.topo.n_segments * T * M
    nets = [trmc.generate_network(2, int(n), int(0.3 * n), 9 + i) for i, n in enumerate(sizes)]
    for i, t in enumerate(nets):
        s, r = slice(so[i], so[i + 1]), slice(ro[i], ro[i + 1])
        for m in range(M):
            want = orc.wavefront_f64(t, w.params[s], w.qlat[m][:, s], w.qdp0[s], w.dp0[s], w.reach_last0[r], T,
                                     mode=mode, variant="cr")
            tol = 1e-12 if mode == 0 else 1e-9
            for got, ref in ((q[m][s], want["qdc"]), (rl[m][r], want["reach_last"])):
                assert np.all(np.abs(got - ref) <= tol * np.abs(ref) + 1e-15), (i, m)


@pytest.mark.gpu
def test_autotuner_picks_a_candidate_and_results_do_not_change(gpu):
    T, M = 6, 2
    w, so, ro = trmc.make_batch_workload([3000, 500, 1500], T, M, seed=21)
    best, ms = trmc.autotune(w.topo, w.params, w.qdp0, w.dp0, w.reach_last0, n_members=M, n_probe=4,
                             qlat=w.qlat)
    assert best["unit_segments"] in (2, 4, 8) and best["batch_iters"] in (6, 8, 12)
    assert (best["schedule"], best["dataflow_variant"]) in ((trmc.SCHED_DATAFLOW, 0), (trmc.SCHED_DATAFLOW, 1),
                                                            (trmc.SCHED_SKEWED, 0))
    assert best["block_size"] in (64, 128, 256) and ms > 0
    outs = []
    for opts in ({}, {k: best[k] for k in ("unit_segments", "batch_iters", "block_size", "schedule",
                                           "dataflow_variant")}):
        with trmc.Network(w.topo, n_members=M, **opts) as net:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            net.set_forcing(w.qlat)
            net.run(T)
            outs.append(net.state())
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [trmc.SCHED_DATAFLOW, trmc.SCHED_SKEWED])
def test_autotuner_keeps_a_given_schedule(gpu, schedule):
    T, M = 4, 2
    w, so, ro = trmc.make_batch_workload([2000, 700], T, M, seed=22)
    best, _ = trmc.autotune(w.topo, w.params, w.qdp0, w.dp0, w.reach_last0, n_members=M, n_probe=3,
                            qlat=w.qlat, schedule=schedule)
    assert best["schedule"] == schedule
