"""CONVERGED and LEARNED modes on the GPU against the FP64 oracle.

Both solve the fixed point of the reference iteration and end every root in
its canonical cell, as orc_wavefront_f64 in converged mode does: every
segment's outflow, depth and velocity and every reach outflow of every step
are the oracle's bits (correctly rounded pow for the velocity), for every
schedule, and LEARNED lands on the same bits as CONVERGED with fewer
residual evaluations."""
import numpy as np
import pytest

from conftest import load_oracle, load_trmc, max_rel

pytestmark = pytest.mark.gpu
trmc = load_trmc()
orc = load_oracle()
TOL = 1e-9


def _log(msg):
    print(msg, flush=True)


def close(got, want, what):
    """The oracle's bits (a failure reports how many differ and by how much)."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    diff = got.view(np.uint64) != want.view(np.uint64)
    bad = np.abs(got - want) > TOL * np.abs(want) + 1e-12
    assert not diff.any(), (f"{what}: {int(diff.sum())} of {diff.size} differ, {int(bad.sum())} beyond 1e-9, "
                            f"worst rel {max_rel(got, want):.3e}")


@pytest.fixture(scope="module")
def case():
    T = 12
    w = trmc.make_workload(2, 10000, T, n_reaches=3000, seed=13)
    want = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, T,
                             mode=trmc.MODE_CONVERGED, history=True, variant="cr")
    return w, T, want


@pytest.fixture(scope="module")
def weights(gpu):
    trmc.mlp_make_dataset(1 << 16, seed=7)
    W, loss = trmc.mlp_train(trmc.mlp_init_weights(3), epochs=30, batch=512, lr=3e-3)
    _log(f"trained: loss {loss[0]:.4g} -> {loss[-1]:.4g}")
    return W


def _run(w, T, mode, W=None, **opt):
    with trmc.Network(w.topo, mode=mode, record_history=1, **opt) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        if W is not None:
            net.set_mlp(W)
        st = net.run(T)
        return net.state(), net.history(T), st


SCHEDULES = [(trmc.SCHED_DATAFLOW, 4, 0), (trmc.SCHED_DATAFLOW, 1, 0), (trmc.SCHED_DATAFLOW, 3, 1),
             (trmc.SCHED_SKEWED, 4, 0), (trmc.SCHED_LEVEL, 2, 0), (trmc.SCHED_DATAFLOW, 2, 0)]


@pytest.mark.parametrize("schedule,units,variant", SCHEDULES)
def test_converged_matches_oracle(gpu, case, schedule, units, variant):
    w, T, want = case
    _log(f"converged schedule={schedule} units={units} variant={variant}")
    (q, d, v, rl), hist, st = _run(w, T, trmc.MODE_CONVERGED, schedule=schedule, unit_segments=units,
                                   dataflow_variant=variant)
    close(hist[0], want["reach_hist"], "reach outflow per step")
    close(q[0], want["qdc"], "qdc")
    close(d[0], want["depth"], "depth")
    close(v[0], want["vel"], "velocity")
    assert st["not_converged"] == 0 and st["noncanonical"] == 0
    _log(f"  evaluations per solve {st['secant_iters'] / st['seg_steps']:.3f}")


@pytest.mark.parametrize("schedule,units,variant", SCHEDULES)
def test_learned_matches_oracle_with_fewer_evaluations(gpu, case, weights, schedule, units, variant):
    w, T, want = case
    _log(f"learned schedule={schedule} units={units} variant={variant}")
    (q, d, v, rl), hist, st = _run(w, T, trmc.MODE_LEARNED, weights, schedule=schedule, unit_segments=units,
                                   dataflow_variant=variant)
    close(hist[0], want["reach_hist"], "reach outflow per step")
    close(q[0], want["qdc"], "qdc")
    close(d[0], want["depth"], "depth")
    assert st["not_converged"] == 0 and st["noncanonical"] == 0
    _, _, stc = _run(w, T, trmc.MODE_CONVERGED, schedule=schedule, unit_segments=units,
                     dataflow_variant=variant)
    el = st["secant_iters"] / st["seg_steps"]
    ec = stc["secant_iters"] / stc["seg_steps"]
    _log(f"  evaluations per solve: learned {el:.3f}, converged {ec:.3f}, fallbacks {st['fallbacks']}")
    assert el < 0.8 * ec
    assert st["fallbacks"] <= 1e-3 * st["seg_steps"]


def test_learned_with_untrained_weights_still_converges(gpu, case):
    """A useless guess costs evaluations (and falls back), never accuracy."""
    w, T, want = case
    W = trmc.mlp_init_weights(1)
    W[1248:1280] *= 30.0                            # a rating curve that swings by e^+-30
    (q, d, v, rl), hist, st = _run(w, T, trmc.MODE_LEARNED, W)
    close(q[0], want["qdc"], "qdc")
    close(hist[0], want["reach_hist"], "reach outflow per step")
    _log(f"  untrained: evaluations per solve {st['secant_iters'] / st['seg_steps']:.3f}, "
         f"fallbacks {st['fallbacks']}")


def test_power_law_guess_runs_without_weights(gpu, case):
    """guess_model 1 (the anchored guess with the exponent 3/5 alone) needs
    no network: no weights are set, and the kernel must not read them."""
    w, T, want = case
    (q, d, v, rl), hist, st = _run(w, T, trmc.MODE_LEARNED, None, guess_model=1)
    close(q[0], want["qdc"], "qdc")
    close(hist[0], want["reach_hist"], "reach outflow per step")
    assert st["mlp_calls"] == 0


def test_learned_needs_weights(gpu, case):
    w, T, _ = case
    with trmc.Network(w.topo, mode=trmc.MODE_LEARNED) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        with pytest.raises(trmc.TrmcError) as e:
            net.run(1)
        assert e.value.code == trmc.ERR_STATE
        with pytest.raises(trmc.TrmcError):
            net.set_mlp(np.full(1281, np.nan, np.float32))


@pytest.mark.parametrize("variant", [0, 1])
def test_learned_ensemble_members(gpu, weights, variant):
    T, M = 6, 3
    w = trmc.make_workload(2, 4000, T, n_reaches=1200, n_members=M, seed=31)
    with trmc.Network(w.topo, mode=trmc.MODE_LEARNED, n_members=M, dataflow_variant=variant) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        net.set_mlp(weights)
        net.run(T)
        q, d, v, rl = net.state()
    for m in range(M):
        want = orc.wavefront_f64(w.topo, w.params, w.qlat[m], w.qdp0, w.dp0, w.reach_last0, T,
                                 mode=trmc.MODE_CONVERGED, variant="cr")
        close(q[m], want["qdc"], f"member {m} qdc")
        close(rl[m], want["reach_last"], f"member {m} reach_last")

