"""The benchmark configurations at their configured sizes on the GPU.

C4  the 1,000,000-segment Shreve tree (300,000 reaches), 288 steps, LEARNED
    with the network trained on 10M pairs: every reach outflow of every step
    and every final segment flow and depth against the FP64 oracle
    (CONVERGED rule, level-parallel on the host threads), and the GPU's own
    CONVERGED run beside it.
C5  the same network x 16 lateral-inflow members, LEARNED, forcing generated
    on the device: properties on every member, and the oracle on sampled
    sub-networks (upstream closures of random reaches, routed alone) of every
    member, every step.
C3  64 independent Shreve networks of 1e4..2e5 segments (log-uniform) x 8
    members x 288 steps in one launch, REFERENCE rule: properties on all, and
    three whole networks x two members bit for bit against the oracle with
    the correctly rounded pow.

The bar of the converged rules: the oracle's bits.  LEARNED and CONVERGED
end every root in its canonical cell (neighbouring depths with the low 28
significand bits zero and a sign change of a residual made of IEEE operations
only), as the oracle does, so every reach outflow of every step and every
final segment flow must equal the oracle's, whatever start the solve took.
A failure reports how many values differ and by how much.
"""
import os
import time

import numpy as np
import pytest

from conftest import flow_scale_error, load_oracle, load_trmc, sample_roots, subnetwork

pytestmark = pytest.mark.gpu
trmc = load_trmc()
orc = load_oracle()
THREADS = max(1, min(16, os.cpu_count() or 1))
SEED = 42


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def net1m():
    return trmc.make_workload(2, 1_000_000, 1, n_reaches=300_000, seed=SEED, forcing=False)


@pytest.fixture(scope="module")
def weights10m(gpu):
    t0 = time.perf_counter()
    trmc.mlp_make_dataset(10_000_000, seed=7)
    W, loss = trmc.mlp_train(trmc.mlp_init_weights(3), epochs=8, batch=1024, lr=3e-3)
    _log(f"network trained on 10M pairs: loss {loss[0]:.4g} -> {loss[-1]:.4g} ({time.perf_counter() - t0:.1f} s)")
    return W


def _route(w, mode, T, M, W=None, forcing=None, units=2):
    with trmc.Network(w.topo, mode=mode, n_members=M, record_history=1, unit_segments=units) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        if forcing is None:
            net.generate_forcing(T, SEED ^ 0xA1A1)
        else:
            net.set_forcing(forcing)
        if W is not None:
            net.set_mlp(W)
        st = net.run(T)
        q, d, v, rl = net.state()
        return dict(q=q, d=d, v=v, rl=rl, hist=net.history(T), stats=st)


def _judge(what, got_hist, want_hist, reach_last0, topo):
    got = np.ascontiguousarray(got_hist, np.float64)
    want = np.ascontiguousarray(want_hist, np.float64)
    n_diff = int((got.view(np.uint64) != want.view(np.uint64)).sum())
    fs, rel = flow_scale_error(got, want, reach_last0, topo)
    n_rel = int((rel > 1e-9).sum())
    _log(f"{what}: {n_diff} of {rel.size} reach-steps not bitwise, {n_rel} beyond 1e-9 relative "
         f"({n_rel / rel.size * 1e6:.2f} per million), max relative {rel.max():.3e}; "
         f"flow-scale error max {fs.max():.3e}")
    assert n_diff == 0, f"{what}: {n_diff} reach-steps differ, {n_rel} beyond 1e-9, max relative {rel.max():.3e}"
    return fs, rel


def test_c4_learned_1m_288_every_reach_outflow_against_the_oracle(net1m, weights10m):
    T = 288
    w = net1m
    ql = trmc.generate_forcing_member(w.topo.n_segments, T, 0, SEED ^ 0xA1A1)
    _log("C4: LEARNED 1M x 288 on the GPU")
    L = _route(w, trmc.MODE_LEARNED, T, 1, W=weights10m)
    P = _route(w, trmc.MODE_CONVERGED, T, 1)
    st = L["stats"]
    assert st["not_converged"] == 0 and P["stats"]["not_converged"] == 0
    _log(f"C4: solves without a canonical cell: learned {st['noncanonical']}, plain {P['stats']['noncanonical']}")
    _log(f"C4: learned {st['secant_iters'] / st['seg_steps']:.3f} evaluations per segment-step, plain "
         f"{P['stats']['secant_iters'] / P['stats']['seg_steps']:.3f}; fallbacks {st['fallbacks']}")
    assert st["secant_iters"] < 0.7 * P["stats"]["secant_iters"]
    t0 = time.perf_counter()
    want = orc.wavefront_f64(w.topo, w.params, ql, w.qdp0, w.dp0, w.reach_last0, T, mode=orc.MODE_CONVERGED,
                             history=True, threads=THREADS)
    _log(f"C4: FP64 oracle, {THREADS} threads, {time.perf_counter() - t0:.0f} s")
    topo = orc.as_topo(w.topo)
    _judge("C4 LEARNED vs oracle", L["hist"][0], want["reach_hist"], w.reach_last0, topo)
    _judge("C4 CONVERGED vs oracle", P["hist"][0], want["reach_hist"], w.reach_last0, topo)
    rel = np.abs(L["q"][0] - want["qdc"]) / np.maximum(np.abs(want["qdc"]), 1e-3)
    _log(f"C4 final segment flows: max rel {rel.max():.3e}, {int((rel > 1e-9).sum())} of {rel.size} beyond 1e-9")
    assert np.array_equal(L["q"][0].view(np.uint64), want["qdc"].view(np.uint64))
    assert np.array_equal(L["d"][0].view(np.uint64), want["depth"].view(np.uint64))
    assert st["noncanonical"] == 0 and P["stats"]["noncanonical"] == 0


def test_c5_learned_1m_x16_288(net1m, weights10m):
    T, M = 288, 16
    w = net1m
    _log("C5: LEARNED 1M x 16 members x 288 on the GPU")
    L = _route(w, trmc.MODE_LEARNED, T, M, W=weights10m, units=1)
    st = L["stats"]
    assert st["seg_steps"] == w.topo.n_segments * M * T and st["not_converged"] == 0
    q, d, v, hist = L["q"], L["d"], L["v"], L["hist"]
    assert np.all(np.isfinite(q)) and np.all(np.isfinite(d)) and np.all(np.isfinite(v)) and np.all(np.isfinite(hist))
    assert q.min() >= 0.0 and hist.min() >= 0.0 and v.min() >= 0.0
    assert np.all(d[q > 0] >= 0.01 - 1e-15)                      # the root is never below mindepth
    assert all(not np.array_equal(hist[0], hist[m]) for m in range(1, M))
    roots = sample_roots(w.topo, 3000, 40000, seed=5)
    sub, seg_map, reach_map = subnetwork(w.topo, roots)
    _log(f"C5: oracle on {len(roots)} sampled closures: {sub['n_reaches']} reaches, {sub['n_segments']} segments")
    for m in range(M):
        ql = trmc.generate_forcing_member(w.topo.n_segments, T, m, SEED ^ 0xA1A1)[:, seg_map]
        want = orc.wavefront_f64(sub, w.params[seg_map], ql, w.qdp0[seg_map], w.dp0[seg_map],
                                 w.reach_last0[reach_map], T, mode=orc.MODE_CONVERGED, history=True,
                                 threads=THREADS)
        _judge(f"C5 member {m} LEARNED vs oracle", hist[m][:, reach_map], want["reach_hist"],
               w.reach_last0[reach_map], sub)


def test_c3_batch_64x8_288_reference_rule():
    T, M = 288, 8
    sizes = trmc.batch_sizes(64, 10_000, 200_000, 3)
    w, so, ro = trmc.make_batch_workload(sizes, 1, M, seed=3, forcing=False)
    _log(f"C3: 64 networks, {w.topo.n_segments} segments, {w.topo.n_reaches} reaches x {M} members x {T}")
    with trmc.Network(w.topo, mode=trmc.MODE_REFERENCE, n_members=M, record_history=1) as net:
        net.set_params(w.params)
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.generate_forcing(T, 3 ^ 0xA1A1)
        st = net.run(T)
        q, d, v, rl = net.state()
        hist = net.history(T)
    assert st["seg_steps"] == w.topo.n_segments * M * T
    assert np.all(np.isfinite(q)) and q.min() >= 0.0 and np.all(np.isfinite(hist)) and hist.min() >= 0.0
    _log(f"C3: {st['kernel_ms']:.1f} ms on the GPU, grid {st['grid_blocks']} blocks")
    smallest = np.argsort(sizes)[:3]
    for i in smallest:
        roots = list(range(int(ro[i]), int(ro[i + 1])))
        sub, seg_map, reach_map = subnetwork(w.topo, roots)
        assert sub["n_segments"] == int(so[i + 1] - so[i])           # networks are independent
        for m in (0, M - 1):
            ql = trmc.generate_forcing_member(w.topo.n_segments, T, m, 3 ^ 0xA1A1)[:, seg_map]
            want = orc.wavefront_f64(sub, w.params[seg_map], ql, w.qdp0[seg_map], w.dp0[seg_map],
                                     w.reach_last0[reach_map], T, history=True, variant="cr")
            got_h = hist[m][:, reach_map]
            same = np.array_equal(got_h.view(np.uint64), want["reach_hist"].view(np.uint64))
            same_q = np.array_equal(q[m][seg_map].view(np.uint64), want["qdc"].view(np.uint64))
            _log(f"C3 network {i} ({sub['n_segments']} segments) member {m}: bitwise history {same}, "
                 f"final flows {same_q}")
            assert same and same_q
