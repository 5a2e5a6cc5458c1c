"""GPU parity of the FP32 REFERENCE rule (the reference's float solve
mc_solve, wavefront_mc.cu:66-193, and its float DAG loop cpu_wavefront,
:413-491).

The GPU computes every float operation of the reference in the same order,
with every powf correctly rounded; the oracle's float restatement with the
correctly rounded powf (oracle "cr") is therefore the bar, bit for bit, on
every segment and reach outflow.  Against the reference routine itself
(golden vectors written by it, and oracle "libm", which is bitwise the same)
the GPU differs exactly as the reference differs from itself when only its
powf is made correctly rounded: the host libm's powf is one ulp off in about
1 call in 1,600, and the rule's 1 % secant stop turns such an ulp into a
different iterate."""
import numpy as np
import pytest

from conftest import TopoView, golden, load_oracle, load_trmc, tiny_cases

pytestmark = pytest.mark.gpu
trmc = load_trmc()
orc = load_oracle()


def _log(msg):
    print(msg, flush=True)


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("T,qk,rk", [(8, "qdc_out_8", "reach_last_out_8"), (288, "qdc_out", "reach_last_out")])
def test_dropin_f32_c1_bitwise(gpu, T, qk, rk):
    c1, z = golden("wavefront_c1.npz"), golden("wavefront_f32_c1.npz")
    d = {k: c1[k] for k in c1.files}
    topo = TopoView(d).t
    _log(f"FP32 drop-in c1 T={T}")
    q, r = trmc.wavefront_f32(d["params"], topo, d["qlat"][:T], d["qdp0"], d["dp0"], d["reach_last0"], T, 300.0)
    want = orc.wavefront_f32(d, d["params"], d["qlat"][:T], d["qdp0"], d["dp0"], d["reach_last0"], T,
                             variant="cr")
    assert same_bits(q, want["qdc"]), "qdc differs from the oracle (correctly rounded powf)"
    assert same_bits(r, want["reach_last"]), "reach_last differs from the oracle (correctly rounded powf)"
    ref_q = z[qk].astype(np.float64)
    rel = np.abs(q.astype(np.float64) - ref_q) / np.maximum(np.abs(ref_q), 1e-3)
    _log(f"c1 T={T}: vs the reference routine (libm powf): bitwise {np.mean(q == z[qk]):.4f}, "
         f"max rel {rel.max():.3e}, median rel {np.median(rel):.3e}")


@pytest.mark.parametrize("name", sorted(tiny_cases()))
def test_dropin_f32_tiny_networks(gpu, name):
    d = tiny_cases()[name]
    T = int(d["n_timesteps"])
    topo = TopoView(d).t
    q, r = trmc.wavefront_f32(d["params"], topo, d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], T, 300.0)
    want = orc.wavefront_f32(d, d["params"], d["qlat"], d["qdp0"], d["dp0"], d["reach_last0"], T, variant="cr")
    assert same_bits(q, want["qdc"]) and same_bits(r, want["reach_last"])


def test_network_f32_shreve_history_bitwise(gpu):
    """The device-resident network in REFERENCE_F32 mode on a 60,000-segment
    Shreve tree, 24 steps: every reach outflow of every step, and the final
    per-segment flows and depths, bit for bit."""
    T = 24
    w = trmc.make_workload(2, 60_000, T, n_reaches=18_000, seed=5)
    p32 = np.asarray(w.params, np.float32)
    ql = np.asarray(w.qlat[0], np.float32)
    q0, d0, r0 = (np.asarray(x, np.float32) for x in (w.qdp0, w.dp0, w.reach_last0))
    _log("FP32 network shreve-60k T=24")
    with trmc.Network(w.topo, mode=trmc.MODE_REFERENCE_F32, record_history=1) as net:
        net.set_params(p32)
        net.set_state(q0, d0, r0)
        net.set_forcing(ql)
        st = net.run(T)
        q, dep, _, rl = net.state()
        hist = net.history(T)[0]
    want = orc.wavefront_f32(orc.as_topo(w.topo), p32, ql, q0, d0, r0, T, variant="cr", threads=8, history=True)
    assert np.array_equal(q[0], q[0].astype(np.float32)), "flows are not float values"
    assert same_bits(q[0], want["qdc"]) and same_bits(dep[0], want["depth"])
    assert same_bits(rl[0], want["reach_last"]) and same_bits(hist, want["reach_hist"])
    assert st["secant_iters"] == want["iters"]


def test_f32_mode_refuses_other_schedules(gpu):
    w = trmc.make_workload(1, 200, 2, seed=1)
    with trmc.Network(w.topo, mode=trmc.MODE_REFERENCE_F32, schedule=trmc.SCHED_SKEWED) as net:
        net.set_params(np.asarray(w.params, np.float32))
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(np.asarray(w.qlat[0], np.float32))
        with pytest.raises(trmc.TrmcError):
            net.run(2)
