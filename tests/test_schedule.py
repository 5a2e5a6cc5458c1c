"""The time-skewed schedule, executed on the CPU from the library's own plan
arrays with the oracle's segment solve, reproduces the reference's level/time
order bit for bit.  This checks the GPU kernel's data flow (unit cutting,
ALAP super-levels, ring slots, in-place segment state, internal vs summed
boundaries) without a GPU; the kernel itself is checked in test_gpu_parity."""
import numpy as np
import pytest

from conftest import TopoView, golden, load_oracle, load_trmc, tiny_cases

trmc = load_trmc()
orc = load_oracle()


def emulate(topo, params, qlat, qdp0, dp0, rl0, T, C, reverse=False, dt=300.0):
    p = trmc.plan(topo, C, T)
    U, L, D = p["n_units"], p["n_levels"], p["ring"]
    sod = p["seg_of_dev"]
    q = np.asarray(qdp0, np.float64)[sod].copy()
    d = np.asarray(dp0, np.float64)[sod].copy()
    ring = np.zeros((D, U))
    for u in range(U):
        src = p["unit_init"][u]
        ring[0, u] = qdp0[src] if src >= 0 else rl0[-(src + 1)]
    P = np.asarray(params, np.float64)[sod]
    ql = np.asarray(qlat, np.float64)[:, sod]
    for k in range(L + T - 1):
        lo = p["level_ptr"][max(0, k - T + 1)]
        hi = p["level_ptr"][min(k, L - 1) + 1]
        items = range(hi - 1, lo - 1, -1) if reverse else range(lo, hi)
        for u in items:
            t = k - p["unit_level"][u]
            prev, cur = ring[t % D], ring[(t + 1) % D]
            ups = p["up_idx"][p["up_ptr"][u]:p["up_ptr"][u + 1]]
            if p["unit_flags"][u] & 1:
                qup_in, quc_in = prev[ups[0]], cur[ups[0]]
            else:
                qup_in = quc_in = 0.0
                for v in ups:
                    qup_in += prev[v]
                    quc_in += cur[v]
            last = qup_next = quc_next = 0.0
            for s in range(p["unit_seg"][u], p["unit_seg"][u + 1]):
                first = s == p["unit_seg"][u]
                qup = qup_in if first else qup_next
                quc = quc_in if first else quc_next
                qdp = q[s]
                qd, dep, _, _ = orc.mc_solve(P[s], qup, quc, qdp, ql[t, s], d[s], dt)
                q[s], d[s] = qd, dep
                qup_next, quc_next, last = qdp, qd, qd
            cur[u] = last
    out = np.array(qdp0, np.float64).copy()
    out[sod] = q
    rl = ring[T % D][p["reach_last_unit"]]
    return out, rl


def bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


@pytest.mark.parametrize("name", sorted(tiny_cases()))
@pytest.mark.parametrize("C", [1, 2, 100])
def test_tiny_networks(name, C):
    dd = tiny_cases()[name]
    T = int(dd["n_timesteps"])
    topo = TopoView(dd).t
    a, b = emulate(topo, dd["params"], dd["qlat"], dd["qdp0"].astype(np.float64),
                   dd["dp0"].astype(np.float64), dd["reach_last0"].astype(np.float64), T, C, reverse=(C == 2))
    covered = np.zeros(topo.n_segments, bool)
    for s, n in zip(topo.reach_seg_start, topo.reach_seg_len):
        covered[s:s + n] = True
    assert np.array_equal(bits(a[covered]), bits(dd["qdc_out"][covered]))
    assert np.array_equal(bits(b), bits(dd["reach_last_out"]))


@pytest.mark.parametrize("C", [1, 3, 40])
def test_golden_c1_first_24_steps(C):
    z = golden("wavefront_c1.npz")
    dd = {k: z[k] for k in z.files}
    T = 24
    qo, ro, _ = orc.wavefront_f32in(dd, dd["params"], dd["qlat"][:T], dd["qdp0"], dd["dp0"], dd["reach_last0"], T)
    a, b = emulate(TopoView(dd).t, dd["params"], dd["qlat"][:T], dd["qdp0"].astype(np.float64),
                   dd["dp0"].astype(np.float64), dd["reach_last0"].astype(np.float64), T, C)
    assert np.array_equal(bits(a), bits(qo))
    assert np.array_equal(bits(b), bits(ro))


def test_shreve_tree_segment_units():
    w = trmc.make_workload(2, 6000, 6, n_reaches=1800, seed=8)
    r = orc.wavefront_f64(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, 6)
    a, b = emulate(w.topo, w.params, w.qlat[0], w.qdp0, w.dp0, w.reach_last0, 6, 1, reverse=True)
    assert np.array_equal(bits(a), bits(r["qdc"]))
    assert np.array_equal(bits(b), bits(r["reach_last"]))
