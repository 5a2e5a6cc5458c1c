"""Linearized Muskingum-Cunge (wavefront/linear_mc.cu): the oracle pinned to
the reference's own cpu_linear_mc_fp64 and to what the coefficient arithmetic
fixes (CPU), then the GPU LINEAR mode and the drop-in trmc_linear_mc_f64
against the oracle, bit for bit (GPU)."""
import numpy as np
import pytest

from conftest import TopoView, golden, load_oracle, load_trmc

trmc = load_trmc()
orc = load_oracle()
DT = 300.0


def _c1():
    z = golden("wavefront_c1.npz")
    d = {k: z[k] for k in z.files}
    return d, TopoView(d).t


def _shreve(n=20000, r=6000, T=24, seed=11):
    w = trmc.make_workload(2, n, T, n_reaches=r, seed=seed)
    return w, np.asarray(w.params, np.float32), np.asarray(w.qlat[0], np.float32), np.asarray(w.qdp0, np.float32)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


# ------------------------------------------------------------------ oracle pins (CPU)
def test_linear_oracle_matches_reference_routine():
    if orc.load_ref() is None or not hasattr(orc.load_ref(), "ref_cpu_linear_mc_fp64"):
        pytest.skip("reference build (oracle/_ref) not present")
    d, topo = _c1()
    c = orc.linear_coeffs(d["params"], DT)
    for T in (1, 8, 288):
        a = orc.linear_mc_fp64(c, topo, d["qlat"][:T], d["qdp0"], T)
        b = orc.linear_mc_fp64(c, topo, d["qlat"][:T], d["qdp0"], T, reference=True)
        assert _bits_equal(a, b), T
    w, p32, q32, Q0 = _shreve()
    c = orc.linear_coeffs(p32, DT)
    a = orc.linear_mc_fp64(c, w.topo, q32, Q0, 24)
    b = orc.linear_mc_fp64(c, w.topo, q32, Q0, 24, reference=True)
    assert _bits_equal(a, b)


def test_linear_coefficients_fixed_by_the_arithmetic():
    """With Qj = 0 the weighting is X = 0.5 for every valid channel, so
    D = Km/2 + dt/2 is computed by the same float operations as C1's
    numerator: C1 == 1 and C3 == -C2 exactly, C1 + C2 + C3 == 1, and
    C4/ql == dt / D; Km follows the celerity's closed form in double."""
    _, p32, _, _ = _shreve(5000, 1500, 1)
    C1, C2, C3, C4 = orc.linear_coeffs(p32, DT)
    assert np.all(C1 == 1.0) and np.all(C3 == -C2)
    assert np.all((C1 + C2) + C3 == 1.0)
    p = p32.astype(np.float64)
    dx, bw, tw, twcc, n, ncc, cs, s0 = p.T
    z = np.where(cs == 0, 1.0, 1.0 / cs)
    bfd = np.where(bw > tw, bw / 1e-5, np.where(bw == tw, bw / (2 * z), (tw - bw) / (2 * z)))
    h = np.maximum(0.5 * bfd, 0.1)
    A = (bw + h * z) * h
    WP = bw + 2 * h * np.sqrt(1 + z * z)
    R = A / WP
    ck = np.maximum(1e-6, np.sqrt(s0) / n * (5 / 3 * R ** (2 / 3) - 2 / 3 * R ** (5 / 3) * (2 * np.sqrt(1 + z * z) / (bw + 2 * h * z))))
    km = np.maximum(DT, dx / ck)
    D = 0.5 * km + 0.5 * DT
    np.testing.assert_allclose(C4, DT / D, rtol=2e-6)
    np.testing.assert_allclose(C2, (0.5 * DT - 0.5 * km) / D, rtol=2e-6, atol=1e-7)


def test_linear_fixed_point_is_accumulated_lateral_inflow():
    """Constant lateral inflow: C1 + C2 = C4/ql = 1 - C3, so the fixed point of
    every segment is its upstream flow plus its own lateral inflow.  A dropped
    term (C2 or C4) or a wrong upstream index moves it far beyond 1e-6."""
    w = trmc.make_workload(1, 200, 1, seed=3)
    t = w.topo
    p32 = np.asarray(w.params, np.float32)
    # short, fast segments: |C3| well below 1, fast convergence
    p32[:, 0] = 200.0
    c = orc.linear_coeffs(p32, DT)
    ql = np.float32(0.25) * (1 + (np.arange(t.n_segments) % 3)).astype(np.float32)
    T = 400
    Q = orc.linear_mc_fp64(c, t, np.tile(ql, (T, 1)), np.ones(t.n_segments, np.float32), T)
    want = np.zeros(t.n_segments)
    last = {}
    for L in range(t.n_levels):
        for r in t.level_reach[t.level_ptr[L]:t.level_ptr[L + 1]]:
            up = sum(last[u] for u in t.reach_up_idx[t.reach_up_ptr[r]:t.reach_up_ptr[r + 1]])
            for i in range(t.reach_seg_len[r]):
                s = t.reach_seg_start[r] + i
                up = up + float(ql[s])
                want[s] = up
            last[r] = up
    # the float coefficients satisfy C4/ql == C1 + C2 to float rounding, so the
    # fixed point matches to a few float ulps (1e-6), not to double precision
    np.testing.assert_allclose(Q, want, rtol=1e-6)


# ------------------------------------------------------------------ GPU parity
@pytest.mark.gpu
def test_linear_coefficients_on_gpu_match_oracle(gpu):
    _, p32, _, _ = _shreve(100000, 30000, 1)
    got = trmc.linear_coeffs(p32, DT)
    want = orc.linear_coeffs(p32, DT)
    for g, wv in zip(got, want):
        same = np.mean(g.view(np.uint32) == wv.view(np.uint32))
        ulps = np.abs(g.view(np.int32).astype(np.int64) - wv.view(np.int32).astype(np.int64))
        assert same >= 0.9999 and ulps.max() <= 1, (same, ulps.max())


@pytest.mark.gpu
def test_linear_dropin_matches_reference_c1(gpu):
    d, topo = _c1()
    c = orc.linear_coeffs(d["params"], DT)
    for T in (1, 8, 288):
        got = trmc.linear_mc(c, topo, d["qlat"][:T], d["qdp0"], T)
        want = orc.linear_mc_fp64(c, topo, d["qlat"][:T], d["qdp0"], T)
        assert _bits_equal(got, want), T
        ref = orc.load_ref()
        if ref is not None and hasattr(ref, "ref_cpu_linear_mc_fp64"):
            assert _bits_equal(got, orc.linear_mc_fp64(c, topo, d["qlat"][:T], d["qdp0"], T, reference=True))


@pytest.mark.gpu
def test_linear_full_size_network_bitwise(gpu):
    """1M segments x 288 steps: every segment's final flow equals the oracle's
    bit for bit (the recurrence has no transcendental, so nothing may differ)."""
    T = 288
    w = trmc.make_workload(2, 1_000_000, T, n_reaches=300_000, seed=42)
    p32 = np.asarray(w.params, np.float32)
    q32 = np.asarray(w.qlat[0], np.float32)
    Q0 = np.asarray(w.qdp0, np.float32)
    c = orc.linear_coeffs(p32, DT)
    got = trmc.linear_mc(c, w.topo, q32, Q0, T)
    want = orc.linear_mc_fp64(c, w.topo, q32, Q0, T)
    assert _bits_equal(got, want)


@pytest.mark.gpu
def test_linear_members_and_chained_runs(gpu):
    w, p32, _, Q0 = _shreve(20000, 6000, 24)
    M, T = 4, 24
    c = orc.linear_coeffs(p32, DT)
    rng = np.random.default_rng(5)
    qlat = rng.lognormal(np.log(0.1) - 0.5, 1.0, (M, T, w.topo.n_segments)).astype(np.float32)
    net = trmc.Network(w.topo, mode=trmc.MODE_LINEAR, n_members=M)
    try:
        net.set_linear_coeffs(c)
        last = np.array([Q0[s + n - 1] for s, n in zip(w.topo.reach_seg_start, w.topo.reach_seg_len)], np.float64)
        net.set_state(np.tile(Q0.astype(np.float64), (M, 1)), np.zeros((M, w.topo.n_segments)),
                      np.tile(last, (M, 1)))
        net.set_forcing(qlat)
        net.run(10)
        net.run(T - 10)
        q = net.state()[0]
        for m in range(M):
            assert _bits_equal(q[m], orc.linear_mc_fp64(c, w.topo, qlat[m], Q0, T)), m
    finally:
        net.close()


@pytest.mark.gpu
def test_linear_mode_needs_coefficients(gpu):
    w, _, _, Q0 = _shreve(2000, 600, 2)
    net = trmc.Network(w.topo, mode=trmc.MODE_LINEAR)
    try:
        net.set_state(w.qdp0, w.dp0, w.reach_last0)
        net.set_forcing(w.qlat)
        with pytest.raises(trmc.TrmcError):
            net.run(1)
    finally:
        net.close()
