"""Matrix form of the linearized MC (wavefront/linear_mc_matrix.cu,
linear_mc_matrix_graph.cu): the oracle's and the library's propagation
matrices against the reference's own build_propagation_matrices, bit for bit,
and the matrix-form steps against the FP64 level-scheduled recurrence they
restate (CPU); then the gfx950 step kernel against the oracle, bit for bit,
for each lane count (GPU)."""
import numpy as np
import pytest

from conftest import golden, load_oracle, load_trmc

trmc = load_trmc()
orc = load_oracle()
DT = 300.0


def _has_ref():
    ref = orc.load_ref()
    return ref is not None and hasattr(ref, "ref_matrix_build")


def _net(n=20000, seed=1, T=24):
    """One segment per reach (the matrix form's network), reaches numbered
    upstream first; lognormal lateral inflow as the forcing generator's."""
    w = trmc.make_workload(2, n, 1, n_reaches=n, seed=seed)
    t = w.topo
    assert t.n_segments == t.n_reaches
    p = np.asarray(w.params, np.float32)
    rng = np.random.default_rng(seed)
    qlat = rng.lognormal(np.log(0.1) - 0.5, 1.0, (T, n)).astype(np.float32)
    Q0 = np.asarray(w.qdp0, np.float32)
    return t, p, qlat, Q0


def _hydrofabric_net():
    z = golden("hydrofabric.npz")
    ids = [None if v == "" else str(v) for v in z["fp_id"]]
    toids = [None if v == "" else str(v) for v in z["fp_toid"]]
    vals = np.where(z["at_null"], np.nan, z["at_vals"])
    topo, params, _ = trmc.build_flowpath_network(ids, toids, [str(v) for v in z["at_link"]], vals)
    return topo, params


def _same(a, b):
    return all(np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
               for m, r in zip(a, b) for x, y in zip(m, r))


# ------------------------------------------------------------------ CPU: builder and oracle pins
@pytest.mark.parametrize("tol", [0.0, 1e-7, 1e-6, 1e-3])
def test_builders_match_reference_routine(tol):
    if not _has_ref():
        pytest.skip("reference build (oracle/_ref) not present")
    t, p, _, _ = _net(20000, seed=3)
    c = orc.matrix_coeffs(p)
    ref = orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, tol, reference=True)
    assert _same(orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, tol), ref)
    assert _same(trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, tol), ref)


def test_threaded_builder_matches_oracle():
    """100k reaches: the library builds the levels on several host threads;
    every row is still the same float sequence, so the matrices equal the
    oracle's (one thread, sorted merges) bit for bit."""
    t, p, _, _ = _net(100000, seed=12)
    c = orc.matrix_coeffs(p)
    assert _same(trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6),
                 orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6))


def test_builders_match_on_hydrofabric_network():
    """The network the hydrofabric builder writes (fan-in up to 12, forest of
    outlets), and the library's builder against the oracle's where the
    reference build is absent."""
    topo, params = _hydrofabric_net()
    c = orc.matrix_coeffs(params)
    lib = trmc.matrix_build(topo.reach_up_ptr, topo.reach_up_idx, c, 1e-6)
    assert _same(lib, orc.matrix_build(topo.reach_up_ptr, topo.reach_up_idx, c, 1e-6))
    if _has_ref():
        assert _same(lib, orc.matrix_build(topo.reach_up_ptr, topo.reach_up_idx, c, 1e-6, reference=True))


def test_coefficients_match_reference_and_arithmetic():
    """Against the reference's compute_mc_coeffs_cpu: equal but where the host
    powf and the rounded double pow differ by an ulp (0.03 % of channels here),
    which the later float steps carry on (C2 = (dt/2 - Km X) / D cancels, so
    up to ~1.5e-6 relative); and X = 0.25 fixes C1 + C2 + C3 = 1 and
    C4/ql = dt / D with D = 0.75 Km + dt/2."""
    _, p, _, _ = _net(20000, seed=4)
    C1, C2, C3, C4 = orc.matrix_coeffs(p)
    if _has_ref():
        for a, b in zip((C1, C2, C3, C4), orc.matrix_coeffs(p, reference=True)):
            assert np.mean(a.view(np.uint32) == b.view(np.uint32)) > 0.999
            np.testing.assert_allclose(a, b, rtol=4e-6)
    s = C1.astype(np.float64) + C2 + C3
    np.testing.assert_allclose(s, 1.0, atol=4e-7)
    Km = (DT / C4.astype(np.float64) - DT / 2) / 0.75
    np.testing.assert_allclose(C1, (0.25 * Km + DT / 2) / (0.75 * Km + DT / 2), rtol=2e-6)
    assert np.all(Km >= DT * (1 - 1e-5))


def test_matrix_form_is_the_linear_recurrence():
    """With tol = 0 nothing is dropped and P_Q Q + P_q ql is exactly the
    one-step linear MC (C1 new upstream + C2 old upstream + C3 old + C4 ql) in
    real arithmetic: the float matrix steps agree with the FP64 level
    recurrence to float rounding.  A dropped term, a wrong column or a
    transposed matrix moves them by orders of magnitude more."""
    t, p, qlat, Q0 = _net(5000, seed=5, T=24)
    c = orc.matrix_coeffs(p)
    PQ, Pq = orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 0.0)
    want = orc.linear_mc_fp64(c, t, qlat, Q0, 24)
    got = orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=1)
    np.testing.assert_allclose(got, want, rtol=2e-5)
    # pruning at 1e-6 (the reference's default) stays within 1 % everywhere
    PQ, Pq = orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    got = orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=1)
    np.testing.assert_allclose(got, want, rtol=1e-2)


def test_summation_orders_agree():
    t, p, qlat, Q0 = _net(5000, seed=6, T=12)
    c = orc.matrix_coeffs(p)
    PQ, Pq = orc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    base = orc.matrix_route(PQ, Pq, qlat, Q0, 12, lanes=1)
    for G in (0, 2, 8, 64):
        np.testing.assert_allclose(orc.matrix_route(PQ, Pq, qlat, Q0, 12, lanes=G), base, rtol=1e-5)


def test_hand_worked_two_reach_chain():
    """Reach 0 drains into reach 1: P_Q = [[C3_0, 0], [C1_1 C3_0 + C2_1, C3_1]],
    P_q = [[C4_0, 0], [C1_1 C4_0, C4_1]] (float products)."""
    f = np.float32
    c = (np.array([0.5, 0.4], f), np.array([0.2, 0.3], f), np.array([0.3, 0.3], f), np.array([0.9, 0.8], f))
    (rQ, cQ, vQ), (rq, cq, vq) = trmc.matrix_build([0, 0, 1], [0], c, 0.0)
    assert list(rQ) == [0, 1, 3] and list(cQ) == [0, 0, 1] and list(rq) == [0, 1, 3] and list(cq) == [0, 0, 1]
    assert vQ[0] == f(0.3) and vQ[1] == f(f(0.3) * f(0.4)) + f(0.3) and vQ[2] == f(0.3)
    assert vq[0] == f(0.9) and vq[1] == f(f(0.9) * f(0.4)) and vq[2] == f(0.8)


def test_builder_rejects_unordered_network():
    c = tuple(np.ones(3, np.float32) * v for v in (0.4, 0.3, 0.3, 0.9))
    with pytest.raises(trmc.TrmcError) as e:
        trmc.matrix_build([0, 1, 1, 1], [2], c, 1e-6)        # reach 0's upstream is reach 2
    assert e.value.code == trmc.ERR_TOPOLOGY


def test_route_rejects_bad_matrices_before_any_device_work():
    PQ = (np.array([0, 1, 2], np.int32), np.array([0, 5], np.int32), np.ones(2, np.float32))
    Pq = (np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.ones(2, np.float32))
    q, Q0 = np.zeros((1, 2), np.float32), np.zeros(2, np.float32)
    with pytest.raises(trmc.TrmcError):
        trmc.matrix_route(PQ, Pq, q, Q0, 1)                  # column 5 of a 2-row matrix
    with pytest.raises(trmc.TrmcError):
        trmc.matrix_route(Pq, Pq, q, Q0, 1, lanes_per_row=3)
    two = (np.array([0, 2, 2], np.int32), np.array([1, 0], np.int32), np.ones(2, np.float32))
    with pytest.raises(trmc.TrmcError):
        trmc.matrix_route(two, Pq, q, Q0, 1)                 # columns of row 0 not ascending


def _q0(t, seed):
    rng = np.random.default_rng(seed)
    return rng.lognormal(np.log(0.1) - 0.5, 1.0, t.n_segments).astype(np.float32)


def _bits(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


def test_secant_coefficients_match_reference_solver():
    """linear_mc_matrix_secant.cu: the library's coefficients equal the
    reference's own mc_secant_solve_cpu (oracle/_ref) run on the per-reach
    inputs of its main, restated in the oracle, bit for bit: a 20k Shreve
    network and the hydrofabric network."""
    if orc.load_ref() is None or not hasattr(orc.load_ref(), "ref_matrix_secant_solve"):
        pytest.skip("reference build (oracle/_ref) not present")
    t, p, _, _ = _net(20000, seed=3)
    q0 = _q0(t, 3)
    lib = trmc.matrix_secant_coeffs(p, t, q0)
    assert _bits(lib, orc.matrix_secant_solve_reference(orc.matrix_secant_inputs(t, p, q0), p))
    topo, params = _hydrofabric_net()
    q0 = _q0(topo, 4)
    lib = trmc.matrix_secant_coeffs(params, topo, q0)
    assert _bits(lib, orc.matrix_secant_solve_reference(orc.matrix_secant_inputs(topo, params, q0), params))


def test_secant_reference_flows_are_subtree_sums():
    """The steady reference flow of a reach is the sum, over the reach and
    every reach upstream of it, of its first-step inflow plus the mean inflow
    (float sums in the reference, double here)."""
    t, p, _, _ = _net(5000, seed=5)
    q0 = _q0(t, 5)
    refq = orc.matrix_secant_inputs(t, p, q0)[5]
    mean = q0.astype(np.float64).mean()
    want = q0.astype(np.float64)[:t.n_reaches] + mean
    for r in range(t.n_reaches):                       # reaches are numbered upstream first
        for u in t.reach_up_idx[t.reach_up_ptr[r]:t.reach_up_ptr[r + 1]]:
            want[r] += want[u]
    np.testing.assert_allclose(refq, want, rtol=2e-5)


def test_secant_falls_back_to_half_bankfull_coefficients():
    """Invalid channels (zero bottom width) take the X = 0.25 half-bankfull
    coefficients: those of the reference's compute_mc_coeffs_cpu."""
    t, p, _, _ = _net(3000, seed=6)
    p = p.copy()
    bad = np.arange(0, t.n_reaches, 7)
    p[bad, 1] = 0.0
    lib = trmc.matrix_secant_coeffs(p, t, _q0(t, 6))
    want = orc.matrix_coeffs(p, reference=orc.load_ref() is not None and hasattr(orc.load_ref(), "ref_matrix_coeffs"))
    for g, w in zip(lib, want):
        assert np.array_equal(g[bad].view(np.uint32), w[bad].view(np.uint32)) or \
            np.allclose(g[bad], w[bad], rtol=4e-6)


def test_secant_needs_an_inflow_per_reach():
    t, p, _, _ = _net(200, seed=8)
    with pytest.raises(trmc.TrmcError) as e:
        trmc.matrix_secant_coeffs(p, t, np.zeros(t.n_reaches - 1, np.float32))
    assert e.value.code == trmc.ERR_ARG


def test_secant_rejects_bad_levels():
    import dataclasses
    t, p, _, _ = _net(200, seed=7)
    lr = np.array(t.level_reach, np.int32).copy()
    lr[0] = lr[1]                                       # a reach listed twice
    with pytest.raises(trmc.TrmcError):
        trmc.matrix_secant_coeffs(p, dataclasses.replace(t, level_reach=lr), _q0(t, 7))


# ------------------------------------------------------------------ GPU parity
@pytest.mark.gpu
def test_gpu_coefficients_match_oracle(gpu):
    _, p, _, _ = _net(100000, seed=7)
    for g, w in zip(trmc.matrix_coeffs(p), orc.matrix_coeffs(p)):
        assert np.mean(g.view(np.uint32) == w.view(np.uint32)) >= 0.9999
        np.testing.assert_allclose(g, w, rtol=4e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [0, 1, 4, 8, 16, 64])
def test_gpu_steps_match_oracle_bitwise(gpu, lanes):
    t, p, qlat, Q0 = _net(20000, seed=8, T=24)
    c = orc.matrix_coeffs(p)
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    got = trmc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes_per_row=lanes)
    want = orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=lanes)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 0])
def test_gpu_capped_grid_loops_over_block_table(gpu, monkeypatch, lanes):
    """A grid of 8 blocks against a block table of hundreds: every block
    of the table is served on the kernel's later loop trips, bit for bit."""
    monkeypatch.setenv("TRMC_MATRIX_MAX_GRID", "8")
    t, p, qlat, Q0 = _net(20000, seed=8, T=24)
    c = orc.matrix_coeffs(p)
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    got = trmc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes_per_row=lanes)
    want = orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=lanes)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_gpu_full_size_network_bitwise(gpu):
    """309,000 reaches x 24 steps at tol 1e-6 (the reference's real-network
    size and its graph variant's default), lanes by row length (the default)."""
    t, p, qlat, Q0 = _net(309000, seed=42, T=24)
    c = orc.matrix_coeffs(p)
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    got = trmc.matrix_route(PQ, Pq, qlat, Q0, 24)
    want = orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # and within 1 % of the FP64 recurrence, as the reference reports
    np.testing.assert_allclose(got, orc.linear_mc_fp64(c, t, qlat, Q0, 24), rtol=1e-2)


@pytest.mark.gpu
def test_gpu_edge_cases(gpu):
    t, p, qlat, Q0 = _net(3000, seed=9, T=3)
    c = orc.matrix_coeffs(p)
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-6)
    assert np.array_equal(trmc.matrix_route(PQ, Pq, qlat, Q0, 0), Q0)               # no step
    for T in (1, 3):                                                               # odd and even step counts
        got = trmc.matrix_route(PQ, Pq, qlat[:T], Q0, T)
        assert np.array_equal(got.view(np.uint32), orc.matrix_route(PQ, Pq, qlat, Q0, T, lanes=0).view(np.uint32))
    b, ms = trmc.matrix_route(PQ, Pq, qlat, Q0, 3, reps=4)                        # timed replays
    assert np.array_equal(b, trmc.matrix_route(PQ, Pq, qlat, Q0, 3)) and ms > 0
    topo, params = _hydrofabric_net()                                              # hydrofabric network
    c = trmc.matrix_coeffs(params)
    PQ, Pq = trmc.matrix_build(topo.reach_up_ptr, topo.reach_up_idx, c, 1e-6)
    rng = np.random.default_rng(1)
    q = rng.lognormal(-2.8, 1.0, (24, topo.n_reaches)).astype(np.float32)
    Q0 = np.full(topo.n_reaches, 0.5, np.float32)
    got = trmc.matrix_route(PQ, Pq, q, Q0, 24, lanes_per_row=16)
    assert np.array_equal(got.view(np.uint32), orc.matrix_route(PQ, Pq, q, Q0, 24, lanes=16).view(np.uint32))


@pytest.mark.gpu
def test_gpu_secant_fitted_matrices_route_bitwise(gpu):
    """The secant-fitted coefficients through the same builder and step
    kernel (tol 1e-7, the secant variant's default): every reach equals the
    oracle's steps bit for bit."""
    t, p, qlat, Q0 = _net(20000, seed=10, T=24)
    c = trmc.matrix_secant_coeffs(p, t, qlat[0])
    PQ, Pq = trmc.matrix_build(t.reach_up_ptr, t.reach_up_idx, c, 1e-7)
    got = trmc.matrix_route(PQ, Pq, qlat, Q0, 24)
    assert np.array_equal(got.view(np.uint32), orc.matrix_route(PQ, Pq, qlat, Q0, 24, lanes=0).view(np.uint32))
