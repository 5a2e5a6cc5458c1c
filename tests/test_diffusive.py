"""Diffusive-wave step (owp_extended_kernels.cu:23-135): the oracle pinned to
the reference's own cpu_diffusive_tridiag and to the linear system it solves
(CPU); the gfx950 kernel against the oracle, bit for bit (GPU)."""
import numpy as np
import pytest

from conftest import load_oracle, load_trmc

trmc = load_trmc()
orc = load_oracle()
NODES = 500


def _has_ref():
    ref = orc.load_ref()
    return ref is not None and hasattr(ref, "ref_cpu_diffusive_tridiag")


def _data(nreach, seed=42, ncomp=None, theta=0.6):
    """Inputs of the reference's shape and distributions (gen_diff_data,
    owp_extended_kernels.cu:347-362) from numpy's generator."""
    rng = np.random.default_rng(seed)
    rp = np.zeros(nreach, orc.DIFF_REACH)
    rp["ncomp"] = rng.integers(20, 100, nreach) if ncomp is None else ncomp
    rp["theta"] = theta
    rp["dtini"] = 60.0
    shape = nreach * NODES
    dx = (100 + 500 * rng.random(shape)).astype(np.float32)
    df = (10 + 100 * rng.random(shape)).astype(np.float32)
    ce = (0.5 + 2 * rng.random(shape)).astype(np.float32)
    qp = (1 + 50 * rng.random(shape)).astype(np.float32)
    return rp, dx, df, ce, qp


def _system_residual(rp, dx, df, qp, out):
    """Largest relative residual of the rows the elimination solves, in double:
    ppi q[i-1] + qqi q[i] + rri q[i+1] = ssi (i = 1 .. nc-1, q[nc] the ghost
    node qp_prev[nc-1]), and the upstream boundary q[0] = q[1]."""
    worst = 0.0
    for j in range(len(rp)):
        nc, th, dt = int(rp["ncomp"][j]), float(rp["theta"][j]), float(rp["dtini"][j])
        b = j * NODES
        q = np.append(out[b:b + nc].astype(np.float64), float(qp[b + nc - 1]))
        p = qp[b:b + nc].astype(np.float64)
        if nc > 1:
            worst = max(worst, abs(q[0] - q[1]) / abs(q[1]))
        for i in range(1, nc):
            dxi = max(float(dx[b + i - 1]), 1.0)
            d = max(float(df[b + i]), 0.01)
            ppi = -th * d * dt / (dxi * dxi)
            nxt = p[i + 1] if i < nc - 1 else p[i]
            ssi = p[i] + dt * d * (1 - th) * (p[i - 1] - 2 * p[i] + nxt) / (dxi * dxi)
            r = ppi * q[i - 1] + (1 - 2 * ppi) * q[i] + ppi * q[i + 1] - ssi
            worst = max(worst, abs(r) / abs(ssi))
    return worst


# ------------------------------------------------------------------ oracle pins (CPU)
@pytest.mark.parametrize("nreach", [1000, 10000])
def test_oracle_matches_reference_routine(nreach):
    if not _has_ref():
        pytest.skip("reference build (oracle/_ref) not present")
    rp, dx, df, ce, qp = orc.gen_diff_data_reference(nreach, 42)
    a = orc.diffusive_tridiag(rp, dx, df, ce, qp)
    b = orc.diffusive_tridiag(rp, dx, df, ce, qp, reference=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_oracle_solves_its_tridiagonal_system():
    for theta in (0.5, 0.6, 1.0):
        rp, dx, df, ce, qp = _data(200, seed=3, theta=theta)
        out = orc.diffusive_tridiag(rp, dx, df, ce, qp)
        assert _system_residual(rp, dx, df, qp, out) < 1e-5, theta


def test_constant_discharge_is_kept():
    rp, dx, df, ce, qp = _data(300, seed=4)
    qp[:] = 7.25
    out = orc.diffusive_tridiag(rp, dx, df, ce, qp)
    for j in range(len(rp)):
        nc = rp["ncomp"][j]
        np.testing.assert_allclose(out[j * NODES:j * NODES + nc], 7.25, rtol=1e-5)


def test_reach_without_nodes_is_rejected():
    rp, dx, df, ce, qp = _data(4, seed=5)
    for bad in (0, NODES + 1):
        rp["ncomp"][2] = bad
        with pytest.raises(trmc.TrmcError):
            trmc.diffusive_tridiag(rp, dx, df, ce, qp)


# ------------------------------------------------------------------ GPU parity
@pytest.mark.gpu
@pytest.mark.parametrize("nreach", [1000, 5000, 10000])
def test_gpu_matches_oracle_bitwise(gpu, nreach):
    if _has_ref():
        rp, dx, df, ce, qp = orc.gen_diff_data_reference(nreach, 42)
    else:
        rp, dx, df, ce, qp = _data(nreach)
    got = trmc.diffusive_tridiag(rp, dx, df, ce, qp)
    want = orc.diffusive_tridiag(rp, dx, df, ce, qp)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_gpu_edge_cases(gpu):
    """ncomp 1, 2 and 500, theta 0.5 and 1, spacing and diffusivity below their
    clamps; entries beyond a reach's ncomp keep the caller's values."""
    rp, dx, df, ce, qp = _data(6, seed=9)
    rp["ncomp"] = [1, 2, 500, 37, 500, 3]
    rp["theta"] = [0.6, 0.5, 1.0, 0.6, 0.6, 1.0]
    dx[3 * NODES:3 * NODES + 10] = 0.25
    df[4 * NODES:4 * NODES + 50] = 0.001
    sentinel = np.full(6 * NODES, -7.0, np.float32)
    got = trmc.diffusive_tridiag(rp, dx, df, ce, qp, qp_out=sentinel)
    want = sentinel.copy()
    o = orc.diffusive_tridiag(rp, dx, df, ce, qp)
    for j in range(6):
        nc = rp["ncomp"][j]
        want[j * NODES:j * NODES + nc] = o[j * NODES:j * NODES + nc]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_gpu_timed_launches_give_the_same_result(gpu):
    rp, dx, df, ce, qp = _data(2000, seed=8)
    a = trmc.diffusive_tridiag(rp, dx, df, ce, qp)
    b, ms = trmc.diffusive_tridiag(rp, dx, df, ce, qp, reps=5)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and ms > 0
