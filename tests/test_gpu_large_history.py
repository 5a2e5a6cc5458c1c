"""A per-step flow history of more than 2^32 entries.  A dispatch counts its
work-items in 32 bits, so the library's element kernels (history set-up and
read-out, forcing scatter) loop over their range with the grid's stride; this
checks a configuration whose history is past that limit against one below
it.  The routing results do not depend on the unit size, so they must agree
bit for bit."""
import numpy as np
import pytest

from conftest import load_trmc

trmc = load_trmc()


@pytest.mark.gpu
def test_history_beyond_two_to_the_32_entries(gpu):
    """The 1M-segment network, 16 members, 288 steps, REFERENCE rule: units
    of one segment give (288 + 1) x 1,000,000 x 16 = 4.6e9 history entries;
    units of four give 1.9e9.  Every member's final flow and reach outflow
    are the same bits."""
    w = trmc.make_workload(2, 1_000_000, 1, n_reaches=300_000, seed=42)
    outs = []
    for units in (1, 4):
        net = trmc.Network(w.topo, mode=trmc.MODE_REFERENCE, n_members=16, unit_segments=units)
        try:
            net.set_params(w.params)
            net.set_state(w.qdp0, w.dp0, w.reach_last0)
            net.generate_forcing(288, 7)
            net.run(288)
            q, _, _, rl = net.state()
            outs.append((q.copy(), rl.copy()))
        finally:
            net.close()
    (qa, ra), (qb, rb) = outs
    assert np.all(np.isfinite(qa))
    assert np.array_equal(qa.view(np.uint64), qb.view(np.uint64))
    assert np.array_equal(ra.view(np.uint64), rb.view(np.uint64))
