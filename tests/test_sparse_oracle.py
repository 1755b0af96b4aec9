"""The oracle above 20 readouts (sparse window cells and final states), on the CPU.

Readouts only label states for the statistics: which nodes are non-internal changes neither
the draws nor the transitions.  So the same network run with 24 readouts and with only the
first 20 of them must give window cells that, summed over the 4 extra readout bits, equal
the 20-readout run's dense table exactly (integer sums of the same sojourn pieces); likewise
the final states, and identical fixed points and transition counts.  The dense path is
pinned on its own (golden vectors, exact CTMC and jump-chain solutions), so this pins the
sparse bookkeeping: keys, windows, state bits, ordering and zero cells."""
import numpy as np
import pytest

from oracle.oracle import OracleModel
from upmabossg_amd import synth


def _pair(discrete, seed):
    net = synth.random_network(40, 2, 3, seed=seed, rates="loguniform", istate="uniform", readouts=24,
                               max_time=12.0, time_tick=1.0, sample_count=1500)
    net.discrete = discrete
    big = OracleModel(net.bnd(), net.cfg())
    net.readouts = list(range(20))
    small = OracleModel(net.bnd(), net.cfg())
    return big, small


@pytest.mark.parametrize("discrete,seed", [(False, 3), (True, 4)])
def test_sparse_cells_marginalise_to_the_dense_table(discrete, seed):
    big, small = _pair(discrete, seed)
    rb = big.run(threads=8)
    rs = small.run(threads=8)
    assert rb.readouts == 24 and rb.prob is None and rb.final is None
    assert rs.readouts == 20 and rs.prob is not None
    assert rb.flips == rs.flips
    assert np.array_equal(rb.fixed_states, rs.fixed_states) and np.array_equal(rb.fixed_counts, rs.fixed_counts)
    w, st, v = rb.cells
    assert len(v) > 0 and np.all(v > 0)
    key = (w.astype(np.uint64) << np.uint64(24)) | st
    assert np.all(key[1:] > key[:-1])  # window-major, states ascending
    assert int(st.max()) >= 1 << 20  # the extra readouts are visited
    dense = np.zeros(rs.prob.shape, np.uint64)
    np.add.at(dense, (w.astype(np.int64), (st & np.uint64((1 << 20) - 1)).astype(np.int64)), v)
    assert np.array_equal(dense, rs.prob)
    fs, fc = rb.final_cells
    assert int(fc.sum()) == 1500 and np.all(fs[1:] > fs[:-1])
    fin = np.zeros(1 << 20, np.uint64)
    np.add.at(fin, (fs & np.uint64((1 << 20) - 1)).astype(np.int64), fc)
    assert np.array_equal(fin, rs.final)
    # the dense run's cell lists are its non-zero table entries
    w2, s2, v2 = rs.cells
    assert np.array_equal(v2, rs.prob[w2.astype(np.int64), s2.astype(np.int64)])
    assert len(v2) == np.count_nonzero(rs.prob)


def test_thread_count_does_not_change_sparse_cells():
    big, _ = _pair(False, 5)
    a = big.run(threads=1).cells
    b = big.run(threads=7).cells
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
