"""The reference octree on the device (SF_TREE_REFERENCE): node array, walk,
and a tree step bit-identical to the reference's."""
import numpy as np
import pytest

from conftest import bits, golden
from starflood_amd import _lib as L
from starflood_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def test_tree_build_walk_and_step_bit_identical_to_reference():
    g = golden("f4_tree_n4096.npz")
    with Simulation(4096, enable_tree=1, logging=0, epsilon=1e-2, dt=1e-3) as s:
        s.mem[:] = g["ic"]
        s.step_number = 0
        s.solve()
        assert s.sim.tree.num_nodes == int(g["num_nodes0"])
        assert np.array_equal(s.tree_nodes(), g["nodes0"])
        assert np.array_equal(bits(s.mem), bits(g["mem0"]))
        s.step_number = 1
        s.step()
        assert s.sim.tree.num_nodes == int(g["num_nodes1"])
        assert np.array_equal(s.tree_nodes(), g["nodes1"])
        assert np.array_equal(bits(s.mem), bits(g["mem1"]))


def test_tree_matches_oracle_on_plummer_and_reports_overflow(oracle):
    N = 6000
    with Simulation(N, enable_tree=1, logging=0, epsilon=1e-2, dt=1e-3) as s:
        s.ic_plummer(seed=2, first=0, count=N)
        pos, mas = s.pos.copy(), s.mas.copy()
        s.step_number = 7
        s.solve()
        rc, bounds, nodes, nn, raw = oracle.tree_ref(pos, mas, 7)
        assert rc == 0 and s.sim.tree.num_nodes == nn
        assert np.array_equal(s.tree_nodes(), raw)
        acc, pot = oracle.tree_walk_ref(pos, mas, nodes, 262144, 1e-2)
        assert np.array_equal(bits(s.acc.ravel()), bits(acc.ravel()))
        assert np.array_equal(bits(s.pot), bits(pot))
    # too few nodes: the build fails (and is reported), the walk of the
    # partial tree still runs, as in the reference
    with Simulation(N, enable_tree=1, logging=0, epsilon=1e-2, dt=1e-3, tree_nodes_max=64) as s:
        s.ic_plummer(seed=2, first=0, count=N)
        with pytest.raises(RuntimeError):
            s.solve()
        rc, bounds, nodes, nn, raw = oracle.tree_ref(s.pos.copy(), s.mas.copy(), 0, max_nodes=64)
        assert rc == -1 and s.sim.tree.num_nodes == nn
        assert np.array_equal(s.tree_nodes(), raw)
    L.set_params()
