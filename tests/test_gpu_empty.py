"""No particles (N = 0): every solver mode initialises, solves, steps, logs
and frees, with zero energies; the reference octree of no particles is its
root alone (src/tree.c:74-318 inserts nothing after the purge)."""
import ctypes as C

import numpy as np
import pytest

from starflood_amd import _lib as L
from starflood_amd.sim import Simulation

pytestmark = pytest.mark.gpu

MODES = {
    "pp_exact": dict(enable_tree=0),
    "pp_fast": dict(enable_tree=0, precision=L.SF_PRECISION_FAST),
    "pp_fast_resident": dict(enable_tree=0, precision=L.SF_PRECISION_FAST, mem_mode=L.SF_MEM_RESIDENT),
    "tree_reference": dict(enable_tree=1),
    "tree_lbvh": dict(enable_tree=1, tree_mode=L.SF_TREE_LBVH, precision=L.SF_PRECISION_FAST),
    "sph_reference": dict(enable_tree=0, enable_sph=1),
    "sph_cubic_grid": dict(enable_tree=0, enable_sph=1, sph_model=L.SF_SPH_CUBIC_IDEAL,
                           precision=L.SF_PRECISION_FAST),
}


@pytest.mark.parametrize("mode", list(MODES))
def test_empty_simulation_solves_and_steps(tmp_path, mode):
    print(f"N=0 mode {mode}", flush=True)
    with Simulation(0, epsilon=1e-2, dt=1e-3, logging=1, out_dir=str(tmp_path), **MODES[mode]) as s:
        assert s.mem.size == 0
        s.solve()
        for n in range(2):
            s.step_number = n
            s.step()
        st = s.stats()
        assert st["T"] == 0.0 and st["V"] == 0.0 and st["K"] == 0.0
    rows = (tmp_path / "statistics.csv").read_text().splitlines()
    assert len(rows) == 3 and rows[1].split(",")[1:] == ["0.000000000000000"] * 3
    L.set_params()


def test_reference_tree_of_no_particles_is_the_root():
    lib = L.load()
    L.set_params(enable_tree=1)
    t = L.TreeT()
    L.check(lib.tree_init(C.byref(t), 64), "tree_init")
    for k in range(3):
        t.bounds[k] = -1.0
        t.bounds[k + 3] = 1.0
    L.check(lib.tree_build(C.byref(t), 0, None, None), "tree_build")
    assert t.num_nodes == 1
    node = np.ctypeslib.as_array(C.cast(t.node, C.POINTER(C.c_uint32)), shape=(20,))
    assert node[1] == 0   # bodies of the root
    L.check(lib.tree_free(C.byref(t)), "tree_free")
    L.set_params()
