"""Host-side behaviour without a GPU: the benchmark's one-GPU contract and the
argument checks of the device-array entry points (which return before any
device call)."""
import ctypes as C
import json
import subprocess
import sys

from conftest import REPO
from starflood_amd import _lib as L


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 4 and "one GPU" in msg["error"]


def test_adjoint_entry_points_validate_arguments():
    lib = L.load()
    tape = C.c_void_p()
    # null arrays, N = 0 and ckpt_every = 0 are rejected before any device work
    assert lib.sf_adjoint_forward(16, None, None, None, 10, 5, None, None, None, C.byref(tape)) == -1
    assert "invalid arguments" in L.last_error()
    assert tape.value is None
    assert lib.sf_adjoint_backward(None, None, None, None, None, None) == -1
    assert "null tape" in L.last_error()
    assert lib.sf_force_vjp_dev(16, None, None, None, None, None) == -1
    assert lib.sf_force_dev(0, None, None, None, None, None) == 0          # N = 0 is a no-op
    lib.sf_tape_free(None)


def test_lbvh_export_without_a_tree_fails_cleanly():
    lib = L.load()
    t = L.TreeT()
    assert lib.sf_lbvh_export(C.byref(t), None, None, None, None, None, None) == -1


def test_sf_run_rejects_bad_configurations_before_any_device_work(tmp_path):
    """sf_run checks its configuration first (no simulation is made)."""
    import ctypes as C
    from starflood_amd import _lib as L
    lib = L.load()
    assert lib.sf_run(None) == L.STARFLOOD_FAILURE
    L.set_params(out_dir=str(tmp_path))
    for pattern in (b"sim/step_%s", b"sim/%u_%u", b"sim/step", b"sim/%.3u", b"sim/%n"):
        cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1, sim_filename=pattern)
        assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE, pattern
        assert "%u" in L.last_error()
    cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1, enable_vis=1, vis_sizex=0, vis_sizey=8)
    assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE
    cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1, vis_file_type=7)
    assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE
    # the cubic SPH model's u is not carried by the run loop or its snapshots
    L.set_params(out_dir=str(tmp_path), enable_sph=1, sph_model=L.SF_SPH_CUBIC_IDEAL)
    cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1)
    assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE
    assert "internal energy" in L.last_error()
    assert list(tmp_path.iterdir()) == []
    L.set_params()


def test_sf_run_output_directory_percent_is_text(tmp_path):
    """A '%' in out_dir is doubled in the file-name pattern, so the name's
    one %u stays the only conversion (no device is used: the pattern checks
    come first, and the pattern-check error does not appear)."""
    import ctypes as C
    from starflood_amd import _lib as L
    lib = L.load()
    L.set_params(out_dir=str(tmp_path / "100%u"))
    cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1, enable_vis=1, vis_sizex=0, vis_sizey=0,
                      sim_filename=b"sim/step_%04u")
    # enable_vis without a frame size is checked after the patterns and fails
    # before any device work: the directory's %u did not fail the pattern
    assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE
    assert "vis_sizex" in L.last_error()
    cfg = L.RunConfig(num_bodies=16, num_timesteps=1, enable_sim=1, sim_filename=b"sim/step_%04u_%u")
    assert lib.sf_run(C.byref(cfg)) == L.STARFLOOD_FAILURE and "%u" in L.last_error()
    L.set_params()


def test_sim_init_fails_without_the_output_directory(tmp_path):
    """As the reference (src/simulation.c:58-76): with logging on, sim_init
    opens <out_dir>/statistics.csv and timings_sim_step.csv first and fails
    when the directory is missing (before any device work)."""
    import ctypes as C
    from starflood_amd import _lib as L
    lib = L.load()
    L.set_params(logging=1, out_dir=str(tmp_path / "missing"))
    s = L.SimT()
    assert lib.sim_init(C.byref(s), 16) == L.STARFLOOD_FAILURE
    assert not (tmp_path / "missing").exists()
    L.set_params()
