"""The reference program's run loop (src/main.c:348-611) through sf_run,
against the files the reference program itself writes (golden fixture F7:
its snapshots, statistics.csv and the step column of timings_sim_step.csv at
N = 256, 10 steps, every step and every 4th step)."""
import hashlib

import numpy as np
import pytest

from conftest import golden
from starflood_amd import _lib as L
from starflood_amd.run import run

pytestmark = pytest.mark.gpu

PHYS = dict(enable_tree=0, epsilon=1e-2, dt=1e-3, logging=1)


def out_dir(tmp_path, name="out"):
    d = tmp_path / name
    (d / "sim").mkdir(parents=True)
    (d / "vis").mkdir()
    return d


def sha(p):
    return hashlib.sha256(p.read_bytes()).hexdigest()


@pytest.mark.parametrize("key,oi", [("every", 0), ("oi4", 4)])
def test_run_writes_the_reference_programs_files(tmp_path, key, oi):
    g = golden("f7_run_n256.npz")
    d = out_dir(tmp_path)
    print(f"sf_run N=256 steps=10 output_interval={oi}", flush=True)
    assert run(256, 10, output_interval=oi, out_dir=str(d), **PHYS) == 0
    files = sorted((d / "sim").glob("*.raw"))
    assert [f.name for f in files] == list(g[f"{key}_names"])
    assert [sha(f) for f in files] == list(g[f"{key}_sha256"])
    assert (d / "statistics.csv").read_text() == str(g[f"{key}_statistics"])
    tl = (d / "timings_sim_step.csv").read_text().splitlines()
    assert tl[0] == "step_number,kick_0,drift,sim_solv,kick_1,ken_sum,pen_sum"
    assert [int(l.split(",")[0]) for l in tl[1:]] == list(g[f"{key}_timing_steps"])
    L.set_params()


def test_run_resumes_from_a_snapshot_bit_for_bit(tmp_path):
    """INIT_COND_FILE (src/main.c:370-381): the snapshot of step 5 carries
    acc, so five more steps reproduce the step-10 snapshot exactly."""
    g = golden("f7_run_n256.npz")
    d = out_dir(tmp_path)
    assert run(256, 10, out_dir=str(d), **PHYS) == 0
    d2 = out_dir(tmp_path, "resume")
    assert run(256, 5, init_cond_file=str(d / "sim" / "step_0005.raw"), out_dir=str(d2), **PHYS) == 0
    assert sha(d2 / "sim" / "step_0005.raw") == g["every_sha256"][10]
    L.set_params()


def test_stop_file_ends_the_run_and_is_removed(tmp_path):
    d = out_dir(tmp_path)
    (d / "stop").write_text("")
    assert run(256, 10, out_dir=str(d), **PHYS) == 0
    assert not (d / "stop").exists()
    assert list((d / "sim").glob("*.raw")) == []
    L.set_params()


def test_replay_reads_snapshots_and_draws_frames(tmp_path):
    """ENABLE_SIM off (src/main.c:414-421, 503-514): steps advance by the
    output interval and snapshot step / interval is read, drawn and saved;
    a missing snapshot ends the run early (return 1)."""
    d = out_dir(tmp_path)
    assert run(256, 10, out_dir=str(d), **PHYS) == 0
    kw = dict(output_interval=4, enable_sim=False, enable_vis=True, vis_filename="vis/step_%04u",
              vis_file_type=L.IMAGE_FILETYPE_PPM, vis_size=(64, 48), out_dir=str(d), **PHYS)
    assert run(256, 10, **kw) == 0
    frames = sorted((d / "vis").glob("*.ppm"))
    assert [f.name for f in frames] == ["step_0000.ppm", "step_0001.ppm", "step_0002.ppm"]
    header = b"P6\n64 48 255\n"
    for f in frames:
        b = f.read_bytes()
        assert b.startswith(header) and len(b) == len(header) + 64 * 48 * 3
        assert max(b[len(header):]) > 0   # the particles are in the frame
    (d / "sim" / "step_0002.raw").unlink()
    assert run(256, 10, **kw) == 1
    L.set_params()
