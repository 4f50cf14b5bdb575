"""Write the golden fixtures under tests/golden/ by running the reference's
own sources, compiled in this container into oracle/_ref/ (make -C oracle
ref).  The fixtures are data: inputs and the reference's outputs.

  f1_pp_n{256,1024}.npz   PP, eps 1e-2, dt 1e-3, the reference's own initial
                          conditions (initcond_generate): sim.mem after the
                          initial sim_solv and after 1 and 10 sim_step calls,
                          T and V of every step
  f2_config1_n4096.npz    uniform ball N = 4096, v = 0.5 x, m = 1/N, eps 1e-2,
                          dt 1e-3: initial state, SHA-256 of sim.mem after the
                          initial solve and after each of 100 steps, T and V
                          per step, final sim.mem
  f3_sph_n512.npz         PP + SPH (Gaussian kernel, polytrope), reference
                          initial conditions: after sim_solv and one step
  f4_tree_n4096.npz       reference octree at N = 4096: node arrays, acc and
                          pot after sim_solv (step 0) and after one step
  f5_pcg4d.npz            pcg4d known answers from the reference's rng.h
  f6_vis_n4096.npz        vis_draw of the F2 uniform ball after its initial
                          sim_solv at step 7, 192 x 128 pixels: the atomic and
                          render buffers, matMVP, the PPM file vis_save
                          writes and the 8-bit pixels of its PNG conversion
  f7_run_n256.npz         the reference program's run loop (src/main.c) at
                          N = 256, 10 steps, eps 1e-2, dt 1e-3, every step and
                          every 4th step: the snapshot files it writes (names,
                          SHA-256, the last one whole), statistics.csv, and
                          the step column of timings_sim_step.csv

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))

from oracle import oracle as O  # noqa: E402

GOLD = Path(__file__).resolve().parent
RNG_H = Path("/root/reference/src/rng.h")


def kahan64(x: np.ndarray) -> float:
    """The reference's FP64 Kahan sum in index order (src/simulation.c:455-497)
    applied to the reference's own per-particle ken/pen outputs."""
    s = 0.0
    c = 0.0
    for v in x.astype(np.float64).tolist():
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def stats_rows(csv: str) -> np.ndarray:
    rows = [l.split(",") for l in csv.strip().splitlines()[1:]]
    return np.array([[float(v) for v in r[1:]] for r in rows], np.float64)


def run_steps(r: O.RefSim, nsteps: int, keep=()):
    """sim_step nsteps times with the reference program's step numbering
    (src/main.c:451: step_number = step index before the step)."""
    kept = {}
    TV = []
    for n in range(nsteps):
        r.sim.step_number = n
        assert r.step() == 0
        TV.append((kahan64(r.block(11)), kahan64(r.block(12))))
        if n + 1 in keep:
            kept[n + 1] = r.mem.copy()
    return kept, np.array(TV, np.float64)


def f1():
    for N in (256, 1024):
        r = O.RefSim("pp_e1e-2_dt1e-3", N)
        assert r.initcond() == 0
        ic = r.mem.copy()
        assert r.solve() == 0
        mem0 = r.mem.copy()
        kept, TV = run_steps(r, 10, keep=(1, 10))
        csv = stats_rows(r.statistics())
        # the reference's own log agrees with the recomputed sums
        assert np.allclose(csv[:, 1], TV[:, 1], atol=2e-15) and np.allclose(csv[:, 2], TV[:, 0], atol=2e-15)
        np.savez_compressed(GOLD / f"f1_pp_n{N}.npz", ic=ic, mem0=mem0, mem1=kept[1], mem10=kept[10], TV=TV,
                            eps=1e-2, dt=1e-3)
        r.close()


def config1_ic(N: int, seed: int = 20261202) -> np.ndarray:
    """Uniform unit ball, v = 0.5 x, m = 1/N, from the library's seeded
    generator run on a host buffer (no device needed)."""
    from starflood_amd import _lib as L
    lib = L.load()
    mem = np.zeros(N * 13, np.float32)
    sim = L.SimT()
    sim.N = N
    sim.mem = mem.ctypes.data
    L.check(lib.sf_ic_uniform_ball(C.byref(sim), seed, 1.0, 1.0, 0.5), "sf_ic_uniform_ball")
    return mem


def f2():
    N = 4096
    ic = config1_ic(N)
    r = O.RefSim("pp_e1e-2_dt1e-3", N)
    r.mem[:] = ic
    r.sim.step_number = 0
    assert r.solve() == 0
    hashes = [hashlib.sha256(r.mem.tobytes()).digest()]
    TV = []
    for n in range(100):
        r.sim.step_number = n
        assert r.step() == 0
        hashes.append(hashlib.sha256(r.mem.tobytes()).digest())
        TV.append((kahan64(r.block(11)), kahan64(r.block(12))))
    np.savez_compressed(GOLD / "f2_config1_n4096.npz", ic=ic, hashes=np.frombuffer(b"".join(hashes), np.uint8)
                        .reshape(101, 32), TV=np.array(TV), final=r.mem.copy(), eps=1e-2, dt=1e-3, seed=20261202)
    r.close()


def f3():
    N = 512
    r = O.RefSim("sph_e1e-2_dt1e-3", N)
    assert r.initcond() == 0
    ic = r.mem.copy()
    assert r.solve() == 0
    mem0 = r.mem.copy()
    kept, TV = run_steps(r, 1, keep=(1,))
    np.savez_compressed(GOLD / "f3_sph_n512.npz", ic=ic, mem0=mem0, mem1=kept[1], TV=TV, eps=1e-2, dt=1e-3)
    r.close()


def f4():
    N = 4096
    r = O.RefSim("tree_e1e-2_dt1e-3", N)
    assert r.initcond() == 0
    ic = r.mem.copy()
    r.sim.step_number = 0
    assert r.solve() == 0
    nn0 = int(re.search(r"built tree with (\d+) nodes", r.last_stdout).group(1))
    nodes0 = r.tree_nodes(nn0)
    mem0 = r.mem.copy()
    r.sim.step_number = 1
    assert r.step() == 0
    nn1 = int(re.search(r"built tree with (\d+) nodes", r.last_stdout).group(1))
    nodes1 = r.tree_nodes(nn1)
    np.savez_compressed(GOLD / "f4_tree_n4096.npz", ic=ic, mem0=mem0, nodes0=nodes0, num_nodes0=nn0, mem1=r.mem.copy(),
                        nodes1=nodes1, num_nodes1=nn1, eps=1e-2, dt=1e-3, theta=0.5)
    r.close()


PCG_PROBE = r"""
#include <stdio.h>
#include "rng.h"
int main(void) {
	unsigned int in[4];
	while (4 == scanf("%u %u %u %u", &in[0], &in[1], &in[2], &in[3])) {
		uint32_t s[4] = {in[0], in[1], in[2], in[3]};
		pcg4d(s);
		printf("%u %u %u %u\n", s[0], s[1], s[2], s[3]);
	}
	return 0;
}
"""


def f5():
    rng = np.random.default_rng(5)
    ins = np.concatenate([
        np.array([[0, 0, 0, 0], [1, 2, 3, 4], [0xFFFFFFFF] * 4,
                  [0x638FE17B, 0xB10BDCE7, 0x7B8ADB9B, 0xD8C5850E],
                  [0xD88D5EFD, 0x203F83D3, 0x710BE493, 0xA294E7F1]], np.uint32),
        rng.integers(0, 2 ** 32, size=(27, 4), dtype=np.uint64).astype(np.uint32)])
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "probe.c"
        src.write_text(PCG_PROBE)
        exe = Path(d) / "probe"
        subprocess.run(["gcc", "-std=c99", "-O2", f"-I{RNG_H.parent}", str(src), "-o", str(exe)], check=True)
        text = "\n".join(" ".join(str(int(v)) for v in row) for row in ins) + "\n"
        out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout
    outs = np.array([[int(v) for v in l.split()] for l in out.strip().splitlines()], np.uint64).astype(np.uint32)
    np.savez_compressed(GOLD / "f5_pcg4d.npz", inputs=ins, outputs=outs)


def f6():
    N = 4096
    r = O.RefSim("pp_e1e-2_dt1e-3", N)
    r.mem[:] = config1_ic(N)
    r.sim.step_number = 0
    assert r.solve() == 0
    r.sim.step_number = 7
    # no PFM: in this PNG-format build the reference's PFM path writes past
    # its 3-byte-per-pixel binary buffer
    with tempfile.TemporaryDirectory() as d:
        ppm, png = Path(d) / "f.ppm", Path(d) / "f.png"
        atomic, render, mvp, png_pixels = r.vis_frame(192, 128, files=((ppm, 1), (png, 2)))
        ppm_bytes = np.frombuffer(ppm.read_bytes(), np.uint8)
    np.savez_compressed(GOLD / "f6_vis_n4096.npz", mem=r.mem.copy(), step=7, sizex=192, sizey=128, atomic=atomic,
                        render=render, mvp=mvp, ppm=ppm_bytes, png_pixels=png_pixels, eps=1e-2, dt=1e-3)
    r.close()


def f7():
    out = {}
    for key, exe in (("every", "ref_main_pp_n256_s10"), ("oi4", "ref_main_pp_n256_s10_oi4")):
        with tempfile.TemporaryDirectory() as d:
            (Path(d) / "out" / "sim").mkdir(parents=True)
            env = {"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"}
            subprocess.run([str(REPO / "oracle" / "_ref" / exe)], cwd=d, env=env, check=True,
                           stdout=subprocess.DEVNULL)
            files = sorted((Path(d) / "out" / "sim").glob("*.raw"))
            out[f"{key}_names"] = np.array([f.name for f in files])
            out[f"{key}_sha256"] = np.array([hashlib.sha256(f.read_bytes()).hexdigest() for f in files])
            out[f"{key}_last"] = np.frombuffer(files[-1].read_bytes(), np.float32).copy()
            out[f"{key}_statistics"] = np.array((Path(d) / "out" / "statistics.csv").read_text())
            tl = (Path(d) / "out" / "timings_sim_step.csv").read_text().splitlines()
            out[f"{key}_timing_steps"] = np.array([int(l.split(",")[0]) for l in tl[1:]])
    np.savez_compressed(GOLD / "f7_run_n256.npz", N=256, num_timesteps=10, eps=1e-2, dt=1e-3, **out)


if __name__ == "__main__":
    O.build(with_reference=True)
    for fn in (f1, f2, f3, f4, f5, f6, f7):
        fn()
        print("wrote", fn.__name__)
