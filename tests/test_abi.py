"""The drop-in boundary without a GPU: the library loads, exports exactly the
functions include/*.h declares, lays out sim_t / tree_t / node_t and enum
sim_conf as the reference does, and its host-only entry points (parameters,
sim_find, initial conditions, raw snapshots) behave as the reference's."""
import ctypes as C
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from conftest import REPO, bits, golden
from starflood_amd import _lib as L
from starflood_amd.build import public_symbols

REF_SRC = Path("/root/reference/src")


def exported(path) -> set:
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_library_exports_exactly_the_header_declarations():
    lib = L.load()
    names = set(public_symbols())
    assert names == exported(L.LIB_PATH)
    assert set(L.exported_names()) == names
    for n in names:
        assert getattr(lib, n) is not None


VIS_PROTO_PROBE = r"""
#include "starflood.h"
#ifdef SOLVER_VIS
/* the reference's prototypes, src/simulation.h:92,95 and src/gravity.h:15 */
int (*p_solv)(sim_t* restrict, f32* restrict, unsigned int, unsigned int) = sim_solv;
int (*p_step)(sim_t* restrict, f32* restrict, unsigned int, unsigned int) = sim_step;
int (*p_tree)(const sim_t* restrict, f32*, unsigned int, unsigned int) = solve_gravity_part_tree;
#else
/* src/simulation.h:98,101 and src/gravity.h:17 */
int (*p_solv)(sim_t* restrict) = sim_solv;
int (*p_step)(sim_t* restrict) = sim_step;
int (*p_tree)(const sim_t* restrict) = solve_gravity_part_tree;
#endif
int main(void) { return 0; }
"""


@pytest.mark.parametrize("solver_vis", [False, True])
def test_solver_vis_prototypes_and_library(solver_vis):
    """Both prototype forms of the reference compile against include/ (a
    mismatch is an error), and the SOLVER_VIS library exports the same names."""
    from starflood_amd.build import LIB_SOLVERVIS
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "p.c"
        src.write_text(VIS_PROTO_PROBE)
        flags = ["-DSOLVER_VIS"] if solver_vis else []
        r = subprocess.run(["gcc", "-std=c99", "-Werror=incompatible-pointer-types", *flags, f"-I{REPO / 'include'}",
                            "-c", str(src), "-o", str(Path(d) / "p.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    if solver_vis:
        assert exported(LIB_SOLVERVIS) == set(public_symbols())


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HEADER
int main(void) {
	sim_t s; tree_t t; node_t n;
	printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %d %d %d\n", sizeof(sim_t), sizeof(tree_t), sizeof(node_t),
	       offsetof(sim_t, mem), offsetof(sim_t, tree), offsetof(tree_t, node), offsetof(node_t, bounds),
	       offsetof(node_t, params), offsetof(tree_t, bounds), (int)SIM_MAS, (int)SIM_PEN, (int)SIM_DOF);
	(void)s; (void)t; (void)n;
	return 0;
}
"""


def probe_layout(include_dir: Path, header: str, defines=()) -> str:
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "p.c"
        src.write_text(LAYOUT_PROBE.replace("HEADER", f'"{header}"'))
        exe = Path(d) / "p"
        subprocess.run(["gcc", "-std=c99", *defines, f"-I{include_dir}", str(src), "-o", str(exe)], check=True,
                       capture_output=True)
        return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()


def test_types_match_reference_layout():
    mine = probe_layout(REPO / "include", "starflood.h")
    assert mine[:3] == ["56", "40", "80"]
    assert mine[-3:] == ["9", "12", "13"]
    mine_sph = probe_layout(REPO / "include", "starflood.h", ["-DENABLE_SPH"])
    assert mine_sph[-1] == "16"
    # ctypes mirror agrees
    assert C.sizeof(L.SimT) == 56 and C.sizeof(L.TreeT) == 40 and C.sizeof(L.NodeT) == 80
    if REF_SRC.is_dir():
        # the reference's headers under its default configuration
        assert probe_layout(REF_SRC, "simulation.h") == mine
        assert probe_layout(REF_SRC, "simulation.h", ["-DENABLE_SPH"])[-1] == "16"


def test_default_parameters_are_the_reference_config():
    p = L.default_params()
    assert (p.epsilon, p.G, p.dt, p.theta) == (1.0e-3, 1.0, 1.0, 0.5)
    assert (p.tree_nodes_max, p.tree_depth_max) == (262144, 16)
    assert (p.enable_tree, p.enable_sph, p.kahan_solver, p.kahan_energy) == (1, 0, 1, 1)
    assert (p.precision, p.acc_mode, p.tree_mode, p.mem_mode) == (0, 0, 0, 0)
    assert p.out_dir == b"./out"
    if REF_SRC.is_dir():
        cfg = (REF_SRC / "config.h").read_text()
        assert "#define EPSILON 1.000e-3" in cfg and "#define TREE_NODES_MAX 262144" in cfg
        assert "#define TIMESTEP_SIZE 1.000e-0" in cfg and "#define TREE_THETA 0.5" in cfg


def test_set_params_rejects_bad_values():
    lib = L.load()
    p = L.default_params()
    p.tree_depth_max = 17
    assert lib.sf_set_params(C.byref(p)) == L.STARFLOOD_FAILURE
    assert "tree_depth_max" in L.last_error()
    p = L.default_params()
    p.precision = 7
    assert lib.sf_set_params(C.byref(p)) == L.STARFLOOD_FAILURE
    for bad in (-1e-3, float("nan"), float("inf")):
        p = L.default_params()
        p.mac_alpha = bad
        assert lib.sf_set_params(C.byref(p)) == L.STARFLOOD_FAILURE
        assert "mac_alpha" in L.last_error()
    p = L.default_params()
    assert p.mac_alpha == 0.0                       # off: the theta test alone, as the reference
    p.mac_alpha = 3e-3
    assert lib.sf_set_params(C.byref(p)) == L.STARFLOOD_SUCCESS
    L.set_params()


def host_sim(N, dof=13):
    mem = np.zeros(N * dof, np.float32)
    s = L.SimT()
    s.N = N
    s.mem = mem.ctypes.data
    return s, mem


def test_sim_find_offsets():
    lib = L.load()
    L.set_params()
    s, mem = host_sim(100)
    base = mem.ctypes.data
    for e, off in [(0, 0), (1, 0), (2, 0), (3, 300), (5, 300), (6, 600), (8, 600), (9, 900), (12, 1200)]:
        assert C.cast(lib.sim_find(C.byref(s), e), C.c_void_p).value == base + 4 * off
    assert not lib.sim_find(C.byref(s), 13)
    L.set_params(enable_sph=1)
    s2, mem2 = host_sim(100, 16)
    assert C.cast(lib.sim_find(C.byref(s2), 15), C.c_void_p).value == mem2.ctypes.data + 4 * 1500
    L.set_params()


@pytest.mark.parametrize("N", [256, 1024])
def test_initcond_generate_matches_reference_bytes(N):
    lib = L.load()
    L.set_params()
    s, mem = host_sim(N)
    L.check(lib.initcond_generate(C.byref(s)), "initcond_generate")
    assert np.array_equal(bits(mem), bits(golden(f"f1_pp_n{N}.npz")["ic"]))


def test_initcond_generate_sph_matches_reference_bytes():
    lib = L.load()
    L.set_params(enable_sph=1)
    s, mem = host_sim(512, 16)
    L.check(lib.initcond_generate(C.byref(s)), "initcond_generate")
    assert np.array_equal(bits(mem), bits(golden("f3_sph_n512.npz")["ic"]))
    L.set_params()


def test_uniform_ball_generator_reproduces_config1_inputs():
    lib = L.load()
    L.set_params()
    g = golden("f2_config1_n4096.npz")
    s, mem = host_sim(4096)
    L.check(lib.sf_ic_uniform_ball(C.byref(s), int(g["seed"]), 1.0, 1.0, 0.5), "ic")
    assert np.array_equal(bits(mem), bits(g["ic"]))
    pos = mem[:3 * 4096].reshape(-1, 3)
    assert np.all(np.linalg.norm(pos, axis=1) <= 1.0 + 1e-6)
    assert np.allclose(mem[3 * 4096:6 * 4096], 0.5 * mem[:3 * 4096], rtol=1e-6)
    assert np.all(mem[9 * 4096:10 * 4096] == np.float32(1.0 / 4096))


def test_plummer_generator_statistics():
    lib = L.load()
    L.set_params()
    N = 20000
    s, mem = host_sim(N)
    L.check(lib.sf_ic_plummer(C.byref(s), 9, 0, N, 1.0, 1.0, None, None), "plummer")
    pos = mem[:3 * N].reshape(N, 3).astype(np.float64)
    vel = mem[3 * N:6 * N].reshape(N, 3).astype(np.float64)
    r = np.linalg.norm(pos, axis=1)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.3048 a
    assert np.median(r) == pytest.approx(1.3048, rel=0.03)
    # virial equilibrium: 2T = -W, W = -3 pi G M^2 / (32 a)
    T = 0.5 * np.sum(vel ** 2) / N
    assert 2 * T == pytest.approx(3 * np.pi / 32, rel=0.05)
    # isotropy
    assert np.abs(pos.mean(0)).max() < 0.1 and np.abs(vel.mean(0)).max() < 0.02


def test_raw_snapshot_round_trip(tmp_path):
    lib = L.load()
    L.set_params()
    s, mem = host_sim(64)
    mem[:] = np.random.default_rng(0).normal(size=mem.size).astype(np.float32)
    f = str(tmp_path / "snap.raw").encode()
    L.check(lib.sim_save_raw(C.byref(s), f), "save")
    assert Path(f.decode()).stat().st_size == 64 * 13 * 4      # headerless sim.mem
    s2, mem2 = host_sim(64)
    L.check(lib.sim_read_raw(C.byref(s2), f), "read")
    assert np.array_equal(bits(mem), bits(mem2))
    # a short file is an error
    Path(f.decode()).write_bytes(b"\0" * 100)
    assert lib.sim_read_raw(C.byref(s2), f) == L.STARFLOOD_FAILURE


def test_entry_points_reject_foreign_sim():
    lib = L.load()
    L.set_params()
    s, mem = host_sim(8)
    assert lib.sim_step(C.byref(s)) == L.STARFLOOD_FAILURE
    assert "sim_init" in L.last_error()
    assert lib.solve_gravity_part_part(C.byref(s)) == L.STARFLOOD_FAILURE


@pytest.mark.parametrize("solver_vis", [False, True])
def test_reference_style_c_driver_links_against_the_library(tmp_path, solver_vis):
    """tests/c/dropin_main.c, a driver making the reference program's calls
    (src/main.c:358-597), compiles against include/ and links against the
    library alone, as the reference's Makefile would link it; every
    reference entry point it calls is resolved from the library."""
    lib = "starflood_hip_solvervis" if solver_vis else "starflood_hip"
    exe = tmp_path / "dropin"
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{REPO / 'include'}",
           str(REPO / "tests" / "c" / "dropin_main.c"), "-o", str(exe), f"-L{L.LIB_PATH.parent}", f"-l{lib}"]
    if solver_vis:
        cmd.insert(1, "-DSOLVER_VIS")
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True,
                         check=True).stdout
    undefined = {l.split()[-1].split("@")[0] for l in out.splitlines() if l.strip()}
    calls = {"sim_init", "initcond_generate", "sim_solv", "sim_step", "sim_save_raw", "sim_free",
             "sf_params_default", "sf_set_params"}
    assert calls <= undefined
    assert calls <= exported(L.LIB_PATH.parent / f"lib{lib}.so")
