"""PP gravity and the KDK step on the device through the C ABI, against the
reference's own outputs (golden fixtures) and the oracle."""
import hashlib

import numpy as np
import pytest

from conftest import bits, golden
from starflood_amd import _lib as L
from starflood_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def exact_sim(N, **kw):
    args = dict(enable_tree=0, logging=0, epsilon=1e-2, dt=1e-3)
    args.update(kw)
    return Simulation(N, **args)


@pytest.mark.parametrize("N", [256, 1024])
def test_exact_pp_and_kdk_bit_identical_to_reference(N):
    g = golden(f"f1_pp_n{N}.npz")
    with exact_sim(N) as s:
        s.mem[:] = g["ic"]
        s.solve()
        assert np.array_equal(bits(s.mem), bits(g["mem0"]))
        for n in range(10):
            s.step_number = n
            s.step()
            st = s.stats()
            assert (st["T"], st["V"]) == tuple(g["TV"][n]), f"step {n}"
            if n == 0:
                assert np.array_equal(bits(s.mem), bits(g["mem1"]))
        assert np.array_equal(bits(s.mem), bits(g["mem10"]))


def test_config1_100_steps_bit_identical_in_step_order():
    g = golden("f2_config1_n4096.npz")
    with exact_sim(4096) as s:
        s.mem[:] = g["ic"]
        s.solve()
        assert hashlib.sha256(s.mem.tobytes()).digest() == g["hashes"][0].tobytes()
        for n in range(100):
            s.step_number = n
            s.step()
            assert hashlib.sha256(s.mem.tobytes()).digest() == g["hashes"][n + 1].tobytes(), f"step {n}"
        assert np.array_equal(bits(s.mem), bits(g["final"]))


def test_solver_vis_library_reproduces_the_reference_steps():
    """libstarflood_hip_solvervis.so (the SOLVER_VIS prototypes, render
    buffer ignored) gives the reference's bits, as the plain library does."""
    import ctypes as C
    import os
    from starflood_amd.build import LIB_SOLVERVIS
    g = golden("f1_pp_n256.npz")
    lib = C.CDLL(str(LIB_SOLVERVIS), mode=os.RTLD_NOW | C.RTLD_LOCAL)
    vis_sig = [C.POINTER(L.SimT), C.c_void_p, C.c_uint, C.c_uint]
    lib.sim_solv.argtypes = lib.sim_step.argtypes = vis_sig
    lib.sim_init.argtypes = [C.POINTER(L.SimT), C.c_uint]
    lib.sim_free.argtypes = [C.POINTER(L.SimT)]
    lib.sf_params_default.argtypes = lib.sf_set_params.argtypes = [C.POINTER(L.Params)]
    p = L.Params()
    lib.sf_params_default(C.byref(p))
    p.enable_tree, p.epsilon, p.dt, p.logging = 0, 1e-2, 1e-3, 0
    assert lib.sf_set_params(C.byref(p)) == 0
    s = L.SimT()
    assert lib.sim_init(C.byref(s), 256) == 0
    mem = np.ctypeslib.as_array((C.c_float * (256 * 13)).from_address(s.mem))
    mem[:] = g["ic"]
    frame = (C.c_float * (4 * 8 * 8))()
    assert lib.sim_solv(C.byref(s), frame, 8, 8) == 0
    assert np.array_equal(bits(mem), bits(g["mem0"]))
    for n in range(10):
        s.step_number = n
        assert lib.sim_step(C.byref(s), frame, 8, 8) == 0
    assert np.array_equal(bits(mem), bits(g["mem10"]))
    assert lib.sim_free(C.byref(s)) == 0


def test_exact_resident_mode_matches_mirror_mode():
    g = golden("f1_pp_n256.npz")
    with exact_sim(256, mem_mode=L.SF_MEM_RESIDENT) as s:
        s.mem[:] = g["ic"]
        s.to_device()
        s.solve()
        for n in range(10):
            s.step_number = n
            s.step()
        s.to_host()
        assert np.array_equal(bits(s.mem), bits(g["mem10"]))
    L.set_params()


@pytest.mark.parametrize("kahan", [0, 1])
@pytest.mark.parametrize("acc_mode", [0, 1])
def test_exact_variants_match_oracle(oracle, kahan, acc_mode):
    rng = np.random.default_rng(10 + kahan + 2 * acc_mode)
    N = 1000
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    mas = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
    with exact_sim(N, kahan_solver=kahan, kahan_energy=kahan, acc_mode=acc_mode, G=0.7) as s:
        s.pos[:] = pos
        s.mas[:] = mas
        s.solve_gravity_pp()
        acc, pot = oracle.pp_ref(pos, mas, 1e-2, G=0.7, kahan_acc=kahan, kahan_pot=kahan, acc_phys=acc_mode)
        assert np.array_equal(bits(s.acc.ravel()), bits(acc.ravel()))
        assert np.array_equal(bits(s.pot), bits(pot))


def test_single_particle_and_zero_mass_edge_cases(oracle):
    with exact_sim(1) as s:
        s.pos[:] = [[0.5, -1.0, 2.0]]
        s.mas[:] = [3.0]
        s.solve_gravity_pp()
        assert np.all(s.acc == 0)
        eps2 = np.float32(0.01 * 0.01)
        assert s.pot[0] == np.float32(-1.0) * (np.float32(3.0) * (np.float32(1.0) / np.sqrt(eps2)))
    with exact_sim(5, precision=L.SF_PRECISION_FAST) as s:
        s.pos[:] = np.arange(15, dtype=np.float32).reshape(5, 3)
        s.mas[:] = 0.0
        s.solve_gravity_pp()
        assert np.all(s.acc == 0) and np.all(s.pot == 0)
    L.set_params()


@pytest.mark.parametrize("N", [1, 100, 1000, 4097])
def test_fast_pp_close_to_fp64_small(oracle, N):
    rng = np.random.default_rng(N)
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    mas = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
    with exact_sim(N, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL) as s:
        s.pos[:] = pos
        s.mas[:] = mas
        s.solve_gravity_pp()
        a64, p64 = oracle.pp_f64(pos, mas, 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max() + 1e-30
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale
        assert np.allclose(s.pot, p64, rtol=2e-6)
    L.set_params()


@pytest.mark.parametrize("N", [127, 128, 1000, 4097])
@pytest.mark.parametrize("masses", ["equal", "one_differs", "zero_first"])
def test_fast_pp_equal_mass_tiles_close_to_fp64(oracle, N, masses):
    """Equal masses take the kernel's mass-free tiles (whole 128-body tiles of
    real particles; the tile holding the padding keeps the general sum); one
    differing mass anywhere sends every tile through the general sum."""
    rng = np.random.default_rng(1000 + N)
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    mas = np.full(N, 1.0 / 3.0 / N, dtype=np.float32)
    if masses == "one_differs":
        mas[N // 2] *= np.float32(1.5)
    elif masses == "zero_first":
        mas[0] = 0.0
    with exact_sim(N, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL) as s:
        s.pos[:] = pos
        s.mas[:] = mas
        s.solve_gravity_pp()
        a64, p64 = oracle.pp_f64(pos, mas, 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max() + 1e-30
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale
        assert np.allclose(s.pot, p64, rtol=2e-6)
    L.set_params()


def test_pack_state_stays_consistent_over_many_launches():
    """The pack kernel's equal-mass detection keeps its block-count words in
    device memory and returns them to zero itself.  After hundreds of packs,
    and after a mass changes mid-run (the general sum must take over, then
    the equal-mass tiles again), the force equals a fresh context's on the
    same state bit for bit.  The masses are not a power of two, so the two
    sums round differently and a stale decision would show."""
    N = 4096
    rng = np.random.default_rng(77)
    kw = dict(precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL, mem_mode=L.SF_MEM_RESIDENT)

    def fresh_force(mem):
        with exact_sim(N, **kw) as b:
            b.mem[:] = mem
            b.to_device()
            b.solve_gravity_pp()
            b.to_host()
            return b.acc.copy(), b.pot.copy()

    with exact_sim(N, **kw) as a:
        a.pos[:] = rng.normal(size=(N, 3)).astype(np.float32)
        a.mas[:] = np.float32(1.0 / 3.0 / N)
        a.to_device()
        a.solve()
        for phase in ("equal", "one_differs", "equal_again"):
            if phase != "equal":
                a.to_host()
                a.mas[7] = np.float32(1.0 / 3.0 / N) * (np.float32(1.25) if phase == "one_differs" else 1)
                a.to_device()
            for _ in range(150):
                a.step()
            a.to_host()
            state = a.mem.copy()
            a.solve_gravity_pp()
            a.to_host()
            acc_b, pot_b = fresh_force(state)
            assert np.array_equal(bits(a.acc.ravel()), bits(acc_b.ravel())), phase
            assert np.array_equal(bits(a.pot), bits(pot_b)), phase
    L.set_params()


def test_f64_direct_sum_kernel_matches_cpu_oracle(oracle):
    N = 8192
    with exact_sim(N, epsilon=2e-3, acc_mode=L.SF_ACC_PHYSICAL) as s:
        s.ic_plummer(seed=3, first=0, count=N)
        idx = np.arange(0, N, 37, dtype=np.uint32)
        a_gpu, p_gpu = s.direct_f64(idx)
        a_cpu, p_cpu = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 2e-3, acc_phys=1, idx=idx)
        assert np.allclose(a_gpu, a_cpu, rtol=1e-11, atol=1e-14)
        assert np.allclose(p_gpu, p_cpu, rtol=1e-12)
    L.set_params()


def test_fast_pp_accuracy_at_262144():
    """The throughput kernel at the stress size against the FP64 direct sum
    on 4096 sampled particles: relative acceleration error <= 1e-5."""
    N = 262144
    with exact_sim(N, epsilon=2e-3, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL,
                   mem_mode=L.SF_MEM_RESIDENT) as s:
        s.ic_plummer(seed=11, first=0, count=N)
        s.solve_gravity_pp()
        s.to_host()
        idx = np.arange(0, N, N // 4096, dtype=np.uint32)
        a64, p64 = s.direct_f64(idx)
        rel = np.linalg.norm(s.acc[idx] - a64, axis=1) / np.linalg.norm(a64, axis=1)
        print(f"rel acc error p50 {np.percentile(rel, 50):.3g} p99 {np.percentile(rel, 99):.3g} max {rel.max():.3g}")
        assert rel.max() <= 1e-5
        assert np.abs(s.pot[idx] / p64 - 1).max() <= 1e-6
        # Newton's third law for equal masses: sum of accelerations ~ 0
        assert np.linalg.norm(s.acc.astype(np.float64).sum(0)) / N <= 1e-6 * np.abs(s.acc).max()
    L.set_params()


def test_fast_kdk_reversible_and_deterministic():
    N = 2048
    with exact_sim(N, epsilon=1e-2, dt=1e-3, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL) as s:
        s.ic_plummer(seed=5, first=0, count=N)
        x0 = s.pos.copy()
        v0 = s.vel.copy()
        s.solve()
        for n in range(20):
            s.step()
        x1 = s.pos.copy()
        s.vel[:] = -s.vel
        for n in range(20):
            s.step()
        assert np.abs(s.pos - x0).max() < 2e-5
        assert np.abs(-s.vel - v0).max() < 2e-5
    with exact_sim(N, epsilon=1e-2, dt=1e-3, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL) as s2:
        s2.ic_plummer(seed=5, first=0, count=N)
        s2.solve()
        for n in range(20):
            s2.step()
        assert np.array_equal(bits(s2.pos), bits(x1))
    L.set_params()


def test_logging_writes_reference_csv_schema(tmp_path):
    g = golden("f1_pp_n256.npz")
    with Simulation(256, enable_tree=0, epsilon=1e-2, dt=1e-3, logging=1, out_dir=str(tmp_path)) as s:
        s.mem[:] = g["ic"]
        s.solve()
        for n in range(3):
            s.step_number = n
            s.step()
    lines = (tmp_path / "statistics.csv").read_text().splitlines()
    assert lines[0] == "step_number,total_energy,potential_energy,kinetic_energy"
    row = lines[2].split(",")
    T, V = g["TV"][1]
    assert row[0] == "1" and row[2] == f"{V:.015f}" and row[3] == f"{T:.015f}"
    tl = (tmp_path / "timings_sim_step.csv").read_text().splitlines()
    assert tl[0] == "step_number,kick_0,drift,sim_solv,kick_1,ken_sum,pen_sum" and len(tl) == 4
    L.set_params()


def test_stats_kinetic_energy_and_momenta_match_host_fp64():
    """sf_get_stats: K = 0.5 sum m v^2 of the current velocities (also before
    any step has written ken), momentum and angular momentum, in FP64."""
    N = 3000
    with exact_sim(N, precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL) as s:
        s.ic_plummer(seed=4, first=0, count=N)
        s.solve()
        for k in range(2):
            st = s.stats()
            m = s.mas.astype(np.float64)
            v = s.vel.astype(np.float64)
            x = s.pos.astype(np.float64)
            K = 0.5 * (m * (v * v).sum(1)).sum()
            assert st["K"] == pytest.approx(K, rel=1e-12)
            assert np.allclose(st["P"], (m[:, None] * v).sum(0), rtol=1e-9, atol=1e-15)
            assert np.allclose(st["L"], (m[:, None] * np.cross(x, v)).sum(0), rtol=1e-7, atol=1e-12)
            s.step()
