"""SPH on the device through the C ABI: the reference's Gaussian/polytrope
model against the reference's own outputs (fixture F3) and the oracle, and
the cubic-spline ideal-gas model against the FP64 oracle and its
conservation laws.

The reference model's exp() and powf() come from glibc on the CPU and from
the device math library here; both are faithfully rounded but not always
equal, so these comparisons carry ulp tolerances (stated per test) instead
of bit equality."""
import numpy as np
import pytest

from conftest import SPH_GAMMA, SPH_RHO0, golden, sph_lattice, ulp_diff
from starflood_amd import _lib as L
from starflood_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def ref_sph_sim(N, **kw):
    args = dict(enable_tree=0, enable_sph=1, logging=0, epsilon=1e-2, dt=1e-3)
    args.update(kw)
    return Simulation(N, **args)


def rel_vec_err(a, b):
    """|a_i - b_i| / max(|b_i|, 1e-3 max |b|), per particle."""
    nb = np.linalg.norm(b, axis=1)
    den = np.maximum(nb, 1e-3 * nb.max())
    return np.linalg.norm(a - b, axis=1) / np.where(den > 0, den, 1.0)   # all zero: absolute error


def test_reference_model_matches_reference_fixture():
    """F3: N=512, initcond_generate ICs (h = 0.1), sim_solv then one step.
    Tolerances: rho <= 4 ulp, prs <= 8 ulp, acc <= 2e-5 relative (the
    force sum cancels, so ulp differences of single terms show up
    relative to the total); one step later pos <= 4 ulp and vel (kicked by
    those accelerations) <= 2e-5 relative."""
    g = golden("f3_sph_n512.npz")
    N = 512
    with ref_sph_sim(N) as s:
        s.mem[:] = g["ic"]
        s.solve()
        m0 = g["mem0"].reshape(16, N)
        got = s.mem.reshape(16, N)
        assert ulp_diff(got[14], m0[14]).max() <= 4        # rho
        assert ulp_diff(got[15], m0[15]).max() <= 8        # prs
        acc_ref = m0[6:9].reshape(N, 3)
        assert rel_vec_err(s.acc, acc_ref).max() <= 2e-5
        assert ulp_diff(got[10], m0[10]).max() == 0        # pot: pure gravity, bit-exact
        s.step_number = 0
        s.step()
        m1 = g["mem1"].reshape(16, N)
        got = s.mem.reshape(16, N)
        assert ulp_diff(got[0:3], m1[0:3]).max() <= 4      # pos
        assert rel_vec_err(s.vel, m1[3:6].reshape(N, 3)).max() <= 2e-5
        assert ulp_diff(got[14], m1[14]).max() <= 4


@pytest.mark.parametrize("kahan", [1, 0])
def test_reference_model_matches_oracle_varied_h(oracle, kahan):
    """Random masses and smoothing radii (N=1536), with and without Kahan
    sums; same tolerances as the fixture test."""
    N = 1536
    rng = np.random.default_rng(31 + kahan)
    pos = rng.normal(scale=0.4, size=(N, 3)).astype(np.float32)
    mas = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
    rad = rng.uniform(0.05, 0.3, N).astype(np.float32)
    acc0 = rng.normal(size=(N, 3)).astype(np.float32)
    rho_o, prs_o, acc_o = oracle.sph_ref(pos, mas, rad, acc0, kahan=kahan)
    with ref_sph_sim(N, kahan_solver=kahan) as s:
        s.pos[:] = pos
        s.mas[:] = mas
        s.rad[:] = rad
        s.acc[:] = acc0
        s.solve_sph()
        assert ulp_diff(s.rho, rho_o).max() <= (4 if kahan else 16)
        assert ulp_diff(s.prs, prs_o).max() <= (8 if kahan else 32)
        assert rel_vec_err(s.acc - acc0, acc_o - acc0).max() <= (2e-5 if kahan else 1e-4)


def cubic_sim(N, **kw):
    args = dict(enable_tree=0, enable_sph=1, sph_model=L.SF_SPH_CUBIC_IDEAL, logging=0, epsilon=1e-2, dt=5e-4,
                precision=L.SF_PRECISION_FAST, acc_mode=L.SF_ACC_PHYSICAL, mem_mode=L.SF_MEM_RESIDENT)
    args.update(kw)
    return Simulation(N, **args)


def cubic_state(N, seed, h):
    rng = np.random.default_rng(seed)
    r = rng.random(N) ** (1 / 3)
    d = rng.normal(size=(N, 3))
    pos = (d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]).astype(np.float32)
    vel = rng.normal(scale=0.1, size=(N, 3)).astype(np.float32)
    mas = np.full(N, 1.0 / N, np.float32)
    rad = np.full(N, h, np.float32)
    u = rng.uniform(0.02, 0.08, N).astype(np.float32)
    return pos, vel, mas, rad, u


def check_cubic_against_oracle(oracle, pos, vel, mas, rad, u, search):
    """rho, P within 2e-5 relative; SPH acceleration and du/dt within 1e-4
    of the per-particle FP64 magnitude (FP32 pair sums with rsqrt)."""
    N = len(mas)
    rho_o, prs_o, acc_o, du_o = oracle.sph_cubic_f64(pos, vel, mas, rad, u, 5.0 / 3.0)
    with cubic_sim(N, mem_mode=L.SF_MEM_MIRROR, sph_search=search) as s:
        s.pos[:] = pos
        s.vel[:] = vel
        s.mas[:] = mas
        s.rad[:] = rad
        s.acc[:] = 0
        s.set_u(u)
        s.solve_sph()
        # h = 0 particles without a partner have rho = 0: relative to 1e-3 of the largest
        assert np.max(np.abs(s.rho - rho_o) / np.maximum(rho_o, 1e-3 * rho_o.max())) <= 2e-5
        assert np.max(np.abs(s.prs - prs_o) / np.maximum(prs_o, 1e-3 * prs_o.max())) <= 2e-5
        assert rel_vec_err(s.acc.astype(np.float64), acc_o).max() <= 1e-4
        du = s.get_dudt().astype(np.float64)
        scale = np.maximum(np.abs(du_o), 1e-3 * np.abs(du_o).max())
        assert np.max(np.abs(du - du_o) / np.where(scale > 0, scale, 1.0)) <= 1e-4


@pytest.mark.parametrize("search", [L.SF_SPH_SEARCH_GRID, L.SF_SPH_SEARCH_ALL])
@pytest.mark.parametrize("h", [0.05, 0.15, 0.3])
def test_cubic_model_matches_fp64_oracle(oracle, h, search):
    """Both pair searches (cell grid, all pairs) against the FP64 all-pairs
    oracle, N = 4096 with mixed smoothing lengths."""
    pos, vel, mas, rad, u = cubic_state(4096, 5, h)
    rad[::7] *= 1.5
    check_cubic_against_oracle(oracle, pos, vel, mas, rad, u, search)


def test_grid_search_on_outliers_duplicates_and_zero_h(oracle):
    """The cell grid where it is stressed: pairs of particles 2e5 and 3e5 away
    from the ball and 0.05 apart (cell coordinates clamped to 21 bits),
    twelve coincident particles, and particles with h = 0, which interact only
    through their partners' h.  Same tolerances against the FP64 oracle."""
    N = 3000
    pos, vel, mas, rad, u = cubic_state(N, 17, 0.1)
    pos[0] = (2.0e5, 1.0, -2.0)
    pos[1] = (2.0e5 + 0.05, 1.0, -2.0)
    pos[2] = (-3.0e5, -3.0e5, 5.0)
    pos[3] = (-3.0e5, -3.0e5 + 0.04, 5.0)
    pos[10:22] = pos[30]
    rad[40:60] = 0.0
    check_cubic_against_oracle(oracle, pos, vel, mas, rad, u, L.SF_SPH_SEARCH_GRID)


def test_cubic_model_conserves_momentum_and_energy_rate():
    """Pair terms are antisymmetric: sum m a = 0 and sum m (v.a) + sum m du/dt
    = 0 up to FP32 rounding of the per-particle sums."""
    N = 8192
    pos, vel, mas, rad, u = cubic_state(N, 9, 0.12)
    with cubic_sim(N, mem_mode=L.SF_MEM_MIRROR) as s:
        s.pos[:] = pos
        s.vel[:] = vel
        s.mas[:] = mas
        s.rad[:] = rad
        s.acc[:] = 0
        s.set_u(u)
        s.solve_sph()
        a = s.acc.astype(np.float64)
        m = mas.astype(np.float64)[:, None]
        p_rate = (m * a).sum(0)
        assert np.abs(p_rate).max() <= 1e-5 * (m * np.abs(a)).sum()
        du = s.get_dudt().astype(np.float64)
        e_rate = (m * vel * a).sum() + (m[:, 0] * du).sum()
        assert abs(e_rate) <= 1e-5 * ((m * np.abs(vel * a)).sum() + (m[:, 0] * np.abs(du)).sum())


def test_cubic_gas_sphere_conserves_energy_and_momentum():
    """A cold-ish gas ball with self-gravity (the PP + SPH step of the
    N=65536 workload at N=8192): 40 KDK steps at dt = 5e-4 keep
    E = K + W + U within 1e-4 relative and the momentum at the FP64 rounding
    floor of its sum."""
    N = 8192
    pos, vel, mas, rad, u = cubic_state(N, 11, 0.1)
    vel[:] = 0
    u[:] = 0.05
    with cubic_sim(N) as s:
        s.pos[:] = pos
        s.vel[:] = vel
        s.mas[:] = mas
        s.rad[:] = rad
        s.to_device()
        s.set_u(u)
        s.solve()
        e0 = s.stats()
        E0 = e0["K"] + e0["W"] + e0["U"]
        for n in range(40):
            s.step_number = n
            s.step()
        e1 = s.stats()
        E1 = e1["K"] + e1["W"] + e1["U"]
        assert abs(E1 - E0) <= 1e-4 * abs(E0)
        assert e1["U"] != e0["U"]                      # the gas did work
        assert np.abs(e1["P"]).max() <= 1e-6


def test_fp64_sph_reference_matches_the_oracle(oracle):
    """sf_sph_direct_f64, the device's FP64 precision reference for the
    cubic model (used by the benchmark), against the CPU oracle: rho, the
    pressure acceleration and du/dt within 1e-12 relative (FP64 sums in a
    different order), mixed smoothing lengths."""
    N = 4096
    pos, vel, mas, rad, u = cubic_state(N, 5, 0.15)
    rad[::7] *= 1.5
    rows = np.arange(1, N, 3, dtype=np.uint32)
    rho_o, _, acc_o, du_o = oracle.sph_cubic_f64(pos, vel, mas, rad, u, 5.0 / 3.0)
    with cubic_sim(N, mem_mode=L.SF_MEM_MIRROR) as s:
        s.pos[:] = pos
        s.vel[:] = vel
        s.mas[:] = mas
        s.rad[:] = rad
        s.set_u(u)
        rho, acc, du = s.sph_direct_f64(rows)
    assert np.max(np.abs(rho - rho_o[rows]) / rho_o[rows]) <= 1e-12
    na = np.linalg.norm(acc_o[rows], axis=1)
    assert np.max(np.linalg.norm(acc - acc_o[rows], axis=1) / np.maximum(na, 1e-6 * na.max())) <= 1e-12
    assert np.max(np.abs(du - du_o[rows]) / np.maximum(np.abs(du_o[rows]), 1e-6 * np.abs(du_o).max())) <= 1e-12


@pytest.mark.parametrize("search", [L.SF_SPH_SEARCH_GRID, L.SF_SPH_SEARCH_ALL])
def test_cubic_model_lattice_closed_forms(search):
    """The closed forms that pin the oracle's pressure force and du/dt, on
    the device: a 24^3 lattice gas at density 3.7 with support radius 4
    spacings.  u linear in x gives a = -(gamma - 1) grad u; the homologous
    flow v = alpha (x - c) gives du/dt = -3 alpha P/rho and no force.  Bound
    2e-3 on the interior particles (lattice moment error 1.3e-4, FP32 pair
    sums)."""
    pos, mas, h, inner = sph_lattice()
    N = len(mas)
    gu, alpha = 0.2, -0.3
    cases = [(0.05 + gu * (pos[:, 0] - 0.5), np.zeros((N, 3))), (np.full(N, 0.05), alpha * (pos - 0.5))]
    with cubic_sim(N, mem_mode=L.SF_MEM_MIRROR, sph_search=search) as s:
        for k, (u, vel) in enumerate(cases):
            s.pos[:] = pos
            s.vel[:] = vel
            s.mas[:] = mas
            s.rad[:] = h
            s.acc[:] = 0
            s.set_u(u.astype(np.float32))
            s.solve_sph()
            rho = s.rho.astype(np.float64)
            assert np.allclose(rho[inner], SPH_RHO0, rtol=2e-3)
            a = s.acc.astype(np.float64)
            du = s.get_dudt().astype(np.float64)
            if k == 0:
                a_exp = -(SPH_GAMMA - 1.0) * gu
                assert np.abs(a[inner, 0] / a_exp - 1.0).max() <= 2e-3
                assert np.abs(a[inner, 1:]).max() <= 1e-4 * abs(a_exp)
            else:
                du_exp = -3.0 * alpha * s.prs.astype(np.float64) / rho
                assert np.abs(du[inner] / du_exp[inner] - 1.0).max() <= 2e-3
                assert np.abs(a[inner]).max() <= 1e-4 * np.abs(du_exp[inner]).max()


@pytest.mark.parametrize("N", [1, 2, 37])
@pytest.mark.parametrize("search", [L.SF_SPH_SEARCH_GRID, L.SF_SPH_SEARCH_ALL])
def test_cubic_model_tiny_systems(oracle, N, search):
    """One particle (self-density only), a pair, and a handful: both searches
    against the FP64 oracle."""
    pos, vel, mas, rad, u = cubic_state(N, 40 + N, 0.8)
    check_cubic_against_oracle(oracle, pos, vel, mas, rad, u, search)
