"""SF_TREE_LBVH Barnes-Hut on the device: bit-exact Morton codes and stable
sort against the oracle, tree invariants, the theta -> 0 limit equal to the
direct sum, and the force error against the FP64 direct sum."""
import numpy as np
import pytest

from conftest import bits
from starflood_amd import _lib as L
from starflood_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def lbvh_sim(N, **kw):
    args = dict(enable_tree=1, tree_mode=L.SF_TREE_LBVH, logging=0, epsilon=1e-2, dt=1e-3,
                acc_mode=L.SF_ACC_PHYSICAL, precision=L.SF_PRECISION_FAST)
    args.update(kw)
    return Simulation(N, **args)


def rel_err(acc, a64):
    return np.linalg.norm(acc - a64, axis=1) / np.linalg.norm(a64, axis=1)


@pytest.mark.parametrize("N", [1, 2, 3, 63, 64, 65, 1000, 4097])
def test_small_trees_equal_direct_sum_when_every_node_opens(oracle, N):
    rng = np.random.default_rng(N)
    with lbvh_sim(N, theta=1e-6) as s:
        s.pos[:] = rng.normal(size=(N, 3)).astype(np.float32)
        s.mas[:] = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
        s.solve()
        ex = s.lbvh_export()
        assert ex["bad"] == 0
        assert ex["cells"] == 0 and ex["parts"] == N * N      # all leaf sums
        a64, p64 = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max() + 1e-30
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale
        assert np.allclose(s.pot, p64, rtol=2e-6)
    L.set_params()


def test_morton_codes_sort_and_tree_invariants(oracle):
    N = 50000
    with lbvh_sim(N) as s:
        s.ic_plummer(seed=4, first=0, count=N)
        s.pos[:5] = s.pos[5:10]          # duplicate positions: equal codes
        s.solve()
        ex = s.lbvh_export(nodes=True)
        keys = oracle.morton63(s.pos.copy())
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ex["keys"], keys[order])
        assert np.array_equal(ex["order"], order.astype(np.uint32))
        # every internal node covers a contiguous range split exactly by its children
        lr = ex["node_lr_range"]
        def rng_of(c):
            return (c - (N - 1), 1) if c >= N - 1 else (lr[c, 2], lr[c, 3])
        assert tuple(lr[0, 2:]) == (0, N)
        f_l, n_l = np.array([rng_of(c) for c in lr[:, 0]]).T
        f_r, n_r = np.array([rng_of(c) for c in lr[:, 1]]).T
        assert np.array_equal(f_l, lr[:, 2]) and np.array_equal(f_l + n_l, f_r)
        assert np.array_equal(n_l + n_r, lr[:, 3])
        # every node id is referenced exactly once as a child
        kids = np.sort(np.concatenate([lr[:, 0], lr[:, 1]]))
        assert np.array_equal(kids, np.arange(1, 2 * N - 1))
        # monopoles: masses and centres of mass of the ranges (FP64 check)
        cm = ex["node_cm"]
        m_sorted = s.mas[order].astype(np.float64)
        p_sorted = s.pos[order].astype(np.float64)
        cum_m = np.concatenate([[0], np.cumsum(m_sorted)])
        cum_mp = np.concatenate([np.zeros((1, 3)), np.cumsum(m_sorted[:, None] * p_sorted, 0)])
        first, count = lr[:, 2], lr[:, 3]
        M = cum_m[first + count] - cum_m[first]
        com = (cum_mp[first + count] - cum_mp[first]) / M[:, None]
        assert np.allclose(cm[:N - 1, 3], M, rtol=1e-5)
        assert np.all(np.abs(cm[:N - 1, :3] - com) <= 1e-5 * (1.0 + np.abs(com)))
        assert np.array_equal(bits(cm[N - 1:]), bits(np.column_stack([s.pos[order], s.mas[order]])))
    L.set_params()


def _check_node_boxes(ex, pos_sorted, N):
    """Internal nodes of at most 64 particles: their opening-test box is the
    particles' bounding box (centre, longest side), bit for bit as FP32 min,
    max, 0.5 (lo + hi) and hi - lo; leaves: the particle, side 0; larger
    nodes: a cell that holds their particles."""
    cell, lr = ex["node_cell"], ex["node_lr_range"]
    assert np.array_equal(bits(cell[N - 1:, :3]), bits(pos_sorted)) and not cell[N - 1:, 3].any()
    small = 0
    for n in range(N - 1):
        f, c = int(lr[n, 2]), int(lr[n, 3])
        p = pos_sorted[f:f + c]
        lo, hi = p.min(0), p.max(0)
        if c <= 64:
            small += 1
            want = np.append(np.float32(0.5) * (lo + hi), (hi - lo).max()).astype(np.float32)
            assert np.array_equal(bits(cell[n]), bits(want)), (n, f, c, cell[n], want)
        else:
            half = 0.5 * np.float64(cell[n, 3]) * (1 + 1e-6) + 1e-6 * np.abs(cell[n, :3]).max()
            assert np.all(np.abs(p - cell[n, :3].astype(np.float64)) <= half), n
    assert small > 0


def test_opening_test_boxes_of_small_nodes_are_their_particles_bounding_boxes(oracle):
    """The boxes of nodes of at most 64 particles come from one wavefront per
    walk group (a sparse table of min and max); they must equal a scan over
    each node's sorted range, on a Plummer sphere with duplicated positions
    and on the degenerate distributions."""
    N = 20000
    with lbvh_sim(N) as s:
        s.ic_plummer(seed=4, first=0, count=N)
        s.pos[:5] = s.pos[5:10]
        s.solve()
        ex = s.lbvh_export(nodes=True)
        _check_node_boxes(ex, s.pos[ex["order"]], N)
    rng = np.random.default_rng(5)
    for name, pos in degenerate_cases(3000, rng):
        with lbvh_sim(3000) as s:
            s.pos[:] = pos.astype(np.float32)
            s.solve()
            ex = s.lbvh_export(nodes=True)
            _check_node_boxes(ex, s.pos[ex["order"]], 3000)
    for n in (2, 3, 64, 65, 129):
        with lbvh_sim(n) as s:
            s.ic_plummer(seed=n, first=0, count=n)
            s.solve()
            ex = s.lbvh_export(nodes=True)
            _check_node_boxes(ex, s.pos[ex["order"]], n)
    L.set_params()


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_force_error_against_fp64_direct_sum(oracle, theta):
    N = 131072
    with lbvh_sim(N, theta=theta, epsilon=5e-3, mem_mode=L.SF_MEM_RESIDENT) as s:
        s.ic_plummer(seed=21, first=0, count=N)
        s.solve()
        s.to_host()
        ex = s.lbvh_export()
        assert ex["bad"] == 0
        idx = np.arange(0, N, 32, dtype=np.uint32)
        a64, p64 = s.direct_f64(idx)
        e = rel_err(s.acc[idx], a64)
        print(f"theta {theta}: rel acc error p50 {np.percentile(e, 50):.3g} p90 {np.percentile(e, 90):.3g} "
              f"p99 {np.percentile(e, 99):.3g} max {e.max():.3g}; interactions/particle "
              f"{(ex['cells'] + ex['parts']) / N:.0f}")
        limit = {0.3: 2e-4, 0.5: 1e-3, 0.7: 3e-3}[theta]
        assert np.percentile(e, 50) <= limit
        assert np.abs(s.pot[idx] / p64 - 1).max() < 1e-2
    L.set_params()


@pytest.mark.parametrize("quad,alpha", [(0, 0.0), (1, 0.0), (0, 1e-3), (1, 2.5e-3)])
def test_walk_is_deterministic(quad, alpha):
    """Wavefronts take their chunks of walk groups from a shared counter, so
    which wavefront walks a group changes from run to run; a group's sums do
    not depend on it, and the forces after sim_solv and two KDK steps are
    bit-identical across runs, for monopole and quadrupole cells, with and
    without the relative opening test."""
    N = 200000
    outs = []
    for _ in range(2):
        with lbvh_sim(N, quadrupole=quad, mac_alpha=alpha, theta=1.15 if alpha else 0.5,
                      leaf_size=6 if quad else 3) as s:
            s.ic_plummer(seed=9, first=0, count=N)
            s.solve()
            for n in range(2):
                s.step_number = n
                s.step()
            assert s.lbvh_export()["bad"] == 0
            outs.append((s.acc.copy(), s.pot.copy(), s.pos.copy()))
    for k in range(3):
        assert np.array_equal(bits(outs[0][k]), bits(outs[1][k]))
    L.set_params()


def test_lbvh_kdk_steps_and_reference_acc_mode(oracle):
    N = 8192
    with lbvh_sim(N, theta=0.4, acc_mode=L.SF_ACC_REFERENCE, precision=L.SF_PRECISION_EXACT) as s:
        s.ic_plummer(seed=1, first=0, count=N)
        s.solve()
        a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=0)
        assert np.median(rel_err(s.acc, a64)) < 1e-3
        for n in range(3):
            s.step()
        assert np.isfinite(s.pos).all() and np.isfinite(s.acc).all()
    L.set_params()


def degenerate_cases(N, rng):
    """(name, positions) pairs that stress the Morton build: coincident
    particles (equal keys, split by index), a line (zero extent in y and z),
    two clusters 1e4 apart, and a grid (many equal-prefix ties)."""
    dup = rng.normal(size=(N, 3))
    dup[: N // 2] = dup[0]                             # half the particles at one point
    line = np.zeros((N, 3))
    line[:, 0] = rng.uniform(-1, 1, N)
    far = rng.normal(scale=0.1, size=(N, 3))
    far[N // 2:, 0] += 1.0e4
    g = int(round(N ** (1 / 3)))
    ax = np.arange(g) / g
    grid = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    grid = np.concatenate([grid, rng.uniform(size=(N - len(grid), 3))])[:N]
    return [("duplicates", dup), ("line", line), ("two_clusters", far), ("grid", grid)]


@pytest.mark.parametrize("case", range(4))
def test_degenerate_distributions_equal_direct_sum_when_every_node_opens(oracle, case):
    N = 3000
    rng = np.random.default_rng(77 + case)
    name, pos = degenerate_cases(N, rng)[case]
    # theta 1e-12: small nodes are sized by their particles' box, and on the
    # line two neighbours 1e-7 apart make a node that theta 1e-6 would let
    # groups 0.1 away accept (harmlessly: a 1e-12 relative monopole error)
    with lbvh_sim(N, theta=1e-12) as s:
        s.pos[:] = pos.astype(np.float32)
        s.mas[:] = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
        s.solve()
        ex = s.lbvh_export()
        assert ex["bad"] == 0, name
        if name in ("line", "grid"):
            # elsewhere nodes of coincident particles have a box of size zero, and the
            # 0.005-wide clusters 1e4 apart tiny ones, which distant groups rightly accept
            # even at theta = 1e-12: the force check below still holds
            assert ex["cells"] == 0 and ex["parts"] == N * N, name
        keys = oracle.morton63(s.pos.copy())
        assert np.array_equal(ex["keys"], np.sort(keys)), name
        assert np.array_equal(ex["order"], np.argsort(keys, kind="stable").astype(np.uint32)), name
        a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max() + 1e-30
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale, name
    L.set_params()


@pytest.mark.parametrize("case", range(4))
def test_degenerate_distributions_at_theta_half(oracle, case):
    """The accepted-cell path on the same distributions: finite forces, no
    malformed-tree report, and the monopole truncation error of a normal
    walk.  The error is measured against the RMS force, since on a line
    the net force nearly cancels for many particles.  On the line a node's
    particle box is a segment, whose monopole error at the same side /
    distance is larger than a filled cube's (its traceless quadrupole does
    not vanish): median |da| / rms|a| 1.5e-2 there (6.5e-3 when nodes were
    sized by their Morton cube), below 1.5e-3 elsewhere."""
    N = 20000
    rng = np.random.default_rng(91 + case)
    name, pos = degenerate_cases(N, rng)[case]
    with lbvh_sim(N, theta=0.5, leaf_size=8) as s:
        s.pos[:] = pos.astype(np.float32)
        s.mas[:] = np.full(N, 1.0 / N, np.float32)
        s.solve()
        assert s.lbvh_export()["bad"] == 0, name
        assert np.isfinite(s.acc).all(), name
        idx = np.arange(0, N, 7)
        a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1, idx=idx.astype(np.uint32))
        d = np.linalg.norm(s.acc[idx] - a64, axis=1)
        rms = np.sqrt((np.linalg.norm(a64, axis=1) ** 2).mean())
        assert np.median(d) / rms < (2e-2 if name == "line" else 1e-2), (name, np.median(d) / rms)
        if name != "line":
            assert np.median(rel_err(s.acc[idx], a64)) < 5e-3, name
    L.set_params()


def _relative_run(N, theta, alpha, steps, acc_mode=L.SF_ACC_PHYSICAL, quadrupole=0):
    """sim_solv then `steps` KDK steps on a Plummer sphere (seed 21); the
    final acc, the FP64 reference on every 32nd particle and the walk's
    interaction count per particle."""
    with lbvh_sim(N, theta=theta, mac_alpha=alpha, epsilon=5e-3, acc_mode=acc_mode, quadrupole=quadrupole,
                  leaf_size=4 if quadrupole else 3, mem_mode=L.SF_MEM_RESIDENT) as s:
        s.ic_plummer(seed=21, first=0, count=N)
        s.solve()
        for n in range(steps):
            s.step_number = n
            s.step()
        s.to_host()
        ex = s.lbvh_export()
        assert ex["bad"] == 0
        idx = np.arange(0, N, 32, dtype=np.uint32)
        a64, _ = s.direct_f64(idx)
        return s.acc[idx].copy(), a64, (ex["cells"] + ex["parts"]) / N, s.mas[idx].copy()


def test_relative_opening_test_bounds_the_error_relative_to_each_acceleration(oracle):
    """mac_alpha > 0 also requires G M l^2 <= alpha |a|_min d^4 of every cell,
    with |a|_min from the group's previous accelerations.  At N = 131072:
    - it only removes cells, so at the same theta it evaluates more
      interactions and every error percentile is lower;
    - at theta 1 and alpha 2e-4 the 99th percentile stays within 2e-3, and the
      errors spread less (99th percentile over median) than with the theta
      test alone at theta 0.5;
    - alpha 1e30 lets every cell pass, and the forces equal the theta walk's
      bit for bit;
    - it holds after two KDK steps and in the reference's acc = m a mode."""
    N = 131072
    acc_t, a64t, n_t, _ = _relative_run(N, 0.7, 0.0, 2)
    acc_r, a64r, n_r, _ = _relative_run(N, 0.7, 2e-4, 2)
    e_t, e_r = rel_err(acc_t, a64t), rel_err(acc_r, a64r)
    print(f"theta 0.7: {n_t:.0f} interactions/particle, p50 {np.percentile(e_t, 50):.3g}, p99 "
          f"{np.percentile(e_t, 99):.3g}; with alpha 2e-4: {n_r:.0f}, p50 {np.percentile(e_r, 50):.3g}, "
          f"p99 {np.percentile(e_r, 99):.3g}")
    assert n_r > n_t
    for q in (50, 90, 99):
        assert np.percentile(e_r, q) < np.percentile(e_t, q)
    acc_1, a64, n_1, _ = _relative_run(N, 1.0, 2e-4, 2)
    acc_5, a645, n_5, _ = _relative_run(N, 0.5, 0.0, 2)
    e_1, e_5 = rel_err(acc_1, a64), rel_err(acc_5, a645)
    print(f"theta 1, alpha 2e-4: {n_1:.0f} interactions/particle, p50 {np.percentile(e_1, 50):.3g}, p99 "
          f"{np.percentile(e_1, 99):.3g}; theta 0.5 alone: {n_5:.0f}, p50 {np.percentile(e_5, 50):.3g}, p99 "
          f"{np.percentile(e_5, 99):.3g}")
    assert np.percentile(e_1, 99) <= 2e-3
    assert np.percentile(e_1, 99) / np.percentile(e_1, 50) < np.percentile(e_5, 99) / np.percentile(e_5, 50)
    acc_h, _, _, _ = _relative_run(N, 0.7, 1e30, 2)
    assert np.array_equal(acc_h.view(np.uint32), acc_t.view(np.uint32))
    acc_m, a64m, _, m = _relative_run(N, 1.0, 2e-4, 2, acc_mode=L.SF_ACC_REFERENCE)
    e_m = rel_err(acc_m, a64m)
    assert np.percentile(e_m, 99) <= 2e-3
    L.set_params()


def test_relative_opening_test_with_no_cells_equals_the_direct_sum(oracle):
    """alpha 1e-30 rejects every cell: only leaf particles are summed, and the
    forces equal the FP64 direct sum to FP32 rounding (solve and one step)."""
    N = 4096
    rng = np.random.default_rng(5)
    with lbvh_sim(N, theta=0.5, mac_alpha=1e-30) as s:
        s.pos[:] = rng.normal(size=(N, 3)).astype(np.float32)
        s.mas[:] = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
        s.solve()
        s.step_number = 0
        s.step()
        ex = s.lbvh_export()
        assert ex["bad"] == 0 and ex["cells"] == 0 and ex["parts"] == N * N
        a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max()
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale
    L.set_params()


def test_quadrupole_walk_sums_cells_that_meet_the_monopole_bound_as_monopoles(oracle):
    """Under the relative opening test a quadrupole walk sums a cell that
    also meets G M l^2 <= 0.2 alpha |a| d^4 as a monopole, through the leaf
    particles' queue.  With alpha 1e30 every accepted cell is such a
    monopole: the forces equal the monopole walk's at the same theta (the
    same cells, summed in another order) to FP32 rounding (median 1e-6, 99th
    percentile 1e-5 of |a|), and the cells count as cell interactions.  At
    alpha 2.5e-3 and theta 1.15 the 99th percentile stays within 1.5e-3,
    far below the theta test's alone at theta 1.15 with either expansion."""
    N = 131072
    def run(quad, alpha, theta):
        with lbvh_sim(N, theta=theta, mac_alpha=alpha, epsilon=5e-3, quadrupole=quad, leaf_size=4,
                      mem_mode=L.SF_MEM_RESIDENT) as s:
            s.ic_plummer(seed=9, first=0, count=N)
            s.solve()
            s.to_host()
            ex = s.lbvh_export()
            assert ex["bad"] == 0
            idx = np.arange(0, N, 32, dtype=np.uint32)
            a64, _ = s.direct_f64(idx)
            return s.acc.copy(), idx, a64, ex
    acc_m, idx, a64, ex_m = run(0, 0.0, 0.7)
    acc_q, _, _, ex_q = run(1, 1e30, 0.7)
    assert (ex_q["cells"], ex_q["parts"]) == (ex_m["cells"], ex_m["parts"])
    d = np.linalg.norm(acc_q - acc_m, axis=1) / np.linalg.norm(acc_m, axis=1)
    print(f"alpha 1e30 against the monopole walk: median {np.median(d):.3g}, p99 {np.percentile(d, 99):.3g}")
    assert np.median(d) <= 1e-6 and np.percentile(d, 99) <= 1e-5
    e_quad = rel_err(run(1, 0.0, 1.15)[0][idx], a64)
    e_mix = rel_err(run(1, 2.5e-3, 1.15)[0][idx], a64)
    e_mono = rel_err(run(0, 0.0, 1.15)[0][idx], a64)
    print(f"p99 at theta 1.15: quadrupole {np.percentile(e_quad, 99):.3g}, relative alpha 2.5e-3 "
          f"{np.percentile(e_mix, 99):.3g}, monopole {np.percentile(e_mono, 99):.3g}")
    assert np.percentile(e_mix, 99) < min(np.percentile(e_mono, 99), np.percentile(e_quad, 99))
    assert np.percentile(e_mix, 99) <= 1.5e-3
    L.set_params()


@pytest.mark.parametrize("quad", [0, 1])
@pytest.mark.parametrize("N", [1, 2, 3, 65, 1000])
def test_relative_opening_test_on_small_and_degenerate_systems(oracle, N, quad):
    """The relative opening test on tiny trees (solve, then a step, whose walk
    takes the previous accelerations): no guard trips, and with theta 1e-12
    every node opens, so the forces equal the FP64 direct sum.  A system of
    coincident particles has |a| = 0 everywhere: the walk falls back to the
    theta test and stays finite."""
    rng = np.random.default_rng(100 + N)
    with lbvh_sim(N, theta=1e-12, mac_alpha=2.5e-3, quadrupole=quad, leaf_size=4) as s:
        s.pos[:] = rng.normal(size=(N, 3)).astype(np.float32)
        s.mas[:] = rng.uniform(0.5, 1.5, N).astype(np.float32) / N
        s.solve()
        s.step_number = 0
        s.step()
        ex = s.lbvh_export()
        assert ex["bad"] == 0 and ex["cells"] == 0 and ex["parts"] == N * N
        a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1)
        scale = np.linalg.norm(a64, axis=1).max() + 1e-30
        assert np.abs(s.acc - a64).max() <= 2e-6 * scale
    with lbvh_sim(N, theta=1.15, mac_alpha=2.5e-3, quadrupole=quad) as s:
        s.pos[:] = 0.25
        s.mas[:] = 1.0 / N
        s.solve()
        s.step_number = 0
        s.step()
        assert s.lbvh_export()["bad"] == 0
        assert np.all(np.isfinite(s.acc)) and np.all(np.abs(s.acc) <= 1e-6)
    L.set_params()


def test_mirror_and_resident_modes_give_identical_forces():
    N = 30000
    outs = []
    for mode in (L.SF_MEM_MIRROR, L.SF_MEM_RESIDENT):
        with lbvh_sim(N, mem_mode=mode) as s:
            s.ic_plummer(seed=13, first=0, count=N)
            s.solve()
            s.step()
            if mode == L.SF_MEM_RESIDENT:
                s.to_host()
            outs.append(s.mem.copy())
    assert np.array_equal(bits(outs[0]), bits(outs[1]))
    L.set_params()


def quad_cell_field(src_pos, src_m, x, eps):
    """FP64 monopole + traceless-quadrupole field and potential of the
    point masses src at points x (the expansion about their centre of mass,
    Plummer-softened like the device: r^2 = |e|^2 + eps^2)."""
    M = src_m.sum()
    c = (src_m[:, None] * src_pos).sum(0) / M
    d = src_pos - c
    Q = (src_m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
    Qt = 3.0 * Q - np.trace(Q) * np.eye(3)
    e = c[None, :] - x
    r2 = (e * e).sum(1) + eps * eps
    inv = 1.0 / np.sqrt(r2)
    qe = e @ Qt
    eqe = (e * qe).sum(1)
    acc = (M * inv ** 3 + 2.5 * eqe * inv ** 7)[:, None] * e - (inv ** 5)[:, None] * qe
    pot = M * inv + 0.5 * eqe * inv ** 5
    return acc, -pot


def test_quadrupole_cell_matches_fp64_expansion(oracle):
    """A compact 64-particle cluster seen from a far 64-particle group is
    accepted as one cell: the device field on the group equals the FP64
    monopole + quadrupole expansion of the cluster plus the group's own
    direct sum."""
    rng = np.random.default_rng(5)
    a = rng.uniform(0.0, 0.1, size=(64, 3))
    b = np.array([10.0, 3.0, 2.0]) + rng.uniform(0.0, 0.01, size=(64, 3))
    pos = np.concatenate([a, b]).astype(np.float32)
    mas = rng.uniform(0.5, 1.5, 128).astype(np.float32) / 128
    eps = 1e-2
    with lbvh_sim(128, theta=0.5, quadrupole=1, epsilon=eps) as s:
        s.pos[:] = pos
        s.mas[:] = mas
        s.solve()
        assert s.lbvh_export()["cells"] >= 64                # the cluster was accepted for the group
        p64, m64 = pos.astype(np.float64), mas.astype(np.float64)
        qa, qp = quad_cell_field(p64[:64], m64[:64], p64[64:], eps)
        da, dp = oracle.pp_f64(pos[64:].copy(), mas[64:].copy(), eps, acc_phys=1)
        ref = qa + da
        got = s.acc[64:].astype(np.float64)
        assert np.abs(got - ref).max() <= 1e-5 * np.linalg.norm(ref, axis=1).max()
        assert np.allclose(s.pot[64:], qp + dp, rtol=1e-5)
    L.set_params()


def test_quadrupole_error_at_two_thetas(oracle):
    """At theta 0.6 and 0.3 the quadrupole walk's median error is at least 4x
    below the monopole walk's (measured 7.5x on this Plummer sphere), and both
    fall at least 3x when theta halves (measured ~10x each: with this
    cell-size criterion the error falls faster than the bare expansion order
    suggests, for either expansion).  The formula itself is pinned by
    test_quadrupole_cell_matches_fp64_expansion."""
    N = 30000
    med = {}
    for quad in (0, 1):
        for theta in (0.6, 0.3):
            with lbvh_sim(N, theta=theta, quadrupole=quad, leaf_size=4) as s:
                s.ic_plummer(seed=17, first=0, count=N)
                s.solve()
                idx = np.arange(0, N, 10).astype(np.uint32)
                a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1, idx=idx)
                med[quad, theta] = float(np.median(rel_err(s.acc[idx], a64)))
    for theta in (0.6, 0.3):
        assert med[1, theta] < med[0, theta] / 4, med
    for quad in (0, 1):
        assert med[quad, 0.6] / med[quad, 0.3] > 3.0, med
    L.set_params()


def test_quadrupole_walk_is_more_accurate_at_the_same_theta(oracle):
    N = 60000
    errs = {}
    for quad in (0, 1):
        with lbvh_sim(N, theta=0.5, quadrupole=quad, leaf_size=4) as s:
            s.ic_plummer(seed=23, first=0, count=N)
            s.solve()
            assert s.lbvh_export()["bad"] == 0
            idx = np.arange(0, N, 15).astype(np.uint32)
            a64, _ = oracle.pp_f64(s.pos.copy(), s.mas.copy(), 1e-2, acc_phys=1, idx=idx)
            e = rel_err(s.acc[idx], a64)
            errs[quad] = (np.median(e), np.percentile(e, 99))
    assert errs[1][0] < errs[0][0] / 3, errs
    assert errs[1][1] < errs[0][1] / 2, errs
    L.set_params()
