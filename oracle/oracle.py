"""ctypes access to the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product library never
loads any of it.

  lib()        liboracle.so, the C restatement (sf_oracle.c)
  RefSim       a simulation driven through a build of the reference's own
               sources (oracle/_ref/libref_<variant>.so), with the
               reference's sim_t layout for that variant
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
import time
from contextlib import contextmanager
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
REF_DIR = ORACLE_DIR / "_ref"
REFERENCE_SRC = Path("/root/reference/src")

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_u = C.POINTER(C.c_uint)


class OrcNode(C.Structure):
    _fields_ = [("parent", C.c_uint32), ("bodies", C.c_uint32), ("branch", C.c_uint32 * 8),
                ("bounds", C.c_float * 6), ("params", C.c_float * 4)]


class OrcSim(C.Structure):
    _fields_ = [("N", C.c_uint), ("dof", C.c_uint), ("step_number", C.c_uint), ("mem", _f),
                ("eps", C.c_double), ("G", C.c_double), ("dt", C.c_double), ("theta", C.c_double),
                ("enable_tree", C.c_int), ("enable_sph", C.c_int), ("kahan_solver", C.c_int),
                ("kahan_energy", C.c_int), ("acc_phys", C.c_int), ("tree_nodes_max", C.c_uint),
                ("tree_depth_max", C.c_int), ("node", C.POINTER(OrcNode)), ("num_nodes", C.c_uint),
                ("tree_status", C.c_int), ("T", C.c_double), ("V", C.c_double)]


def build(with_reference: bool | None = None) -> None:
    """Compile liboracle.so and, when the reference sources are present,
    the reference variants and the reference program's run configurations
    into _ref/."""
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "-B", "all"], check=True)
    if with_reference is None:
        with_reference = REFERENCE_SRC.is_dir()
    if with_reference:
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "ref", "ref_main", f"REF_SRC={REFERENCE_SRC}"],
                       check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = ORACLE_DIR / "liboracle.so"
        if not path.exists():
            build(with_reference=False)
        L = C.CDLL(str(path))
        L.orc_pcg4d.argtypes = [_u]
        L.orc_pp_ref.argtypes = [C.c_uint, _f, _f, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _f, _f]
        L.orc_sph_ref.argtypes = [C.c_uint, _f, _f, _f, C.c_double, C.c_int, _f, _f, _f]
        L.orc_tree_bounds_ref.argtypes = [C.c_uint, _f, C.c_uint, _f]
        L.orc_tree_build_ref.argtypes = [C.c_uint, _f, _f, _f, C.c_uint, C.c_int, C.POINTER(OrcNode), _u]
        L.orc_tree_walk_ref.argtypes = [C.c_uint, _f, _f, C.POINTER(OrcNode), C.c_uint, C.c_uint, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _f, _f]
        L.orc_sim_solv.argtypes = [C.POINTER(OrcSim)]
        L.orc_sim_step.argtypes = [C.POINTER(OrcSim)]
        L.orc_pp_f64.argtypes = [C.c_uint, _f, _f, C.c_double, C.c_double, C.c_int, _u, C.c_uint, _d, _d]
        L.orc_sph_cubic_f64.argtypes = [C.c_uint, _d, _d, _d, _d, _d, C.c_double, _d, _d, _d, _d]
        L.orc_sph_cubic_rows_f64.argtypes = [C.c_uint, _d, _d, _d, _d, _d, C.c_double, C.c_uint, _u, _d, _d, _d, _d]
        L.orc_cubic_w.argtypes = [C.c_double, C.c_double]
        L.orc_cubic_w.restype = C.c_double
        L.orc_cubic_dw.argtypes = [C.c_double, C.c_double]
        L.orc_cubic_dw.restype = C.c_double
        L.orc_force_f64.argtypes = [C.c_uint, _d, _d, C.c_double, C.c_double, C.c_int, _d]
        L.orc_force_vjp_f64.argtypes = [C.c_uint, _d, _d, C.c_double, C.c_double, C.c_int, _d, _d]
        L.orc_kdk_f64.argtypes = [C.c_uint, _d, _d, _d, C.c_double, C.c_double, C.c_double, C.c_int, C.c_uint]
        L.orc_kdk_adjoint_f64.argtypes = [C.c_uint, _d, _d, _d, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_uint, _d, _d, _d]
        L.orc_kdk_adjoint_f64.restype = C.c_double
        L.orc_morton63.argtypes = [C.c_uint, _f, C.POINTER(C.c_ulonglong)]
        L.orc_now.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f)


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_d)


def up(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u)


def pcg4d(s) -> np.ndarray:
    a = np.ascontiguousarray(s, dtype=np.uint32).copy()
    lib().orc_pcg4d(up(a))
    return a


def pp_ref(pos, mas, eps, G=1.0, kahan_acc=1, kahan_pot=1, acc_phys=0):
    pos = np.ascontiguousarray(pos, np.float32)
    mas = np.ascontiguousarray(mas, np.float32)
    N = len(mas)
    acc = np.zeros((N, 3), np.float32)
    pot = np.zeros(N, np.float32)
    lib().orc_pp_ref(N, fp(pos), fp(mas), eps, G, kahan_acc, kahan_pot, acc_phys, fp(acc), fp(pot))
    return acc, pot


def pp_f64(pos, mas, eps, G=1.0, acc_phys=0, idx=None):
    pos = np.ascontiguousarray(pos, np.float32)
    mas = np.ascontiguousarray(mas, np.float32)
    N = len(mas)
    idx = np.arange(N, dtype=np.uint32) if idx is None else np.ascontiguousarray(idx, np.uint32)
    acc = np.zeros((len(idx), 3), np.float64)
    pot = np.zeros(len(idx), np.float64)
    lib().orc_pp_f64(N, fp(pos), fp(mas), eps, G, acc_phys, up(idx), len(idx), dp(acc), dp(pot))
    return acc, pot


def morton63(pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    N = len(pos)
    key = np.zeros(N, np.uint64)
    lib().orc_morton63(N, fp(pos), key.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return key


def sph_ref(pos, mas, rad, acc_in, eps=1e-2, kahan=1):
    pos = np.ascontiguousarray(pos, np.float32)
    mas = np.ascontiguousarray(mas, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    N = len(mas)
    acc = np.ascontiguousarray(acc_in, np.float32).copy()
    rho = np.zeros(N, np.float32)
    prs = np.zeros(N, np.float32)
    lib().orc_sph_ref(N, fp(pos), fp(mas), fp(rad), eps, kahan, fp(rho), fp(prs), fp(acc))
    return rho, prs, acc


def sph_cubic_f64(pos, vel, mas, h, u, gamma):
    """Cubic-spline ideal-gas SPH in FP64: rho, P, SPH acceleration (physical
    units) and du/dt."""
    d = lambda a: np.ascontiguousarray(a, np.float64).ravel()
    pos, vel, mas, h, u = d(pos), d(vel), d(mas), d(h), d(u)
    N = len(mas)
    rho, prs, dudt = np.zeros(N), np.zeros(N), np.zeros(N)
    acc = np.zeros(3 * N)
    lib().orc_sph_cubic_f64(N, dp(pos), dp(vel), dp(mas), dp(h), dp(u), gamma, dp(rho), dp(prs), dp(acc), dp(dudt))
    return rho, prs, acc.reshape(N, 3), dudt


def sph_cubic_rows_f64(pos, vel, mas, h, u, gamma, rows):
    """sph_cubic_f64 with the force rows limited to `rows`: rho and P of all
    particles, acceleration and du/dt of the listed ones."""
    d = lambda a: np.ascontiguousarray(a, np.float64).ravel()
    pos, vel, mas, h, u = d(pos), d(vel), d(mas), d(h), d(u)
    rows = np.ascontiguousarray(rows, np.uint32)
    N = len(mas)
    rho, prs = np.zeros(N), np.zeros(N)
    acc, dudt = np.zeros(3 * len(rows)), np.zeros(len(rows))
    lib().orc_sph_cubic_rows_f64(N, dp(pos), dp(vel), dp(mas), dp(h), dp(u), gamma, len(rows),
                                 rows.ctypes.data_as(C.POINTER(C.c_uint32)), dp(rho), dp(prs), dp(acc), dp(dudt))
    return rho, prs, acc.reshape(-1, 3), dudt


def _d64(a):
    return np.ascontiguousarray(a, np.float64).ravel().copy()


def force_f64(x, m, eps, G=1.0, acc_phys=1):
    x, m = _d64(x), _d64(m)
    acc = np.zeros_like(x)
    lib().orc_force_f64(len(m), dp(x), dp(m), eps, G, acc_phys, dp(acc))
    return acc.reshape(-1, 3)


def force_vjp_f64(x, m, mu, eps, G=1.0, acc_phys=1):
    x, m, mu = _d64(x), _d64(m), _d64(mu)
    g = np.zeros_like(x)
    lib().orc_force_vjp_f64(len(m), dp(x), dp(m), eps, G, acc_phys, dp(mu), dp(g))
    return g.reshape(-1, 3)


def kdk_f64(x, v, m, eps, G, dt, acc_phys, nsteps):
    x, v, m = _d64(x), _d64(v), _d64(m)
    lib().orc_kdk_f64(len(m), dp(x), dp(v), dp(m), eps, G, dt, acc_phys, nsteps)
    return x.reshape(-1, 3), v.reshape(-1, 3)


def kdk_adjoint_f64(x0, v0, m, eps, G, dt, acc_phys, nsteps, target):
    """(L, dL/dx0, dL/dv0) for L = mean((x_N - target)^2)."""
    x0, v0, m, target = _d64(x0), _d64(v0), _d64(m), _d64(target)
    gx, gv = np.zeros_like(x0), np.zeros_like(v0)
    L = lib().orc_kdk_adjoint_f64(len(m), dp(x0), dp(v0), dp(m), eps, G, dt, acc_phys, nsteps, dp(target), dp(gx),
                                  dp(gv))
    return L, gx.reshape(-1, 3), gv.reshape(-1, 3)


def tree_ref(pos, mas, step_number, max_nodes=262144, depth_max=16, bounds=None):
    """Bounds, build and node array of the reference octree (bounds from the
    step's jittered mean radius unless given)."""
    pos = np.ascontiguousarray(pos, np.float32)
    mas = np.ascontiguousarray(mas, np.float32)
    N = len(mas)
    if bounds is None:
        bounds = np.zeros(6, np.float32)
        lib().orc_tree_bounds_ref(N, fp(pos), step_number, fp(bounds))
    else:
        bounds = np.ascontiguousarray(bounds, np.float32).copy()
    nodes = (OrcNode * max_nodes)()
    nn = C.c_uint(0)
    rc = lib().orc_tree_build_ref(N, fp(pos), fp(mas), fp(bounds), max_nodes, depth_max, nodes, C.byref(nn))
    raw = np.frombuffer(nodes, dtype=np.uint8, count=nn.value * 80).copy()
    return rc, bounds, nodes, nn.value, raw


def tree_walk_ref(pos, mas, nodes, max_nodes, eps, G=1.0, theta=0.5, acc_phys=0):
    pos = np.ascontiguousarray(pos, np.float32)
    mas = np.ascontiguousarray(mas, np.float32)
    N = len(mas)
    acc = np.zeros((N, 3), np.float32)
    pot = np.zeros(N, np.float32)
    lib().orc_tree_walk_ref(N, fp(pos), fp(mas), nodes, max_nodes, 2 * 8 * max_nodes, eps, G, theta, 1, 1,
                            acc_phys, fp(acc), fp(pot))
    return acc, pot


class OracleSim:
    """The restated reference simulation on a numpy sim.mem."""

    def __init__(self, N, *, eps, dt, G=1.0, theta=0.5, enable_tree=0, enable_sph=0, kahan_solver=1,
                 kahan_energy=1, acc_phys=0, tree_nodes_max=262144, tree_depth_max=16):
        self.N = N
        self.dof = 16 if enable_sph else 13
        self.mem = np.zeros(N * self.dof, np.float32)
        self.nodes = (OrcNode * tree_nodes_max)() if enable_tree else None
        self.s = OrcSim(N=N, dof=self.dof, step_number=0, mem=fp(self.mem), eps=eps, G=G, dt=dt, theta=theta,
                        enable_tree=enable_tree, enable_sph=enable_sph, kahan_solver=kahan_solver,
                        kahan_energy=kahan_energy, acc_phys=acc_phys, tree_nodes_max=tree_nodes_max,
                        tree_depth_max=tree_depth_max,
                        node=C.cast(self.nodes, C.POINTER(OrcNode)) if self.nodes is not None else None)

    def block(self, e, count=1):
        return self.mem[self.N * e:self.N * (e + count)]

    def solve(self):
        return lib().orc_sim_solv(C.byref(self.s))

    def step(self):
        return lib().orc_sim_step(C.byref(self.s))


# ---------------------------------------------------------------- reference
class _RefSimNoTree(C.Structure):
    """reference sim_t without ENABLE_TREE (src/simulation.h:53-67)"""
    _fields_ = [("step_number", C.c_uint), ("N", C.c_uint), ("mem", C.c_void_p)]


class _RefTree(C.Structure):
    _fields_ = [("max_nodes", C.c_uint32), ("num_nodes", C.c_uint32), ("bounds", C.c_float * 6),
                ("node", C.c_void_p)]


class _RefSimTree(C.Structure):
    _fields_ = [("step_number", C.c_uint), ("N", C.c_uint), ("mem", C.c_void_p), ("tree", _RefTree)]


def ref_vis_save(render, filetype: int, path=None):
    """The reference build's vis_save of a (h, w, 4) float render buffer:
    the 8-bit pixels it converts for PPM and PNG (PNG: the conversion only,
    the build has no PNG writer), and the file it writes when path is
    given.  Not PFM: in this (PNG-format) build the binary buffer holds
    3 bytes per pixel and the PFM path writes 12 (src/visualization.c:76-87,
    src/visualization_io.c:195-204), past the end of the allocation."""
    if filetype == 0:
        raise ValueError("the PFM path of the reference's PNG-format configuration overflows its buffer")
    lib = C.CDLL(str(REF_DIR / "libref_pp_e1e-2_dt1e-3.so"))
    render = np.ascontiguousarray(render, np.float32)
    h, w = render.shape[:2]
    vis = VisT()
    with tempfile.TemporaryDirectory() as d, _in_dir(d), capture_stdout():
        os.makedirs("out", exist_ok=True)
        assert lib.vis_init(C.byref(vis), C.c_uint(w), C.c_uint(h)) == 0
        try:
            C.memmove(vis.render_buffer, render.ctypes.data, render.nbytes)
            target = str(path) if path is not None else os.path.join(d, "unused")
            assert lib.vis_save(C.byref(vis), target.encode(), C.c_int(filetype)) == 0
            pixels = np.frombuffer((C.c_uint8 * (3 * w * h)).from_address(vis.binary_buffer), np.uint8).copy()
        finally:
            lib.vis_free(C.byref(vis))
    return pixels


def ref_available(variant: str) -> bool:
    return (REF_DIR / f"libref_{variant}.so").exists()


def run_ref_program(name: str, workdir: Path, threads: int) -> float:
    """Run the reference program build _ref/ref_main_<name> (src/main.c with
    that run configuration) in workdir, where it writes ./out/; create
    out/sim and out/vis first.  Returns the wall time in seconds."""
    exe = REF_DIR / f"ref_main_{name}"
    (Path(workdir) / "out" / "sim").mkdir(parents=True, exist_ok=True)
    (Path(workdir) / "out" / "vis").mkdir(parents=True, exist_ok=True)
    env = {"OMP_NUM_THREADS": str(threads), "PATH": os.environ.get("PATH", "/usr/bin:/bin")}
    t0 = time.perf_counter()
    subprocess.run([str(exe)], cwd=str(workdir), env=env, check=True, stdout=subprocess.DEVNULL)
    return time.perf_counter() - t0


def ref_program_available(name: str) -> bool:
    return (REF_DIR / f"ref_main_{name}").exists()


@contextmanager
def _in_dir(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


@contextmanager
def capture_stdout():
    """Capture what C code writes to fd 1; yields a list filled on exit."""
    out = []
    import sys
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            yield out
        finally:
            C.CDLL(None).fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
            tmp.seek(0)
            out.append(tmp.read().decode(errors="replace"))


class VisT(C.Structure):
    """vis_t, reference src/visualization.h:19-27."""
    _fields_ = [("sizex", C.c_uint), ("sizey", C.c_uint), ("mem", C.c_void_p), ("atomic_buffer", C.c_void_p),
                ("render_buffer", C.c_void_p), ("binary_buffer", C.c_void_p)]


class RefSim:
    """A simulation run by a build of the reference's own sources.

    The reference opens ./out/statistics.csv in sim_init, so each instance
    works in a private temporary directory."""

    def __init__(self, variant: str, N: int):
        path = REF_DIR / f"libref_{variant}.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = C.CDLL(str(path))
        self.tree = variant.startswith("tree")
        self.sph = variant.startswith("sph")
        self.dof = 16 if self.sph else 13
        self.N = N
        self.tmp = tempfile.TemporaryDirectory()
        os.makedirs(os.path.join(self.tmp.name, "out"))
        self.sim = (_RefSimTree if self.tree else _RefSimNoTree)()
        for fn in ("sim_init", "sim_free", "sim_solv", "sim_step", "initcond_generate", "solve_gravity_part_part"):
            getattr(self.lib, fn).restype = C.c_int
        with _in_dir(self.tmp.name), capture_stdout():
            rc = self.lib.sim_init(C.byref(self.sim), C.c_uint(N))
        if rc != 0:
            raise RuntimeError("reference sim_init failed")
        self.mem = np.ctypeslib.as_array((C.c_float * (N * self.dof)).from_address(self.sim.mem))

    def block(self, e, count=1):
        return self.mem[self.N * e:self.N * (e + count)]

    def _call(self, name):
        with _in_dir(self.tmp.name), capture_stdout() as out:
            rc = getattr(self.lib, name)(C.byref(self.sim))
        self.last_stdout = out[0]
        return rc

    def initcond(self):
        return self._call("initcond_generate")

    def solve(self):
        return self._call("sim_solv")

    def step(self):
        return self._call("sim_step")

    def solve_pp(self):
        return self._call("solve_gravity_part_part")

    def statistics(self) -> str:
        with open(os.path.join(self.tmp.name, "out", "statistics.csv")) as f:
            return f.read()

    def vis_frame(self, sizex: int, sizey: int, files=()):
        """vis_init, vis_draw of the current state and step number, vis_save
        to each (path, image_filetype_t) of files, vis_free: the atomic and
        render buffers (sizey, sizex, 4) and the reference's matMVP."""
        vis = VisT()
        for fn in ("vis_init", "vis_free", "vis_draw", "vis_save"):
            getattr(self.lib, fn).restype = C.c_int
        with _in_dir(self.tmp.name), capture_stdout():
            assert self.lib.vis_init(C.byref(vis), C.c_uint(sizex), C.c_uint(sizey)) == 0
            try:
                assert self.lib.vis_draw(C.byref(vis), C.byref(self.sim)) == 0
                n = 4 * sizex * sizey
                atomic = np.frombuffer((C.c_int32 * n).from_address(vis.atomic_buffer), np.int32).copy()
                render = np.frombuffer((C.c_float * n).from_address(vis.render_buffer), np.float32).copy()
                mvp = np.frombuffer((C.c_float * 16).in_dll(self.lib, "matMVP"), np.float32).copy()
                for path, ftype in files:
                    assert self.lib.vis_save(C.byref(vis), str(path).encode(), C.c_int(ftype)) == 0
                # the 8-bit pixels of the last PPM or PNG save (written even
                # where the build has no PNG writer)
                pixels = np.frombuffer((C.c_uint8 * (3 * sizex * sizey)).from_address(vis.binary_buffer),
                                       np.uint8).copy()
            finally:
                self.lib.vis_free(C.byref(vis))   # joins the image-writer thread
        return atomic.reshape(sizey, sizex, 4), render.reshape(sizey, sizex, 4), mvp, pixels

    def vis_draw_seconds(self, sizex: int, sizey: int, reps: int = 3) -> float:
        """Mean wall time of the reference's vis_draw of the current state."""
        vis = VisT()
        for fn in ("vis_init", "vis_free", "vis_draw"):
            getattr(self.lib, fn).restype = C.c_int
        with _in_dir(self.tmp.name), capture_stdout():
            assert self.lib.vis_init(C.byref(vis), C.c_uint(sizex), C.c_uint(sizey)) == 0
            try:
                assert self.lib.vis_draw(C.byref(vis), C.byref(self.sim)) == 0
                t0 = time.perf_counter()
                for _ in range(reps):
                    assert self.lib.vis_draw(C.byref(vis), C.byref(self.sim)) == 0
                dt = (time.perf_counter() - t0) / reps
            finally:
                self.lib.vis_free(C.byref(vis))
        return dt

    def tree_nodes(self, num_nodes: int) -> np.ndarray:
        raw = (C.c_uint8 * (num_nodes * 80)).from_address(self.sim.tree.node)
        return np.frombuffer(raw, dtype=np.uint8).copy()

    def close(self):
        if self.sim.mem:
            with _in_dir(self.tmp.name), capture_stdout():
                self.lib.sim_free(C.byref(self.sim))
            self.sim.mem = None
        self.tmp.cleanup()
