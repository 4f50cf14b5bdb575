"""vis_draw on the device (src/visualization.c:253-736) against the
reference's own output (fixture F6: 4096 particles x 128 samples at
192 x 128 pixels), and properties of a large frame.

cosf, sinf and logf in the splat are the device library's; on F6 every
sample lands where the reference's does, so the comparison is bit for bit.
"""
import numpy as np
import pytest

from conftest import golden, ulp_diff
from starflood_amd import _lib as L
from starflood_amd.sim import Simulation
from starflood_amd.vis import Visualization
from test_vis_host import decode_png, pfm_bytes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mem_mode", [L.SF_MEM_MIRROR, L.SF_MEM_RESIDENT])
def test_vis_draw_and_files_bit_identical_to_reference(tmp_path, mem_mode):
    g = golden("f6_vis_n4096.npz")
    sx, sy = int(g["sizex"]), int(g["sizey"])
    with Simulation(4096, enable_tree=0, logging=0, epsilon=float(g["eps"]), dt=float(g["dt"]),
                    mem_mode=mem_mode) as s:
        s.mem[:] = g["mem"]
        s.to_device()
        s.step_number = int(g["step"])
        with Visualization(sx, sy) as v:
            v.draw(s)
            assert np.array_equal(v.atomic, g["atomic"])
            assert np.array_equal(v.render.view(np.uint32), g["render"].view(np.uint32))
            v.save(tmp_path / "f.ppm", L.IMAGE_FILETYPE_PPM)
            v.save(tmp_path / "f.pfm", L.IMAGE_FILETYPE_PFM)
            v.save(tmp_path / "f.png", L.IMAGE_FILETYPE_PNG)
    assert np.array_equal(np.frombuffer((tmp_path / "f.ppm").read_bytes(), np.uint8), g["ppm"])
    assert (tmp_path / "f.pfm").read_bytes() == pfm_bytes(g["render"])
    w, h, pixels = decode_png((tmp_path / "f.png").read_bytes())
    assert (w, h) == (sx, sy) and np.array_equal(pixels.ravel(), g["png_pixels"])


def test_large_frame_is_deterministic_and_tone_mapped_from_its_counts():
    """Plummer sphere N = 65536 at 1080 x 1080 (the reference program's frame):
    two draws are identical; every sample lands once (alpha sums to at most
    128 N); colour counts stay within 100 x alpha; the render buffer is
    tanh(0.01 / 128 * 2^-1 * counts) to 1 ulp (double tanh on the device and
    in numpy, rounded to float), alpha 1."""
    N = 65536
    with Simulation(N, enable_tree=0, logging=0, epsilon=1e-2, dt=1e-3, precision=L.SF_PRECISION_FAST) as s:
        s.ic_plummer(seed=3, first=0, count=N)
        s.solve()
        s.step_number = 123
        with Visualization(1080, 1080) as v:
            v.draw(s)
            a1, r1 = v.atomic.copy(), v.render.copy()
            v.draw(s)
            assert np.array_equal(v.atomic, a1) and np.array_equal(v.render.view(np.uint32), r1.view(np.uint32))
    alpha = a1[..., 3].astype(np.int64)
    assert 0.5 * 128 * N < alpha.sum() <= 128 * N
    assert (a1[..., :3] >= 0).all() and (a1[..., :3].astype(np.int64) <= 100 * alpha[..., None]).all()
    scale = 0.01 / 128 * 2.0 ** -1
    expect = np.tanh(scale * a1[..., :3].astype(np.float64)).astype(np.float32)
    assert ulp_diff(r1[..., :3], expect).max() <= 1
    assert (r1[..., 3] == 1.0).all()


def test_degenerate_frames_and_positions():
    """A 0 x 0 and a 1 x 1 frame; a particle at 3e38 (the mean radius, and so
    the zoom, follows it); and NaN or infinite coordinates, which make the
    reference's mean radius, hence its whole camera matrix, NaN, so that no
    sample lands (a NaN pixel coordinate converts to INT_MIN on its host)."""
    N = 256
    with Simulation(N, enable_tree=0, logging=0, epsilon=1e-2, dt=1e-3) as s:
        s.ic_uniform_ball(seed=9, radius=1.0, total_mass=1.0, v_hubble=0.0)
        s.solve()
        s.pos[2] = (3e38, 3e38, 3e38)
        with Visualization(0, 0) as v:
            v.draw(s)
        with Visualization(1, 1) as v:
            v.draw(s)
            a = v.atomic.copy()
            assert 0 <= a[0, 0, 3] <= 128 * N and (a[..., :3] <= 100 * a[..., 3:4]).all()
        with Visualization(64, 48) as v:
            v.draw(s)
            assert 0 < v.atomic[..., 3].sum() <= 128 * (N - 1)
            for bad in ((np.nan, 0.0, 0.0), (np.inf, -np.inf, 0.0)):
                s.pos[0] = bad
                v.draw(s)
                assert v.atomic.sum() == 0
                assert (v.render[..., :3] == 0.0).all() and (v.render[..., 3] == 1.0).all()
