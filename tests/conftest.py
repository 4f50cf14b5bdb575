import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
GOLD = REPO / "tests" / "golden"
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def golden(name: str) -> dict:
    """Load a golden fixture (plain arrays; no pickles)."""
    with np.load(GOLD / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a.view(np.uint64)


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in FP32 units in the last place (sign-magnitude aware)."""
    def key(x):
        i = np.ascontiguousarray(x, np.float32).view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    return np.abs(key(a) - key(b))


# Cubic-lattice gas for the SPH closed forms: 24^3 particles filling the unit
# cube at density RHO0 (far from 1, so P/rho and P/rho^2 differ 3.7-fold),
# support radius 4 spacings.  Particles at least 2h from every face have a
# full neighbourhood whose members also have full neighbourhoods, so there
# rho is uniform and the SPH sums reduce to their lattice moments.  The
# lattice moment sum_j V (x_j - x_i)_x dW/dx, which is 1 in the continuum,
# is 1 - 1.3e-4 at 4 spacings (measured; 1.7e-2 at 2.5, 3e-3 at 3.5).
SPH_LAT_N, SPH_LAT_H, SPH_RHO0, SPH_GAMMA = 24, 4.0 / 24, 3.7, 5.0 / 3.0


def sph_lattice():
    n = SPH_LAT_N
    g = (np.arange(n) + 0.5) / n
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    mas = np.full(len(pos), SPH_RHO0 / len(pos))
    h = np.full(len(pos), SPH_LAT_H)
    inner = np.all(np.abs(pos - 0.5) < 0.5 - 2.0 * SPH_LAT_H, axis=1)
    return pos, mas, h, inner
