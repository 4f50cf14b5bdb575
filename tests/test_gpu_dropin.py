"""The reference-style C driver (tests/c/dropin_main.c, built by
`make -C tests/c`) run as a program on the device: the snapshots and
statistics.csv it writes at N = 256 over 10 steps equal, bit for bit, the
files the reference program itself writes (golden fixture F7), with the
plain and the SOLVER_VIS library, every step and every 4th step."""
import hashlib
import subprocess

import pytest

from conftest import REPO, golden

pytestmark = pytest.mark.gpu

BIN = REPO / "tests" / "c" / "bin"


@pytest.mark.parametrize("exe,key", [("dropin_main", "every"), ("dropin_main_solvervis", "every"),
                                     ("dropin_main_oi4", "oi4")])
def test_c_driver_writes_the_reference_programs_files(tmp_path, exe, key):
    g = golden("f7_run_n256.npz")
    path = BIN / exe
    if not path.exists():
        pytest.fail(f"{path} is not built (make -C tests/c)")
    (tmp_path / "out" / "sim").mkdir(parents=True)
    print(f"running {exe} N=256 steps=10", flush=True)
    p = subprocess.run([str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    files = sorted((tmp_path / "out" / "sim").glob("*.raw"))
    assert [f.name for f in files] == list(g[f"{key}_names"])
    assert [hashlib.sha256(f.read_bytes()).hexdigest() for f in files] == list(g[f"{key}_sha256"])
    assert (tmp_path / "out" / "statistics.csv").read_text() == str(g[f"{key}_statistics"])
