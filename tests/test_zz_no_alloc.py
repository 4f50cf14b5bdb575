"""sf_force_dev_noalloc through the C ABI: the allocation-free variant of
sf_force_dev (scratch in a static device array) gives the same acc and pot,
bit for bit, as sf_force_dev on the same inputs, on consecutive calls of
decreasing size and changing masses (the scratch keeps the last call's
contents), and refuses sizes beyond its maximum."""
import numpy as np
import pytest
import torch

from starflood_amd import _lib as L

pytestmark = pytest.mark.gpu


def _inputs(N, seed, equal_mass):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(N, 3)).astype(np.float32)
    m = np.full(N, 1.0 / N, np.float32) if equal_mass else rng.uniform(0.5, 1.5, N).astype(np.float32) / N
    return torch.from_numpy(x).cuda(), torch.from_numpy(m).cuda()


def _force(fn, N, x, m):
    acc = torch.full((N, 3), float("nan"), device="cuda")
    pot = torch.full((N,), float("nan"), device="cuda")
    rc = fn(N, x.data_ptr(), m.data_ptr(), acc.data_ptr(), pot.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, acc.cpu().numpy(), pot.cpu().numpy()


def test_noalloc_force_equals_sf_force_dev_on_consecutive_calls():
    lib = L.load()
    L.set_params(epsilon=1e-2, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    try:
        # full size with equal masses (the pack kernel's uniform tiles), then
        # smaller sizes, one not a multiple of 1024, with differing masses
        for N, seed, equal in ((L.SF_NOALLOC_MAX_N, 1, True), (5000, 2, False), (1024, 3, True)):
            print(f"sf_force_dev_noalloc N={N} equal_mass={equal}", flush=True)
            x, m = _inputs(N, seed, equal)
            rc_ref, acc_ref, pot_ref = _force(lib.sf_force_dev, N, x, m)
            assert rc_ref == 0, L.last_error()
            rc, acc, pot = _force(lib.sf_force_dev_noalloc, N, x, m)
            assert rc == 0, L.last_error()
            assert np.array_equal(acc.view(np.uint32), acc_ref.view(np.uint32))
            assert np.array_equal(pot.view(np.uint32), pot_ref.view(np.uint32))
    finally:
        L.set_params()


def test_noalloc_force_refuses_sizes_beyond_its_scratch():
    lib = L.load()
    x, m = _inputs(16, 4, True)
    N = L.SF_NOALLOC_MAX_N + 1
    assert lib.sf_force_dev_noalloc(N, x.data_ptr(), m.data_ptr(), x.data_ptr(), m.data_ptr(), None) == -1
    assert "exceeds" in L.last_error()
    assert lib.sf_force_dev_noalloc(0, None, None, None, None, None) == 0
