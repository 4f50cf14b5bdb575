"""Force, force VJP and the adjoint KDK on the device (sf_force_dev,
sf_force_vjp_dev, sf_adjoint_forward/backward through the torch binding),
against the FP64 oracle, FP64 central finite differences and an FP64 torch
autograd of a plain KDK forward."""
import numpy as np
import pytest
import torch

from starflood_amd import _lib as L
from starflood_amd import adjoint as A

pytestmark = pytest.mark.gpu

EPS, DT = 1e-2, 1e-3


def two_spheres(N, seed, a=0.5, sep=1.5, vrel=0.3):
    """Two Plummer-profile spheres (radii by inverse CDF, isotropic
    directions) at (+-sep, 0, 0) approaching at vrel, small random
    velocities, total mass 1."""
    rng = np.random.default_rng(seed)
    h = N // 2
    u = np.clip(rng.random(N), 1e-12, 0.999)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, 20 * a)
    d = rng.normal(size=(N, 3))
    x = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    x[:h, 0] -= sep
    x[h:, 0] += sep
    v = rng.normal(scale=0.3, size=(N, 3))
    v[:h, 0] += 0.5 * vrel
    v[h:, 0] -= 0.5 * vrel
    m = np.full(N, 1.0 / N)
    return x.astype(np.float32), v.astype(np.float32), m.astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def rel_rows(a, b):
    nb = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(nb, 1e-3 * nb.max())


@pytest.fixture(params=[L.SF_ACC_PHYSICAL, L.SF_ACC_REFERENCE], ids=["physical", "reference"])
def acc_mode(request):
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=request.param, precision=L.SF_PRECISION_FAST)
    return request.param


def test_force_dev_matches_fp64(oracle, acc_mode):
    N = 1000                                           # not a multiple of the 1024 padding
    x, _, m = two_spheres(N, 1)
    m *= np.random.default_rng(2).uniform(0.5, 1.5, N).astype(np.float32)
    acc, _ = A.force(dev(x), dev(m))
    ref = oracle.force_f64(x, m, EPS, 1.0, int(acc_mode == L.SF_ACC_PHYSICAL))
    assert rel_rows(acc.cpu().numpy().astype(np.float64), ref).max() <= 1e-5


def test_force_vjp_matches_fp64(oracle, acc_mode):
    """(da/dx)^T mu against the FP64 restatement (itself pinned by central
    differences of the FP64 force): 1e-5 of the per-particle magnitude."""
    N = 1000
    x, _, m = two_spheres(N, 3)
    m *= np.random.default_rng(4).uniform(0.5, 1.5, N).astype(np.float32)
    mu = np.random.default_rng(5).normal(size=(N, 3)).astype(np.float32)
    g = A.force_vjp(dev(x), dev(m), dev(mu)).cpu().numpy().astype(np.float64)
    ref = oracle.force_vjp_f64(x, m, mu, EPS, 1.0, int(acc_mode == L.SF_ACC_PHYSICAL))
    assert rel_rows(g, ref).max() <= 1e-5


def run_adjoint(x0, v0, m, nsteps, ckpt, gxN):
    """GPU (x_N, v_N) and the gradient of <gxN, x_N> w.r.t. (x0, v0)."""
    tx, tv = dev(x0).requires_grad_(True), dev(v0).requires_grad_(True)
    xN, vN = A.kdk(tx, tv, dev(m), nsteps, ckpt)
    xN.backward(dev(gxN))
    return (xN.detach().cpu().numpy(), vN.detach().cpu().numpy(), tx.grad.cpu().numpy().astype(np.float64),
            tv.grad.cpu().numpy().astype(np.float64))


def test_adjoint_matches_fp64_adjoint_and_central_differences(oracle):
    """Loss L = mean((x_N - target)^2), target from ICs perturbed by
    N(0, 1e-3).  The GPU backward is driven with the FP64 forward's
    cotangent dL/dx_N so that the FP32 forward's rounding does not enter the
    comparison; its gradient must agree with the FP64 discrete adjoint and
    with FP64 central differences of L along random directions to 1e-4."""
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    N, K = 1024, 60
    x0, v0, m = two_spheres(N, 7)
    rng = np.random.default_rng(8)
    xt, vt = oracle.kdk_f64(x0 + rng.normal(scale=1e-3, size=x0.shape), v0, m, EPS, 1.0, DT, 1, K)
    xN64, _ = oracle.kdk_f64(x0, v0, m, EPS, 1.0, DT, 1, K)
    gxN = 2.0 * (xN64 - xt) / xN64.size
    xN, _, gx, gv = run_adjoint(x0, v0, m, K, 20, gxN)
    assert np.abs(xN - xN64).max() <= 1e-5              # FP32 forward vs FP64
    loss64, gx64, gv64 = oracle.kdk_adjoint_f64(x0, v0, m, EPS, 1.0, DT, 1, K, xt)
    for g, r in ((gx, gx64), (gv, gv64)):
        assert np.linalg.norm(g - r) <= 1e-4 * np.linalg.norm(r)
    # central differences in FP64
    for k in range(3):
        dx = rng.normal(size=x0.shape)
        dv = rng.normal(size=v0.shape)
        for d, g in ((dx, gx), (dv, gv)):
            hstep = 1e-5
            pert = (d, 0) if d is dx else (0, d)
            lp = ((oracle.kdk_f64(x0 + hstep * pert[0], v0 + hstep * pert[1], m, EPS, 1.0, DT, 1, K)[0] - xt) ** 2).mean()
            lm = ((oracle.kdk_f64(x0 - hstep * pert[0], v0 - hstep * pert[1], m, EPS, 1.0, DT, 1, K)[0] - xt) ** 2).mean()
            fd = (lp - lm) / (2 * hstep)
            assert abs(float((g * d).sum()) - fd) <= 1e-4 * abs(fd), (k, fd)


def test_adjoint_reference_acc_mode_matches_fp64_adjoint(oracle):
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_REFERENCE)
    N, K = 512, 25
    x0, v0, m = two_spheres(N, 9)
    m *= 8.0                                            # acc = m a: forces of order 1
    rng = np.random.default_rng(10)
    xt = x0 + rng.normal(scale=1e-2, size=x0.shape)
    xN64, _ = oracle.kdk_f64(x0, v0, m, EPS, 1.0, DT, 0, K)
    gxN = 2.0 * (xN64 - xt) / xN64.size
    _, _, gx, gv = run_adjoint(x0, v0, m, K, 6, gxN)
    _, gx64, gv64 = oracle.kdk_adjoint_f64(x0, v0, m, EPS, 1.0, DT, 0, K, xt)
    for g, r in ((gx, gx64), (gv, gv64)):
        assert np.linalg.norm(g - r) <= 1e-4 * np.linalg.norm(r)


def test_checkpoint_interval_does_not_change_result():
    """Recomputed segments replay the forward bit for bit, so the gradient
    does not depend on the checkpoint interval (1, a divisor, a
    non-divisor, the whole run)."""
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    N, K = 700, 30
    x0, v0, m = two_spheres(N, 11)
    gxN = np.random.default_rng(12).normal(size=x0.shape)
    runs = [run_adjoint(x0, v0, m, K, c, gxN) for c in (1, 5, 7, K, 1000)]
    for r in runs[1:]:
        for a, b in zip(r, runs[0]):
            assert np.array_equal(a, b)



def test_backward_twice_on_one_tape():
    """A second backward on the same tape (autograd's retain_graph)
    recomputes the last segment, whose end position x_N the first backward
    read from the forward's tape: both gradients are bit-identical, for a
    checkpoint every step, a non-divisor and the whole run."""
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    N, K = 700, 30
    x0, v0, m = two_spheres(N, 11)
    gxN = dev(np.random.default_rng(12).normal(size=x0.shape))
    gvN = dev(np.random.default_rng(13).normal(size=x0.shape))
    for ckpt in (1, 7, K):
        tx, tv = dev(x0).requires_grad_(True), dev(v0).requires_grad_(True)
        xN, vN = A.kdk(tx, tv, dev(m), K, ckpt)
        torch.autograd.backward((xN, vN), (gxN, gvN), retain_graph=True)
        first = (tx.grad.clone(), tv.grad.clone())
        tx.grad, tv.grad = None, None
        torch.autograd.backward((xN, vN), (gxN, gvN))
        assert torch.equal(tx.grad, first[0]), ckpt
        assert torch.equal(tv.grad, first[1]), ckpt

def test_matches_fp64_torch_autograd_of_plain_forward():
    """Reference gradient: autograd through a plain FP64 torch KDK (CPU),
    with both the cotangent on x_N and on v_N."""
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    N, K = 64, 12
    x0, v0, m = two_spheres(N, 13)

    def acc(x, mm):
        d = x[None, :, :] - x[:, None, :]                 # x_j - x_i
        s2 = (d * d).sum(-1) + EPS * EPS
        return (mm[None, :, None] * d * s2[..., None] ** -1.5).sum(1)

    tx = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    tm = torch.tensor(m, dtype=torch.float64)
    x, v = tx, tv
    a = acc(x, tm)
    for _ in range(K):
        v = v + 0.5 * DT * a
        x = x + DT * v
        a = acc(x, tm)
        v = v + 0.5 * DT * a
    rng = np.random.default_rng(14)
    cx, cv = rng.normal(size=x0.shape), rng.normal(size=v0.shape)
    ((x * torch.tensor(cx)).sum() + (v * torch.tensor(cv)).sum()).backward()

    gx_t, gv_t = dev(x0).requires_grad_(True), dev(v0).requires_grad_(True)
    xN, vN = A.kdk(gx_t, gv_t, dev(m), K, 5)
    torch.autograd.backward([xN, vN], [dev(cx), dev(cv)])
    for g, r in ((gx_t.grad, tx.grad), (gv_t.grad, tv.grad)):
        g, r = g.cpu().numpy().astype(np.float64), r.numpy()
        assert np.linalg.norm(g - r) <= 1e-5 * np.linalg.norm(r)


def test_zero_steps_and_null_cotangent():
    L.set_params(epsilon=EPS, dt=DT, G=1.0, acc_mode=L.SF_ACC_PHYSICAL)
    x0, v0, m = two_spheres(256, 15)
    tx, tv = dev(x0).requires_grad_(True), dev(v0).requires_grad_(True)
    xN, vN = A.kdk(tx, tv, dev(m), 0, 4)
    assert torch.equal(xN, tx.detach()) and torch.equal(vN, tv.detach())
    g = torch.ones_like(vN)
    vN.backward(g)                                      # no cotangent on x_N
    assert torch.equal(tv.grad, g) and torch.count_nonzero(tx.grad) == 0
