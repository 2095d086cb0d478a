"""GPU tests of the learned preconditioner (U-Net kernels against the FP64
CPU oracle) and of the equal-size batched solvers (plain CG, mixed CG,
flexible CG) against an FP64 residual computed on the CPU."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from oracle import unet_oracle as U  # noqa: E402

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def learned():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from similie_amd import learned as Lr

    return Lr


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_params(rng):
    p = np.zeros(U.N_PARAMS)
    p[: 9 * U.C * 2] = rng.standard_normal(9 * U.C * 2) * np.sqrt(2 / 18)
    p[9 * U.C * 2: U.N_IN] = rng.standard_normal(U.C) * 0.05
    for l in range(1, U.N_LAYERS + 1):
        ws, bs = U.layer_slices(l)
        p[ws] = rng.standard_normal(9 * U.C * U.C) * np.sqrt(2 / (9 * U.C))
        p[bs] = rng.standard_normal(U.C) * 0.05
    p[U.OUT_BASE:] = rng.standard_normal(U.N_OUT) * np.sqrt(1 / (9 * U.C))
    return p.astype(np.float32)


def weights_of(kappa):
    wx = np.empty_like(kappa)
    wy = np.empty_like(kappa)
    wx[:, :-1, :] = 0.5 * (kappa[:, :-1, :] + kappa[:, 1:, :])
    wx[:, -1, :] = kappa[:, -1, :]
    wy[:, :, :-1] = 0.5 * (kappa[:, :, :-1] + kappa[:, :, 1:])
    wy[:, :, -1] = kappa[:, :, -1]
    return wx, wy


# BF16 weights and activations with FP32 accumulation against FP64: the
# tolerance is a few BF16 roundings (2^-8) accumulated over 14 layers
FWD_TOL = 3e-2
GRAD_TOL = 6e-2


@pytest.mark.parametrize("shape", [(2, 32, 32), (1, 16, 40), (3, 64, 64)])
def test_unet_forward_matches_oracle(learned, shape):
    B, H, W = shape
    rng = np.random.default_rng(10 + H)
    params = make_params(rng)
    x = rng.standard_normal((B, H, W, 2)).astype(np.float32)
    net = learned.UNet(B, H, W, seed=1)
    net.set_params(torch.from_numpy(params))
    out = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = U.forward(params.astype(np.float64), x)
    assert rel(out, ref) <= FWD_TOL


def test_unet_backward_matches_oracle(learned):
    B, H, W = 2, 32, 32
    rng = np.random.default_rng(21)
    params = make_params(rng)
    x = rng.standard_normal((B, H, W, 2)).astype(np.float32)
    dout = rng.standard_normal((B, H, W)).astype(np.float32)
    net = learned.UNet(B, H, W, seed=1)
    net.set_params(torch.from_numpy(params))
    xt = torch.from_numpy(x).cuda()
    net.forward(xt)
    g = net.backward(xt, torch.from_numpy(dout).cuda()).cpu().numpy()
    _, act = U.forward(params.astype(np.float64), x)
    gref = U.backward(params.astype(np.float64), act, dout)
    # every parameter block separately: conv_in, the twelve layers, conv_out
    blocks = [slice(0, U.N_IN)] + [slice(U.layer_slices(l)[0].start, U.layer_slices(l)[1].stop)
                                    for l in range(1, U.N_LAYERS + 1)] + [slice(U.OUT_BASE, U.N_PARAMS)]
    errs = [rel(g[s], gref[s]) for s in blocks]
    assert max(errs) <= GRAD_TOL, errs


def test_unet_direction_loss_matches_oracle(learned):
    B, n = 2, 32
    rng = np.random.default_rng(31)
    kappa = np.exp(rng.standard_normal((B, n, n)))
    wx, wy = weights_of(kappa)
    e = rng.standard_normal((B, n, n))
    e[:, [0, -1], :] = 0
    e[:, :, [0, -1]] = 0
    r = U.apply5(e, wx, wy)
    N = rng.standard_normal((B, n, n))
    x = np.stack([r, np.log(kappa)], axis=-1).astype(np.float32)
    net = learned.UNet(B, n, n)
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    loss, dN = net.loss(f32(N), f32(x), f32(e), f32(wx), f32(wy))
    lref, dref = U.direction_loss(N, x[..., 0].astype(np.float64), e.astype(np.float32).astype(np.float64),
                                  wx.astype(np.float32).astype(np.float64), wy.astype(np.float32).astype(np.float64))
    assert np.allclose(loss.cpu().numpy(), lref, rtol=1e-4, atol=1e-5)
    assert rel(dN.cpu().numpy(), dref) <= 1e-4


def test_unet_training_reduces_the_loss(learned):
    """A few Adam steps on one fixed batch lower the direction loss."""
    B, n = 4, 32
    rng = np.random.default_rng(41)
    kappa = np.exp(rng.standard_normal((B, n, n)))
    wx, wy = weights_of(kappa)
    e = rng.standard_normal((B, n, n))
    e[:, [0, -1], :] = 0
    e[:, :, [0, -1]] = 0
    r = U.apply5(e, wx, wy)
    s = np.sqrt(np.mean(r * r, axis=(1, 2)))[:, None, None]
    x = np.stack([r / s, np.log(kappa)], axis=-1)
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    net = learned.UNet(B, n, n, seed=3)
    xt, et, wxt, wyt = f32(x), f32(e / s), f32(wx), f32(wy)
    first = net.train_step(xt, et, wxt, wyt, 1e-3).mean().item()
    for _ in range(60):
        last = net.train_step(xt, et, wxt, wyt, 1e-3).mean().item()
    assert np.isfinite(last) and last < 0.8 * first


# ----------------------------------------------------------- batch solves ---

def problems(B, n, seed):
    rng = np.random.default_rng(seed)
    kappa = np.exp(0.8 * rng.standard_normal((B, n, n)))
    b = rng.standard_normal((B, n, n))
    b[:, [0, -1], :] = 0
    b[:, :, [0, -1]] = 0
    return kappa, b


def test_batch_apply_matches_oracle(learned):
    kappa, _ = problems(3, 19, 5)
    x = np.random.default_rng(6).standard_normal(kappa.shape)
    pb = learned.PoissonBatch(torch.from_numpy(kappa))
    y = pb.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    wx, wy = weights_of(kappa)
    assert rel(y, U.apply5(x, wx, wy)) <= 1e-14


@pytest.mark.parametrize("method", ["cg", "cg_mixed", "fcg_learned", "fcg_learned_mixed"])
def test_batch_solvers_reach_fp64_verified_tolerance(learned, method):
    B, n = 5, 32
    kappa, b = problems(B, n, 7)
    pb = learned.PoissonBatch(torch.from_numpy(kappa))
    net = learned.UNet(B, n, n, seed=2) if method.startswith("fcg") else None
    x, res, secs = pb.solve(torch.from_numpy(b).cuda(), method=method, tol=1e-10, max_iterations=4000, net=net)
    wx, wy = weights_of(kappa)
    xr = x.cpu().numpy()
    for p in range(B):
        r = b[p] - U.apply5(xr[p: p + 1], wx[p: p + 1], wy[p: p + 1])[0]
        true_rel = np.linalg.norm(r) / np.linalg.norm(b[p])
        assert res[p]["converged"], (method, p, res[p])
        assert true_rel <= 1e-10 * 1.01, (method, p, true_rel)
        assert abs(true_rel - res[p]["true_relative_residual"]) <= 1e-3 * 1e-10 + 1e-3 * true_rel


def test_batch_record_returns_residual_history(learned):
    B, n = 2, 32
    kappa, b = problems(B, n, 8)
    pb = learned.PoissonBatch(torch.from_numpy(kappa))
    x, res, secs, (rr, xr) = pb.solve(torch.from_numpy(b).cuda(), method="cg", tol=1e-10, record_every=4,
                                      max_records=400)
    assert rr.shape[0] >= 2
    # record 0 is the initial residual b (x0 = 0)
    assert rel(rr[0].cpu().numpy(), b) <= 1e-6
    wx, wy = weights_of(kappa)
    k = rr.shape[0] // 2
    r_k = b - U.apply5(xr[k].cpu().numpy().astype(np.float64), wx, wy)
    assert rel(rr[k].cpu().numpy(), r_k) <= 1e-4
