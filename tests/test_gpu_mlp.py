"""The surrogate network's hand-written MFMA forward/backward kernels against
the CPU FP64 reference (oracle/mlp_ref.py) on the same parameters and batch,
and a short training run that must reduce the loss."""
import numpy as np
import pytest
import torch

from chaos_oneapi_amd import lyap as L
from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu

RANGES = (0.05, 0.4, 2.0, 20.0)
YM, YS = (0.02, -0.01, -8.0), (0.05, 0.02, 4.0)


def batch(n, seed):
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(0.05, 0.4, n), rng.uniform(2, 20, n)], axis=1)
    lam = np.stack([rng.normal(0.02, 0.05, n), rng.normal(-0.01, 0.02, n), rng.normal(-8, 4, n)], 1)
    status = (rng.uniform(size=n) < 0.15).astype(np.uint32)
    return p, lam, status


def dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("n", [64, 200])
def test_gradients_match_cpu_reference(n):
    net = L.Mlp(max_batch=256, hidden=256, n_fourier=32, fourier_sigma=4.0, seed=11, ranges=RANGES,
                y_mean=YM, y_std=YS)
    p, lam, st = batch(n, n)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    print(f"mlp backward n={n}", flush=True)
    net.backward(dev(p, torch.float64), dev(lam, torch.float64), dev(st.view(np.int32), torch.int32), loss)
    torch.cuda.synchronize()
    g_gpu = net.get(1).astype(np.float64)
    params = net.get(0)
    fourier = net.get(2)
    l_ref, g_ref, Z = R.forward_backward(params, fourier, p, lam, st, RANGES, YM, YS, hidden=256)
    assert abs(loss.item() - l_ref) <= 2e-3 * abs(l_ref)
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    cos = g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref))
    print(f"grad rel err {rel:.2e} cos {cos:.6f}", flush=True)
    assert rel < 1e-2 and cos > 0.9999
    # per layer as well (a transposed operand would scramble one layer only)
    dims = [64, 256, 256, 256, 16]
    off = 0
    for l in range(4):
        sz = dims[l + 1] * dims[l] + dims[l + 1]
        gl, rl = g_gpu[off:off + sz], g_ref[off:off + sz]
        assert np.linalg.norm(gl - rl) <= 2e-2 * np.linalg.norm(rl) + 1e-9
        off += sz
    out = torch.empty((n, 5), dtype=torch.float32, device="cuda")
    net.forward(dev(p, torch.float64), out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    np.testing.assert_allclose(o[:, 0], Z[:, 0] * YS[0] + YM[0], rtol=2e-3, atol=2e-4)
    np.testing.assert_allclose(o[:, 3], np.exp(0.5 * Z[:, 3]) * YS[0], rtol=2e-3)
    np.testing.assert_allclose(o[:, 4], 1 / (1 + np.exp(-Z[:, 4])), atol=1e-3)
    net.close()


def test_training_reduces_loss():
    # a smooth synthetic spectrum over the (a, c) plane
    net = L.Mlp(max_batch=1024, seed=5, ranges=RANGES, y_mean=(0, 0, -11), y_std=(0.1, 0.1, 5))
    rng = np.random.default_rng(0)
    n = 1024
    p = np.stack([rng.uniform(0.05, 0.4, n), rng.uniform(2, 20, n)], axis=1)
    lam = np.stack([0.1 * np.sin(8 * p[:, 0]) * np.cos(p[:, 1] / 3), 0 * p[:, 0], -p[:, 1] + p[:, 0]], 1)
    st = (p[:, 0] > 0.38).astype(np.uint32)
    pd, ld, sd = dev(p, torch.float64), dev(lam, torch.float64), dev(st.view(np.int32), torch.int32)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    losses = []
    for it in range(300):
        net.train_step(pd, ld, sd, loss)
        if it % 50 == 0 or it == 299:
            torch.cuda.synchronize()
            losses.append(loss.item())
    print("losses", losses, flush=True)
    assert losses[-1] < losses[0] - 1.0
    net.close()


def test_adam_step_and_bf16_issue_match_cpu():
    """One train_step = backward at the current parameters, Adam on every
    trainable parameter (step 1: m = (1-b1) g, v = (1-b2) g^2, bias-corrected),
    then the new weights issued to the BF16 copies that the forward and the
    input-gradient products read.  Pins: the updated parameters against a
    NumPy Adam of the step's own gradient; the frozen head rows untouched; the
    forward output and the next gradient at the updated parameters against the
    CPU reference (a stale or transposed BF16 copy fails them)."""
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    net = L.Mlp(max_batch=256, hidden=256, n_fourier=32, fourier_sigma=4.0, seed=13, ranges=RANGES,
                y_mean=YM, y_std=YS, lr=lr)
    p, lam, st = batch(200, 3)
    pd, ld, sd = dev(p, torch.float64), dev(lam, torch.float64), dev(st.view(np.int32), torch.int32)
    w0 = net.get(0).astype(np.float64)
    print("mlp train_step n=200", flush=True)
    net.train_step(pd, ld, sd)
    torch.cuda.synchronize()
    w1 = net.get(0).astype(np.float64)
    g = net.get(1).astype(np.float64)
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    ref = w0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    dims = [64, 256, 256, 256, 16]
    # the head's padded rows (outputs 5..15) are frozen at zero
    off = sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(3))
    frozen = np.zeros_like(w0, dtype=bool)
    wh = w0[off:off + 16 * 256].reshape(16, 256)
    frozen[off:off + 16 * 256] = np.repeat(np.arange(16) >= 5, 256)
    frozen[off + 16 * 256 + 5:off + 16 * 256 + 16] = True
    assert np.all(w1[frozen] == w0[frozen]) and np.all(wh[5:] == 0.0)
    np.testing.assert_allclose(w1[~frozen], ref[~frozen], rtol=0, atol=1e-6 * lr + 1e-7)
    fourier = net.get(2)
    out = torch.empty((200, 5), dtype=torch.float32, device="cuda")
    net.forward(pd, out)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    net.backward(pd, ld, sd, loss)
    torch.cuda.synchronize()
    l_ref, g_ref, Z = R.forward_backward(net.get(0), fourier, p, lam, st, RANGES, YM, YS, hidden=256)
    np.testing.assert_allclose(out.cpu().numpy()[:, 0], Z[:, 0] * YS[0] + YM[0], rtol=2e-3, atol=2e-4)
    g2 = net.get(1).astype(np.float64)
    assert np.linalg.norm(g2 - g_ref) <= 1e-2 * np.linalg.norm(g_ref)
    net.close()
