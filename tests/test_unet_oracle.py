"""The CPU oracle of the warm-start U-Net (oracle/unet_oracle.py): pinned by central finite
differences and by torch's CPU autograd of the same architecture (a library routine, not the
code under test), in FP32 without BF16 rounding; the BF16 rounding itself against known values."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def _small_batch(seed=0, B=2, S=16):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, S, S, 1)).astype(np.float32)
    t = rng.standard_normal((B, S, S, 1)).astype(np.float32)
    t -= t.mean(axis=(1, 2, 3), keepdims=True)
    return x, t


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.0e-3, 65504.0], dtype=np.float32)
    r = U.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 1.5*2^-8 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == x[3]
    assert abs(r[4] - x[4]) <= 2.0 ** -8 * abs(x[4])  # half a unit in the 8-bit significand
    assert (U.bf16_round(r) == r).all()


def test_parameter_count_is_about_one_and_a_half_million():
    assert 1.4e6 < U.n_params() < 1.6e6


def test_gradients_match_finite_differences():
    """float64 network (no BF16 rounding): central differences with a small step, so that no
    ReLU changes state inside it, agree with the analytic gradient."""
    P = {k: v.astype(np.float64) for k, v in U.init_params(3).items()}
    x, t = _small_batch(1)
    x, t = x.astype(np.float64), t.astype(np.float64)
    out, acts = U.forward(P, x, bf16=False, keep=True)
    _, dout = U.mse_loss(out, t)
    G = U.backward(P, acts, dout, bf16=False)
    rng = np.random.default_rng(7)

    def loss_of(Q):
        return U.mse_loss(U.forward(Q, x, bf16=False), t)[0]

    checked = 0
    for name in ("e0a.w", "e2b.w", "e3a.b", "d1a.w", "d0b.b", "out.w", "out.b"):
        for _ in range(3):
            idx = tuple(int(rng.integers(0, s)) for s in P[name].shape)
            Q = {k: v.copy() for k, v in P.items()}
            h = 1e-6
            Q[name][idx] += h
            lp = loss_of(Q)
            Q[name][idx] -= 2 * h
            lm = loss_of(Q)
            fd = (lp - lm) / (2 * h)
            g = float(G[name][idx])
            assert abs(fd - g) <= 1e-4 * max(abs(g), abs(fd)) + 1e-9, (name, idx, fd, g)
            checked += 1
    assert checked == 21


def test_forward_and_gradients_match_torch_autograd():
    torch = pytest.importorskip("torch")
    import torch.nn.functional as F

    P = U.init_params(5)
    x, t = _small_batch(2, B=2, S=16)
    out, acts = U.forward(P, x, bf16=False, keep=True)
    loss, dout = U.mse_loss(out, t)
    G = U.backward(P, acts, dout, bf16=False)

    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in P.items()}
    spec = {L[0]: L for L in U.LAYERS}

    def cv(name, inp):
        _, k, ci, co, relu = spec[name]
        w = T[name + ".w"].reshape(k, k, ci, co).permute(3, 2, 0, 1)  # [co, ci, ky, kx]
        y = F.conv2d(inp, w, T[name + ".b"], padding=k // 2)
        return torch.relu(y) if relu else y

    xt = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    s0 = cv("e0b", cv("e0a", xt))
    s1 = cv("e1b", cv("e1a", F.avg_pool2d(s0, 2)))
    s2 = cv("e2b", cv("e2a", F.avg_pool2d(s1, 2)))
    s3 = cv("e3b", cv("e3a", F.avg_pool2d(s2, 2)))
    upf = lambda z: F.interpolate(z, scale_factor=2, mode="nearest")
    t2 = cv("d2b", cv("d2a", torch.cat([upf(s3), s2], 1)))
    t1 = cv("d1b", cv("d1a", torch.cat([upf(t2), s1], 1)))
    t0 = cv("d0b", cv("d0a", torch.cat([upf(t1), s0], 1)))
    o = cv("out", t0)
    on = o.permute(0, 2, 3, 1).reshape(2, -1)
    e = on - on.mean(1, keepdim=True) - torch.tensor(t.reshape(2, -1), dtype=torch.float64)
    L = (e * e).mean()
    L.backward()
    np.testing.assert_allclose(o.permute(0, 2, 3, 1).detach().numpy(), out, rtol=1e-4, atol=1e-5)
    assert abs(L.item() - loss) <= 1e-5 * abs(loss)
    scale = max(np.abs(T[k].grad.numpy()).max() for k in P)
    for k in P:
        ref = T[k].grad.numpy()
        # the output bias's gradient is zero (the loss ignores the mean of the prediction)
        np.testing.assert_allclose(G[k], ref, rtol=2e-3,
                                   atol=2e-4 * np.abs(ref).max() + 1e-6 * scale, err_msg=k)


def test_bf16_oracle_differs_from_fp32_only_at_bf16_level():
    P = U.init_params(2)
    x, _ = _small_batch(4)
    a = U.forward(P, x, bf16=True)
    b = U.forward(P, x, bf16=False)
    rel = np.abs(a - b).max() / np.abs(b).max()
    assert 1e-5 < rel < 5e-2


def test_field_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    out = rng.standard_normal((2, 8, 8, 1))
    t = rng.standard_normal((2, 8, 8, 1))
    t -= t.mean(axis=(1, 2, 3), keepdims=True)
    loss, d = U.field_loss(out, t, grad_weight=3.0)
    for _ in range(10):
        idx = tuple(int(rng.integers(0, s)) for s in out.shape)
        h = 1e-6
        o2 = out.copy()
        o2[idx] += h
        lp = U.field_loss(o2, t, 3.0)[0]
        o2[idx] -= 2 * h
        lm = U.field_loss(o2, t, 3.0)[0]
        assert abs((lp - lm) / (2 * h) - d[idx]) <= 1e-6 * max(1.0, abs(d[idx]))
    # a constant shift of the prediction changes nothing (the gauge)
    assert abs(U.field_loss(out + 5.0, t, 3.0)[0] - loss) < 1e-12
