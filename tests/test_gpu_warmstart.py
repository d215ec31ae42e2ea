"""The warm-start network on the GPU against its CPU oracle (BF16-rounded operands, float64
sums), and its training."""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _setup(seed, B=2, S=64):
    rng = np.random.default_rng(seed)
    P = U.init_params(seed)
    for k in P:
        if k.endswith(".b"):
            P[k] = (0.05 * rng.standard_normal(P[k].shape)).astype(np.float32)
    x = rng.standard_normal((B, S, S, 1)).astype(np.float32)
    t = rng.standard_normal((B, S, S, 1)).astype(np.float32)
    t -= t.mean(axis=(1, 2, 3), keepdims=True)
    return P, x, t


def _net(S=64, B=2):
    from caustic_amd import warmstart as WS

    return WS, WS.WarmStartNet(S=S, max_batch=B, seed=None)


def test_forward_matches_oracle(gpu_available):
    print("[case] network forward S=64 B=2", flush=True)
    P, x, _ = _setup(1)
    WS, net = _net()
    net.set_params(WS.dict_to_flat(P))
    y = net.forward(x[..., 0])
    ref = U.forward(P, x, bf16=True)[..., 0]
    err = np.abs(y - ref).max() / np.abs(ref).max()
    # identical BF16 operands; only the FP32 summation order differs, which can flip the BF16
    # rounding of an activation now and then
    assert err < 5e-3, err
    net.close()


def test_gradients_match_oracle(gpu_available):
    print("[case] network gradients S=64 B=2", flush=True)
    P, x, t = _setup(2)
    WS, net = _net()
    net.set_params(WS.dict_to_flat(P))
    net.set_loss(7.0)
    g, loss = net.gradients(x[..., 0], t[..., 0])
    lam = 7.0  # the loss's neighbouring-difference weight (set on the network above)

    def oracle(acc):
        U.ACC = acc
        try:
            out, acts = U.forward(P, x, bf16=True, keep=True)
            ref_loss, dout = U.field_loss(out, t, lam)
            return ref_loss, U.backward(P, acts, dout, bf16=True)
        finally:
            U.ACC = np.float64

    ref_loss, G = oracle(np.float64)
    _, G32 = oracle(np.float32)  # the same products summed in FP32: the spread of the order
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    gd = WS.flat_to_dict(g)
    rows = []
    for k in G:
        if k == "out.b":  # exactly zero (the loss ignores the prediction's mean)
            assert abs(float(gd[k][0])) <= 1e-6 * np.abs(G["out.w"]).max()
            continue
        nrm = max(np.linalg.norm(G[k]), 1e-30)
        rel = np.linalg.norm(gd[k] - G[k]) / nrm
        spread = np.linalg.norm(G32[k] - G[k]) / nrm
        cos = float(np.dot(gd[k].ravel(), G[k].ravel()) / (np.linalg.norm(gd[k]) * nrm))
        rows.append((k, rel, spread, cos))
    for r in rows:
        print("  %-6s gpu-vs-oracle %.3e  fp32-order spread %.3e  cosine %.6f" % r)
    for k, rel, spread, cos in rows:
        # the GPU's FP32 order is one more order: its distance stays within the spread that
        # another order shows, with a floor for layers whose spread happens to be tiny
        assert rel < max(4.0 * spread, 5e-3) and cos > 0.999, (k, rel, spread, cos)
    net.close()


def test_training_reduces_the_loss(gpu_available):
    """Adam on 32 smooth synthetic pairs (target a fixed linear map of the input's smoothed
    field): the loss falls by a factor of 4 within 60 steps."""
    print("[case] network training S=64", flush=True)
    from caustic_amd import warmstart as WS

    rng = np.random.default_rng(3)
    S, n = 64, 32
    c = (np.arange(S) + 0.5) / S
    X, Y = np.meshgrid(c, c, indexing="xy")
    xs, ts = [], []
    for _ in range(n):
        a, b, f = rng.normal(size=3)
        field = a * np.sin(2 * np.pi * X * (1 + abs(f))) + b * np.cos(2 * np.pi * Y)
        xs.append(field.astype(np.float32))
        tt = 0.5 * field + 0.2 * field ** 2
        ts.append((tt - tt.mean()).astype(np.float32))
    net = WS.WarmStartNet(S=S, max_batch=8, seed=0)
    net.set_dataset(np.stack(xs), np.stack(ts))
    losses = []
    for ep in range(15):
        losses += list(net.train(rng.permutation(n), batch=8, lr=1e-3))
    assert np.isfinite(losses).all()
    assert np.mean(losses[-4:]) < 0.25 * np.mean(losses[:4]), (losses[:4], losses[-4:])
    st = net.train_stats()
    assert st["samples"] == 15 * n and st["seconds"] > 0
    net.close()


def test_learned_start_gains_over_zero_start(gpu_available):
    """The committed network (caustic_amd/warmstart_unet128.npz: trained on the GPU by this
    package's kernels on 1,152 converged N = 16,384 solves of seeds 100000 and up) against the
    zero start on two held-out targets (seeds 900000 and 900001), each solved to max-rel 1e-5
    (L-BFGS to max-rel 1, then Newton): the learned start is taken (its |mass - nu|_2 is below
    the zero start's), needs at least 10% fewer iterations, and its solution is confirmed by the
    FP64 oracle."""
    import os

    from caustic_amd import warmstart as WS
    from caustic_amd import workloads as W
    from oracle import oracle as O

    assert os.path.exists(WS.DEFAULT_MODEL), "the trained network is missing"
    print("[case] learned vs zero start, N=16384, seeds 900000-900001", flush=True)
    net = WS.WarmStartNet.load(WS.DEFAULT_MODEL)
    rows = WS.evaluate_warm_start(net, [900000, 900001], log=print)
    net.close()
    for r in rows:
        assert r["converged"] and r["learned_used"], r
        assert r["warm_iters"] < 0.9 * r["cold_iters"], (r["warm_iters"], r["cold_iters"])
        p = W.warm_start_sample(r["seed"])
        ev = O.evaluate(p.sites, r["w_warm"], p.nu, p.rho, p.proj_domain)
        assert O.max_rel_error(ev.mass, p.nu) < 1.1e-5
