"""The Gaussian-splat receiver renderer (utils/basic_function.py:674-772) on the GPU against
the float64 oracle: image, loss, and the points' gradient by central differences with the
window sums held fixed, as the reference's detached normalisation does."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_splat_two_pass_matches_oracle(gpu_available):
    import torch

    from caustic_amd import splat

    print("[case] gaussian splat two-pass, 400 points, 24x32", flush=True)
    rng = np.random.default_rng(21)
    proj = (0.0, 2.0, 0.0, 1.5)
    ry, rx, sigma, radius = 24, 32, 0.7, 2
    pts = np.column_stack([rng.uniform(0, 2, 400), rng.uniform(0, 1.5, 400)])
    pts = np.concatenate([pts, [[2.5, 0.5], [np.nan, 0.2], [-0.1, 1.0]]])  # masked out
    pts = pts.astype(np.float32).astype(np.float64)
    en = rng.uniform(0.5, 1.5, pts.shape[0]).astype(np.float32).astype(np.float64)
    target = rng.uniform(0, 2, (ry, rx))
    gray, rms = 3.0, 0.8
    dev = torch.device("cuda:0")
    vp = torch.tensor(pts, dtype=torch.float64, device=dev, requires_grad=True)
    img, loss = splat.render_gaussian_splat_two_pass_backward(
        None, gray, vp, torch.tensor(en, dtype=torch.float64, device=dev), (ry, rx), proj,
        torch.tensor(target, device=dev), rms, sigma=sigma, radius=radius)
    legal = np.isfinite(pts).all(axis=1)
    ins = legal & (pts[:, 0] >= 0) & (pts[:, 0] < 2) & (pts[:, 1] >= 0) & (pts[:, 1] < 1.5)
    norm = gray / en[legal].sum()

    # the reference detaches the window sums (utils/basic_function.py:699): its gradient, and
    # so the differences below, hold them at the unperturbed points
    sums = O.splat_window_sums(pts[ins], ry, rx, proj, sigma, radius)

    def oracle_loss(p):
        im = O.gaussian_splat(p[ins], en[ins], ry, rx, proj, sigma, radius, sums) * norm
        return im, np.sqrt(np.mean((im - target) ** 2)) / rms

    ref, L0 = oracle_loss(pts)
    np.testing.assert_allclose(img.cpu().numpy(), ref, rtol=0, atol=2e-6 * ref.max())
    assert abs(float(loss) - L0) < 1e-5 * L0
    g = vp.grad.cpu().numpy()
    assert not g[~ins].any()
    h = 1e-6
    for i in list(np.nonzero(ins)[0][:12]):
        for c in range(2):
            p1, p2 = pts.copy(), pts.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = (oracle_loss(p1)[1] - oracle_loss(p2)[1]) / (2 * h)
            assert abs(g[i, c] - fd) <= 2e-3 * np.abs(g).max() + 1e-3 * abs(fd), (i, c, g[i, c], fd)
