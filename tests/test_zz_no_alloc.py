"""cot_irra_render_grad_noalloc (no device allocation: the Sobel weight image in the library's
fixed device array) against cot_irra_render_grad, through the C ABI: two calls in a row, the
second with a smaller mesh and image, each to the FP32 atomics' reordering (1e-6 of each
output's largest magnitude; the same kernels, summed in a different order)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(seed, n, pm, pn):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.uniform(0.02, 0.98, (n, 2)),
                          [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]])
    tri = Delaunay(pts).simplices.astype(np.int32)
    e = rng.uniform(0.5, 1.5, tri.shape[0]) / tri.shape[0]
    real = rng.uniform(0.0, 3.0, pm * pn)
    return pts, tri, e, real


def _both(pts, tri, e, real, pm, pn):
    import torch

    from caustic_amd import irra_render

    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731
    args = (t(e), t(tri, torch.int32), t(pts), t([1.7]), t(real), t([True], torch.bool),
            t((0.0, 1.0, 0.0, 1.0)), t([pm, pn]))
    ref = [o.cpu().numpy() for o in irra_render.render_grad(*args)]
    got = [o.cpu().numpy() for o in irra_render.render_grad(*args, no_alloc=True)]
    return ref, got


def test_noalloc_render_grad_matches_allocating_entry_point(gpu_available):
    for seed, n, pm, pn in ((5, 300, 64, 80), (6, 100, 32, 40)):  # the second call smaller
        print(f"[case] irra_render_grad no-alloc vs allocating, {n} points {pm}x{pn}", flush=True)
        pts, tri, e, real = _case(seed, n, pm, pn)
        ref, got = _both(pts, tri, e, real, pm, pn)
        assert len(ref) == len(got) == 9
        for k, (a, b) in enumerate(zip(ref, got)):
            assert a.shape == b.shape, k
            np.testing.assert_allclose(b, a, rtol=0, atol=1e-6 * max(np.abs(a).max(), 1e-30),
                                       err_msg=f"output {k}")
        assert np.abs(ref[7]).max() > 0  # the Sobel branch ran (its scratch image was used)


def test_noalloc_render_grad_rejects_oversize_and_bad_faces(gpu_available):
    """An image over COT_NOALLOC_MAX_PIXELS with the Sobel branch is refused before any
    kernel runs; a face naming a point that does not exist is refused by the device check."""
    import torch

    from caustic_amd import _lib as L

    lib = L.load()
    dev = torch.device("cuda:0")
    f = torch.zeros(16, dtype=torch.float32, device=dev)
    faces = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=dev)
    xy = torch.tensor([[0.1, 0.1], [0.9, 0.1], [0.5, 0.9]], dtype=torch.float32, device=dev)
    dom = (ctypes.c_double * 4)(0.0, 1.0, 0.0, 1.0)
    p = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

    def call(pm, pn, fc, sobel=1):
        return lib.cot_irra_render_grad_noalloc(p(f), p(fc), 1, p(xy), 3, 1.0, p(f), dom, pm, pn,
                                                *([p(f)] * 7), sobel, p(f), p(f), stream)

    assert call(2048, 1024, faces) == L.COT_EINVAL
    assert b"exceeds" in lib.cot_last_error()
    bad = torch.tensor([[0, 1, 3]], dtype=torch.int32, device=dev)
    assert call(4, 4, bad) == L.COT_EINVAL
    assert b"face index" in lib.cot_last_error()
    torch.cuda.synchronize()
