"""The irra_render extension's forward pass (render, render_exe_program/render.cu:93-171,
553-591) on the GPU against the FP64 oracle's pixel clipping."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mesh(seed, n=300):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.uniform(0.02, 0.98, (n, 2)),
                          [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]]])
    tri = Delaunay(pts).simplices.astype(np.int32)
    return pts, tri, rng.uniform(0.5, 1.5, tri.shape[0]) / tri.shape[0]


def test_render_matches_oracle(gpu_available):
    """A Delaunay mesh of the unit square on a 40 x 56 grid: every pixel against the oracle's
    clip of each triangle (FP32 output and atomics: 2e-6 of the brightest pixel); the image
    holds the whole energy; the output layout is the reference's (row pm - 1 - ix, column iy)."""
    import torch

    from caustic_amd import irra_render

    print("[case] irra_render forward, Delaunay mesh 40x56", flush=True)
    pts, tri, e = _mesh(3)
    gray, pm, pn = 2.5, 40, 56
    proj = (0.0, 1.0, 0.0, 1.0)
    dev = torch.device("cuda:0")
    (img,) = irra_render.render(torch.tensor(e, dtype=torch.float32, device=dev),
                                torch.tensor(tri, device=dev),
                                torch.tensor(pts, dtype=torch.float32, device=dev),
                                torch.tensor([gray], dtype=torch.float32, device=dev),
                                torch.tensor(proj, dtype=torch.float32, device=dev),
                                torch.tensor([pm, pn], dtype=torch.float32, device=dev))
    img = img.cpu().numpy().astype(np.float64)
    assert img.shape == (pm, pn)
    # the oracle clips the triangles of the FP32-rounded points, image [iy][ix]
    p32 = pts.astype(np.float32).astype(np.float64)
    cells = (np.arange(tri.shape[0], dtype=np.int32) * 3, np.full(tri.shape[0], 3, np.int32),
             np.ascontiguousarray(p32[tri].reshape(-1)))
    e32 = e.astype(np.float32).astype(np.float64)
    grid = O.render(cells, e32 * gray, (pn, pm), proj)
    ref = grid.T[::-1]
    np.testing.assert_allclose(img, ref, rtol=0, atol=2e-6 * ref.max())
    assert abs(img.sum() - gray * e32.sum()) < 1e-5 * gray * e32.sum()


def test_render_triangle_partly_outside_and_bad_index(gpu_available):
    """A triangle reaching past x = 1 keeps only the energy of its part inside, spread over its
    whole area (compute_render divides by the unclipped area): 3/4 here; a face naming a point
    that does not exist is refused before any kernel runs."""
    import torch

    from caustic_amd import _lib as L
    from caustic_amd import irra_render

    print("[case] irra_render forward, clipped triangle + bad index", flush=True)
    dev = torch.device("cuda:0")
    pts = torch.tensor([[0.5, 0.2], [1.5, 0.2], [0.5, 0.7]], dtype=torch.float32, device=dev)
    faces = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=dev)
    e = torch.tensor([0.8], dtype=torch.float32, device=dev)
    args = (torch.tensor([1.0], device=dev), torch.tensor([0.0, 1.0, 0.0, 1.0], device=dev),
            torch.tensor([16.0, 16.0], device=dev))
    (img,) = irra_render.render(e, faces, pts, *args)
    assert abs(float(img.sum()) - 0.75 * 0.8) < 1e-6
    with pytest.raises(L.CausticLibError):
        irra_render.render(e, torch.tensor([[0, 1, 3]], dtype=torch.int32, device=dev), pts,
                           *args)
