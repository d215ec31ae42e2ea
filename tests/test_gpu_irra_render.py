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


def _losses(pts, tri, e, gray, real, pm, pn, proj):
    """The oracle's render (grid layout [iy, ix]), sum_p (render - real)^2 and
    sum_f (inte_f - e_f gray)^2 for the mesh."""
    a_s = O.irra_face_areas(pts, tri, pm, pn, proj, signed=True)  # a face's pieces share its sign
    a = np.abs(a_s)
    u, v = pts[tri[:, 1]] - pts[tri[:, 0]], pts[tri[:, 2]] - pts[tri[:, 0]]
    A = 0.5 * np.abs(u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0])
    img = np.tensordot(e * gray / A, a, axes=1)
    cell = (proj[1] - proj[0]) * (proj[3] - proj[2]) / (pm * pn)
    # the per-face integrals with the faces' orientations (computer_grad_inte's signed areas)
    inte = np.tensordot(a_s, real, axes=([1, 2], [0, 1])) / cell
    return img, float(np.sum((img - real) ** 2)), inte, float(np.sum((inte - e * gray) ** 2))


def test_render_grad_matches_finite_differences(gpu_available):
    """render_grad on a small mesh inside the domain, Sobel branch off and on: image and image - target in the
    output layout, the per-face integrals of the target, and the derivatives of both losses by
    central differences of the FP64 oracle (the clipped areas are piecewise quadratic in the
    vertices, so the differences are exact away from the kinks; FP32 outputs: 2e-3 of the
    largest component)."""
    import torch

    from caustic_amd import irra_render

    print("[case] irra_render render_grad, 24-point mesh 12x10", flush=True)
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(11)
    pts = rng.uniform(0.05, 0.95, (24, 2)).astype(np.float32).astype(np.float64)
    tri = Delaunay(pts).simplices.astype(np.int32)
    e = rng.uniform(0.5, 1.5, tri.shape[0]).astype(np.float32).astype(np.float64)
    gray, pm, pn = 0.75, 12, 10
    proj = (0.0, 1.0, 0.0, 1.0)
    real = rng.uniform(0.0, 3.0, (pn, pm)).astype(np.float32).astype(np.float64)  # [iy, ix]
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731
    out = irra_render.render_grad(t(e), t(tri, torch.int32), t(pts), t([gray]),
                                  t(real.reshape(-1)), t([False], torch.bool), t(proj),
                                  t([pm, pn]))
    out = [o.cpu().numpy().astype(np.float64) for o in out]
    img, L0, inte, J0 = _losses(pts, tri, e, gray, real, pm, pn, proj)
    np.testing.assert_allclose(out[0], img.T[::-1], rtol=0, atol=2e-6 * img.max())
    np.testing.assert_allclose(out[1], (img - real).T[::-1], rtol=0, atol=2e-6 * img.max())
    np.testing.assert_allclose(out[6], inte, rtol=1e-5)
    np.testing.assert_allclose(out[4], inte - e * gray, rtol=0, atol=1e-5 * np.abs(inte).max())
    h = 1e-6
    gL, gJ = np.zeros_like(pts), np.zeros_like(pts)
    for i in range(pts.shape[0]):
        for c in range(2):
            p1, p2 = pts.copy(), pts.copy()
            p1[i, c] += h
            p2[i, c] -= h
            _, l1, _, j1 = _losses(p1, tri, e, gray, real, pm, pn, proj)
            _, l2, _, j2 = _losses(p2, tri, e, gray, real, pm, pn, proj)
            gL[i, c], gJ[i, c] = (l1 - l2) / (2 * h), (j1 - j2) / (2 * h)
    np.testing.assert_allclose(out[2], gL, rtol=0, atol=2e-3 * np.abs(gL).max())
    np.testing.assert_allclose(out[5], gJ, rtol=0, atol=2e-3 * np.abs(gJ).max())
    gE = np.zeros_like(e)
    for f in range(e.shape[0]):  # the image loss is quadratic in the energies
        e1, e2 = e.copy(), e.copy()
        e1[f] += 1e-4
        e2[f] -= 1e-4
        gE[f] = (_losses(pts, tri, e1, gray, real, pm, pn, proj)[1]
                 - _losses(pts, tri, e2, gray, real, pm, pn, proj)[1]) / 2e-4
    np.testing.assert_allclose(out[3], gE, rtol=0, atol=2e-4 * np.abs(gE).max())
    assert not out[7].any() and not out[8].any()

    # the Sobel branch: the same outputs plus the derivatives of the smoothness loss
    # sum over 4-neighbouring pixel pairs of (diff_p - diff_q)^2
    def smooth(p_, e_):
        d = _losses(p_, tri, e_, gray, real, pm, pn, proj)[0] - real
        return float(np.sum((d[:, 1:] - d[:, :-1]) ** 2) + np.sum((d[1:, :] - d[:-1, :]) ** 2))

    sob = irra_render.render_grad(t(e), t(tri, torch.int32), t(pts), t([gray]),
                                  t(real.reshape(-1)), t([True], torch.bool), t(proj),
                                  t([pm, pn]))
    sob = [o.cpu().numpy().astype(np.float64) for o in sob]
    for k in range(7):
        np.testing.assert_allclose(sob[k], out[k], rtol=1e-6, atol=1e-6 * np.abs(out[k]).max())
    gS = np.zeros_like(pts)
    for i in range(pts.shape[0]):
        for c in range(2):
            p1, p2 = pts.copy(), pts.copy()
            p1[i, c] += h
            p2[i, c] -= h
            gS[i, c] = (smooth(p1, e) - smooth(p2, e)) / (2 * h)
    np.testing.assert_allclose(sob[7], gS, rtol=0, atol=2e-3 * np.abs(gS).max())
    gSE = np.zeros_like(e)
    for f in range(e.shape[0]):
        e1, e2 = e.copy(), e.copy()
        e1[f] += 1e-4
        e2[f] -= 1e-4
        gSE[f] = (smooth(pts, e1) - smooth(pts, e2)) / 2e-4
    np.testing.assert_allclose(sob[8], gSE, rtol=0, atol=2e-4 * np.abs(gSE).max())


def test_render_grad_inverted_faces_use_signed_integrals(gpu_available):
    """Faces wound clockwise (the pipeline produces them: its proj_area_loss penalises inverted
    projected triangles): the image and its derivatives do not depend on a face's winding, the
    per-face integral inte_f, inte_diff and the derivative of sum_f inte_diff_f^2 take the
    reference's signed clipped areas (computer_grad_inte, render_exe_program/render.cu:422-542
    with compute_grad_area, clip.h:277-292): inte_f of a reversed face is minus its
    counter-clockwise value, and the gradient matches central differences of the oracle's
    signed integrals."""
    import torch

    from caustic_amd import irra_render
    from scipy.spatial import Delaunay

    print("[case] irra_render render_grad, every third face reversed", flush=True)
    rng = np.random.default_rng(13)
    pts = rng.uniform(0.05, 0.95, (24, 2)).astype(np.float32).astype(np.float64)
    tri0 = Delaunay(pts).simplices.astype(np.int32)
    tri = tri0.copy()
    tri[::3, 1], tri[::3, 2] = tri0[::3, 2], tri0[::3, 1]
    e = rng.uniform(0.5, 1.5, tri.shape[0]).astype(np.float32).astype(np.float64)
    gray, pm, pn = 0.75, 12, 10
    proj = (0.0, 1.0, 0.0, 1.0)
    real = rng.uniform(0.0, 3.0, (pn, pm)).astype(np.float32).astype(np.float64)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731

    def run(faces):
        out = irra_render.render_grad(t(e), t(faces, torch.int32), t(pts), t([gray]),
                                      t(real.reshape(-1)), t([False], torch.bool), t(proj),
                                      t([pm, pn]))
        return [o.cpu().numpy().astype(np.float64) for o in out]

    ccw, out = run(tri0), run(tri)
    rev = np.zeros(tri.shape[0], dtype=bool)
    rev[::3] = True
    img, _, inte, _ = _losses(pts, tri, e, gray, real, pm, pn, proj)
    np.testing.assert_allclose(out[0], ccw[0], rtol=0, atol=1e-6 * np.abs(ccw[0]).max())
    np.testing.assert_allclose(out[2], ccw[2], rtol=0, atol=1e-5 * np.abs(ccw[2]).max())
    np.testing.assert_allclose(out[6][rev], -ccw[6][rev], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(out[6][~rev], ccw[6][~rev], rtol=1e-6, atol=1e-9)
    assert (out[6][rev] < 0).all() and (inte[rev] < 0).all()
    np.testing.assert_allclose(out[6], inte, rtol=1e-5)
    np.testing.assert_allclose(out[4], inte - e * gray, rtol=0, atol=1e-5 * np.abs(inte).max())
    h = 1e-6
    gJ = np.zeros_like(pts)
    for i in range(pts.shape[0]):
        for c in range(2):
            p1, p2 = pts.copy(), pts.copy()
            p1[i, c] += h
            p2[i, c] -= h
            gJ[i, c] = (_losses(p1, tri, e, gray, real, pm, pn, proj)[3]
                        - _losses(p2, tri, e, gray, real, pm, pn, proj)[3]) / (2 * h)
    np.testing.assert_allclose(out[5], gJ, rtol=0, atol=2e-3 * np.abs(gJ).max())


def test_render_grad_edge_on_a_grid_line(gpu_available):
    """Two triangles sharing an edge that lies on the interior grid line x = 1/2 (a mesh aligned
    with the image): each point of the edge belongs to one pixel, so the boundary transport is
    counted once and the x-derivatives of the edge's vertices are the one-sided derivatives for
    moving them towards +x (forward differences of the FP64 oracle); the y-derivatives, where
    the edge's normal has no component, match central differences."""
    import torch

    from caustic_amd import irra_render

    print("[case] irra_render render_grad, edge on the grid line x = 1/2", flush=True)
    pts = np.array([[0.5, 0.2], [0.5, 0.8], [0.2, 0.55], [0.85, 0.4]])
    tri = np.array([[0, 1, 2], [0, 3, 1]], dtype=np.int32)
    e = np.array([0.6, 1.1])
    gray, pm, pn = 1.0, 4, 5
    proj = (0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(17)
    real = rng.uniform(0.0, 3.0, (pn, pm)).astype(np.float32).astype(np.float64)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float32: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731
    out = irra_render.render_grad(t(e), t(tri, torch.int32), t(pts), t([gray]),
                                  t(real.reshape(-1)), t([False], torch.bool), t(proj),
                                  t([pm, pn]))
    g = out[2].cpu().numpy().astype(np.float64)
    L = lambda p_: _losses(p_, tri, e, gray, real, pm, pn, proj)[1]  # noqa: E731
    h = 1e-7
    for i in (0, 1):
        p1 = pts.copy()
        p1[i, 0] += h
        fwd = (L(p1) - L(pts)) / h
        assert abs(g[i, 0] - fwd) <= 1e-3 * max(abs(fwd), 1e-3), (i, g[i, 0], fwd)
        p1, p2 = pts.copy(), pts.copy()
        p1[i, 1] += 1e-6
        p2[i, 1] -= 1e-6
        ctr = (L(p1) - L(p2)) / 2e-6
        assert abs(g[i, 1] - ctr) <= 1e-3 * max(abs(ctr), 1e-3), (i, g[i, 1], ctr)
