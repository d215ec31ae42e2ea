"""The HIP path at the reference's own configuration (the shipped run: N = 160,978 lattice
sites, 640 x 640 ustc_centerL target with 76.6 % zero pixels, proj_domain 160 units wide,
weights in [-512, 2588]), checked against the FP64 oracle and the shipped outputs."""
import numpy as np
import pytest

import shipped
from caustic_amd.workloads import Problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    S = shipped.load()
    nu = O.source_masses(S.points, S.sample_domain)
    p = Problem(sites=S.proj_xy, nu=nu, rho=S.rho, proj_domain=S.proj_domain)
    ref = O.evaluate(p.sites, S.weights, nu, p.rho, p.proj_domain, centroid_min_piece=0.0)
    return S, p, ref


@pytest.mark.parametrize("precision,mass_tol,sum_tol,cen_tol", [("fp64", 1e-7, 1e-9, 1e-9),
                                                                 ("mixed", 1e-6, 1e-7, 1e-5)])
def test_shipped_weights_evaluation(gpu_available, case, precision, mass_tol, sum_tol, cen_tol):
    """cot_evaluate and cot_centroids at the shipped t = 0 weights: masses within mass_tol of
    the oracle relative to each cell's own mass, every cell present, and the rho-weighted
    centroids reproduce the shipped 0/out.txt as the oracle's do (median 3e-4, 99 % within 1e-2
    units; the residual is the weights' 6-digit truncation)."""
    from caustic_amd.ot import OTBatch

    S, p, ref = case
    print(f"[case] shipped run evaluation {precision}", flush=True)
    with OTBatch([p], precision=precision) as b:
        ev = b.evaluate(S.weights)
        cen, ex = b.centroids(S.weights)
    assert ex.all() and ev["exist"].all()
    rel = np.abs(ev["mass"] - ref.mass) / ref.mass
    assert rel.max() < mass_tol, rel.max()
    assert abs(ev["mass"].sum() - S.area_scalar) < sum_tol * S.area_scalar
    assert np.abs(cen - ref.centroid).max() < cen_tol
    med, p99, mx = shipped.centroid_agreement(cen, S.centroids)
    assert med < 1e-3 and p99 < 1e-2 and mx < 0.2, (med, p99, mx)
