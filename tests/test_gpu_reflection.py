"""GPU parity of the t >= 1 path (Fresnel-weighted masses, L-BFGS, Anderson acceleration)
against the FP64 CPU oracle, through the C ABI (cot_centroids, cot_solve_reflection)."""
import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BETA = -0.5  # aa_fixed_para for the unit-square scale of these problems


@pytest.fixture(scope="module")
def lens_case():
    return W.lens_problem()


def _batch(p):
    from caustic_amd.ot import OTBatch

    return OTBatch([p], precision="fp64")


def test_geometric_centroids_match_oracle(gpu_available, lens_case):
    """The Fresnel masses are taken at the cells' geometric centroids (compute_center with
    weight_false, basic_function.cuh:1056-1191)."""
    p, _ = lens_case
    print("[case] geometric centroids, lens problem", flush=True)
    rng = np.random.default_rng(11)
    w = rng.normal(0.0, 2e-4, p.n)
    orc = O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain)
    with _batch(p) as b:
        cen, ex = b.centroids(w, weighted=False)
    assert np.array_equal(ex, orc.exist.astype(bool))
    np.testing.assert_allclose(cen[ex], orc.geo[ex], rtol=0, atol=1e-13)


@pytest.mark.parametrize("stage_a", [True, False])
def test_reflection_iterates_follow_the_oracle(gpu_available, lens_case, stage_a):
    """Three iterations of each stage from zero weights: the device's weights are the
    oracle's to FP64 summation-order accuracy (the solve stops at the iteration limit)."""
    p, lens = lens_case
    print(f"[case] reflection iterates, stage_a={stage_a}", flush=True)
    kw = dict(max_iter=3, tol=1e-14, aa_history_size=4, aa_fixed_para=BETA,
              if_aa_op_initial=stage_a)
    with _batch(p) as b:
        w, rep = b.solve_reflection(np.zeros(p.n), lens, history=7, **kw)
    orc = O.reflection_solve(p.sites, np.zeros(p.n), p.rho, p.proj_domain, lens, m=7,
                             total_iter=3, tol=1e-14, aa_history_size=4, aa_fixed_para=BETA,
                             if_aa_op_initial=stage_a, max_vertices=24)
    assert not rep["converged"]
    assert rep["iterations"] == (6 if stage_a else 3)
    scale = np.abs(orc.w).max()
    assert scale > 0
    err = np.abs(w - orc.w).max() / scale
    assert err < 1e-9, f"weights differ by {err:.3e} of max |w|"


def test_reflection_solve_converges_to_the_oracle_residual(gpu_available, lens_case):
    """Both stages to |mass - nu|_2 < 1e-6, nu the Fresnel masses of the final cells; the FP64
    oracle, evaluated at the device's weights, confirms the residual."""
    p, lens = lens_case
    print("[case] reflection solve to 1e-6", flush=True)
    tol = 1e-6
    with _batch(p) as b:
        w, rep = b.solve_reflection(np.zeros(p.n), lens, history=7, max_iter=2000, tol=tol,
                                    aa_history_size=7, aa_fixed_para=BETA)
    assert rep["converged"], rep
    ev = O.evaluate(p.sites, w, np.zeros(p.n), p.rho, p.proj_domain)
    nu = O.fresnel_masses(ev, lens)
    res = np.linalg.norm(ev.mass - nu)
    assert res < tol * (1 + 1e-6), f"oracle residual {res:.3e}"
    assert ev.exist.all()
    # the Fresnel masses differ from the clip areas' proportions by ~20%: the coupling of the
    # masses to the cells is exercised
    share = lens["clip_area"] / lens["clip_area"].sum()
    assert np.abs(nu - share).max() > 0.05 * share.max()
