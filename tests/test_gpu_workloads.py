"""The benchmark configurations solved end to end on the GPU (cot_solve_ladder, mixed
precision) and confirmed problem by problem by the FP64 oracle at the returned weights."""
import numpy as np
import pytest

from caustic_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _confirm(p, w):
    ev = O.evaluate(p.sites, w, p.nu, p.rho, p.proj_domain)
    return O.max_rel_error(ev.mass, p.nu), int(np.count_nonzero(ev.mass == 0.0))


def test_large_shapes_solve_confirmed_by_oracle(gpu_available):
    """N = 65,536 jittered sites, 1024^2 disks/rectangles target (100:1): the ladder solve to
    max-rel 1e-5 in mixed precision; the FP64 oracle's masses at the returned weights meet
    1e-5 to within the FP32 geometry's 1e-6."""
    from caustic_amd.ot import OTBatch

    print("[case] large_shapes N=65536 ladder solve", flush=True)
    p = W.large_shapes()
    with OTBatch([p], precision="mixed") as b:
        w, rep = b.solve_ladder(tol=1e-5, verbose=1)
    assert rep[0]["converged"], rep
    mr, zeros = _confirm(p, w)
    print(f"large_shapes: {rep[0]}, oracle max-rel {mr:.3e}", flush=True)
    assert mr < 1e-5 + 1e-6 and zeros == 0


def test_batch64_solve_confirmed_by_oracle(gpu_available):
    """The batched configuration: 64 problems, N_k = 2,048 .. 32,768 (log-uniform), 512^2
    Gaussian-mixture targets, in one context; every problem reaches max-rel 1e-5, confirmed by
    the FP64 oracle at its weights."""
    from caustic_amd.ot import OTBatch

    print("[case] batch64 ladder solve", flush=True)
    probs = W.batched_uneven()
    sizes = [p.n for p in probs]
    assert len(probs) == 64 and min(sizes) >= 2048 and max(sizes) <= 32768 + 2 * 182
    with OTBatch(probs, precision="mixed") as b:
        w, rep = b.solve_ladder(tol=1e-5, verbose=1)
        ws = b.split(w)
    assert all(r["converged"] for r in rep), [r for r in rep if not r["converged"]][:3]
    # every problem's Newton steps converge: when the batch's conjugate gradients stopped on the
    # summed residual at the loosest problem's tolerance, small problems' steps failed and
    # L-BFGS took them the rest of the way (up to 1,051 iterations at the final level, against
    # at most 87 with per-problem stopping)
    its = [r["iterations"] for r in rep]
    assert max(its) < 300, sorted(its)[-5:]
    worst = 0.0
    for k, (p, wk) in enumerate(zip(probs, ws)):
        mr, zeros = _confirm(p, wk)
        worst = max(worst, mr)
        assert mr < 1e-5 + 1e-6 and zeros == 0, (k, p.n, mr, zeros)
    print(f"batch64: {rep[0]['seconds']:.2f} s, oracle worst max-rel {worst:.3e}", flush=True)


@pytest.mark.parametrize("tuned", [False, True])
def test_curves262k_ladder_solve_confirmed_by_oracle(gpu_available, tuned):
    """N = 262,144 sites on the 30-degree tilted source plane, 2048^2 Bezier-stroke target: the
    library's ladder (16 .. 256 levels) to max-rel 1e-5, with the default launch shapes and with
    the shapes the bench's autotuner picked on MI355X for this workload (under them the final
    level's first evaluation, at the 256^2 level's transferred weights, failed the conformity
    check at the largest ring; the solve must still report the verified result its accepted
    trials reach).  Confirmed by the FP64 oracle."""
    from caustic_amd.ot import OTBatch

    print(f"[case] curves262k ladder solve tuned={tuned}", flush=True)
    p = W.tilted_curves()
    shape = dict(int_lanes=8, repair_blocks=512, sites_per_bucket=1.0, full_pass_ring=2,
                 pass_mode=0)
    lanes = [dict(shape, repair_blocks=256, pass_mode=1)] * 2 + [
        dict(shape, pass_mode=1), dict(shape, full_pass_ring=1), dict(shape)]
    final = dict(shape, int_lanes=4)
    with OTBatch([p], precision="mixed", tuning=final if tuned else None) as b:
        lad = b.ladder_options(level_tunings=lanes) if tuned else None
        w, rep = b.solve_ladder(tol=1e-5, verbose=1, ladder=lad)
    assert rep[0]["converged"] and rep[0]["status"] == 0, rep
    mr, zeros = _confirm(p, w)
    print(f"curves262k: {rep[0]}, oracle max-rel {mr:.3e}", flush=True)
    assert mr < 1e-5 + 1e-6 and zeros == 0
