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
    worst = 0.0
    for k, (p, wk) in enumerate(zip(probs, ws)):
        mr, zeros = _confirm(p, wk)
        worst = max(worst, mr)
        assert mr < 1e-5 + 1e-6 and zeros == 0, (k, p.n, mr, zeros)
    print(f"batch64: {rep[0]['seconds']:.2f} s, oracle worst max-rel {worst:.3e}", flush=True)
