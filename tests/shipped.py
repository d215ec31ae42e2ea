"""The reference's shipped run (N = 160,978, 640 x 640 ustc_centerL target) rebuilt from the
fixture tests/golden/ustc_centerL_t0.npz: the seed files spline_irra.py would write
(utils/data_utils.py:374-435, utils/basic_function.py:774-780: "%f" text), the normalised
target, and the reference's t = 0 outputs (6-digit weights and centroids)."""
import json
import os

import numpy as np

from caustic_amd import workloads as W

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "ustc_centerL_t0.npz")


def _as_printed(a):
    """Values as np.savetxt(fmt="%f") writes them and the reference's reader parses them."""
    return np.array([float("%f" % v) for v in np.ravel(a)]).reshape(np.shape(a))


def _from_text(a32):
    return np.array([float("%g" % v) for v in np.ravel(a32)]).reshape(np.shape(a32))


class Shipped:
    def __init__(self):
        z = np.load(FIXTURE, allow_pickle=False)
        self.settings = json.loads(str(z["settings"]))
        self.text = z["picture"].astype(np.float64)       # picture text matrix, rows as in the file
        # the 6-digit decimals of the shipped text files, exactly as the reference's reader
        # parses them ("%g" of the stored float32 gives the text back)
        self.weights = _from_text(z["weights"])       # t = 0 final_proj_point_weights.txt
        self.centroids = _from_text(z["centroids"])   # t = 0 out.txt
        self.in_sha256 = str(z["in_sha256"])
        self.in_lines = int(z["in_lines"])
        s = self.settings
        self.sample_domain = tuple(float(v) for v in s["sample_domain"])
        self.proj_domain = tuple(float(v) for v in s["proj_domain"])
        self.area_scalar = ((self.sample_domain[1] - self.sample_domain[0]) *
                            (self.sample_domain[3] - self.sample_domain[2]))
        lattice = W.triangular_lattice(s["point_num"], self.sample_domain)
        self.points = _as_printed(lattice)
        self.proj_xy = _as_printed(W.similarity_map(self.points, self.sample_domain,
                                                    self.proj_domain))
        pic = W.picture_from_text_rows(self.text)
        self.rho = pic / pic.sum() * self.area_scalar   # ot_op.cuh:81-88,154
        self.n = self.points.shape[0]


_CACHE = {}


def load() -> Shipped:
    if "s" not in _CACHE:
        _CACHE["s"] = Shipped()
    return _CACHE["s"]


def centroid_agreement(cen, ref):
    """(median, 99th percentile, max) of the per-cell max-norm distance, in domain units."""
    d = np.abs(np.asarray(cen) - np.asarray(ref)).max(axis=1)
    return float(np.median(d)), float(np.percentile(d, 99)), float(d.max())
