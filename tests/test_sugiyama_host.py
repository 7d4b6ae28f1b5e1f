"""Host side of the Sugiyama comparison: point-to-polyline distances and the
curve comparison on shapes whose distances are known; the digitized path
loads as 221 points inside the unit square.  CPU only."""
import math

import numpy as np

from eulerianfsi_amd import sugiyama as S


def test_distances_to_polyline():
    poly = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    pts = np.array([[0.5, 0.3], [2.0, 0.5], [-1.0, 0.0], [1.2, 1.5]])
    d = S.distances_to_polyline(pts, poly)
    assert np.allclose(d, [0.3, 1.0, 1.0, math.hypot(0.2, 0.5)], atol=1e-15)


def test_compare_paths_shifted_circle():
    th = np.linspace(0.0, 2.0 * math.pi, 2001)
    ref = np.c_[0.5 + 0.2 * np.cos(th), 0.5 + 0.2 * np.sin(th)]
    # every 8th of 2001 points keeps both ends; the largest gap is the chord
    # sagitta r (1 - cos(4 dtheta))
    same = S.compare_paths(ref[::8], ref)
    assert abs(same["hausdorff"] - 0.2 * (1.0 - math.cos(4.0 * (th[1] - th[0])))) < 1e-9
    bigger = np.c_[0.5 + 0.23 * np.cos(th), 0.5 + 0.23 * np.sin(th)]
    c = S.compare_paths(bigger, ref)
    assert abs(c["max_to_reference"] - 0.03) < 1e-5 and abs(c["mean_from_reference"] - 0.03) < 1e-5
    # a computed path over the upper half only: the lowest reference point
    # (0.5, 0.3) is r sqrt(2) from the nearest end of the half circle
    half = ref[: len(ref) // 2 + 1]
    c = S.compare_paths(half, ref)
    assert c["max_to_reference"] < 1e-12 and abs(c["max_from_reference"] - 0.2 * math.sqrt(2.0)) < 1e-9


def test_digitized_path_loads():
    p = S.load_digitized()
    assert p.shape == (221, 2)
    assert np.all((p > 0.0) & (p < 1.0))
    assert abs(p[0, 0] - 0.6) < 1e-3 and abs(p[0, 1] - 0.5) < 2e-3  # starts at the disk centre
