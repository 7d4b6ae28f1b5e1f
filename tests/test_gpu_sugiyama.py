"""The Sugiyama benchmark (row f2) on the device: the lid-driven cavity
with a deformable disk at N = 128 run with the driver's adaptive dt to
T = 4 (about 16,000 steps), and the computed centroid path compared with
the digitized path of Sugiyama et al.: every computed point lies close to
the reference curve (the full T = 20 runs of N = 128 ... 1024 stay within
0.015 of it, mean about 3e-3), and the ACDI phase volume is conserved."""
import numpy as np
import pytest

from eulerianfsi_amd import sugiyama

pytestmark = pytest.mark.gpu


def test_sugiyama_path_follows_the_digitized_reference(tmp_path):
    res = sugiyama.run(128, str(tmp_path / "sugiyama"), T=4.0, progress_every=1.0)
    cmp = res["path_vs_digitized"]
    print(res, flush=True)
    assert abs(res["T"] - 4.0) < 1e-9
    traj = np.loadtxt(tmp_path / "sugiyama" / "centroid_trajectory.csv", delimiter=",", skiprows=1)
    # the disk has moved along the lid: more than a quarter of the way round
    assert np.hypot(traj[-1, 1] - traj[0, 1], traj[-1, 2] - traj[0, 2]) > 0.15
    assert cmp["max_to_reference"] <= 0.02
    assert cmp["mean_to_reference"] <= 5e-3
    lo, hi = res["phase_integral_ratio_range"]
    assert abs(lo - 1.0) <= 1e-10 and abs(hi - 1.0) <= 1e-10
