"""The Sugiyama lid-driven cavity benchmark on the device, with the geometric
comparison of the disk's centroid path against the digitized reference.

Configuration of scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl
(BenchmarkConfig :40-128, build_models :386-420): unit square, lid speed 1,
nu = 0.01, neo-Hookean disk (G = 0.1) of radius 0.2 centred at (0.6, 0.5),
ACDI with epsilon = 0.51 h and a tanh start of width 2 epsilon, fixed Gamma,
T = 20, the driver's adaptive dt.  The digitized path
(data/sugiyama_lid_driven_cavity_1024_digitized.csv, a copy of the data file
data/benchmarks/sugiyama_lid_driven_cavity_1024_digitized.csv of the
reference) holds centroid positions without time stamps, so the comparison
is between curves, as the driver's own plot is (:1940-1950): the distance
from each computed centroid to the digitized polyline and back.

    python -m eulerianfsi_amd.sugiyama --N 512 --out DIR
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

from . import diagnostics as D
from . import driver, setups

DIGITIZED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                         "sugiyama_lid_driven_cavity_1024_digitized.csv")


def load_digitized(path: str = DIGITIZED) -> np.ndarray:
    return np.loadtxt(path, delimiter=",", skiprows=1)


def distances_to_polyline(points: np.ndarray, poly: np.ndarray) -> np.ndarray:
    """Distance of each point to the polyline through poly's vertices."""
    a, b = poly[:-1], poly[1:]
    ab = b - a
    L2 = np.maximum(np.sum(ab * ab, axis=1), 1e-300)
    d = np.empty(len(points))
    for k, p in enumerate(points):
        t = np.clip(np.sum((p - a) * ab, axis=1) / L2, 0.0, 1.0)
        q = a + t[:, None] * ab
        d[k] = float(np.min(np.hypot(*(p - q).T)))
    return d


def compare_paths(path: np.ndarray, reference: np.ndarray) -> dict:
    """Distances between a computed centroid path and the reference curve:
    computed -> reference (max, mean) and reference -> computed (max, mean);
    the symmetric Hausdorff distance is the larger maximum."""
    fwd = distances_to_polyline(path, reference)
    back = distances_to_polyline(reference, path)
    return dict(max_to_reference=float(fwd.max()), mean_to_reference=float(fwd.mean()),
                max_from_reference=float(back.max()), mean_from_reference=float(back.mean()),
                hausdorff=float(max(fwd.max(), back.max())))


def run(N: int, out: str, T: float = 20.0, method: str = "acdi", acdi_epsilon_over_h: float = 0.51,
        progress_every: float = 1.0) -> dict:
    case = setups.lid_driven_cavity(N, method, radius=0.2, center=(0.6, 0.5), shear_modulus=0.1, viscosity=0.01,
                                    acdi_eps_over_h=acdi_epsilon_over_h)
    cfg = D.BenchmarkConfig(T=T)
    h = driver.conservative_mesh_spacing(case.metrics)
    setup = driver.phase_field_setup(h, acdi_epsilon_over_h=acdi_epsilon_over_h)
    mark = [progress_every]

    def progress(step, t, dt):
        if t >= mark[0] - 1e-9:
            print(f"N={N} t={t:.2f} step={step} dt={dt:.3e}", flush=True)
            mark[0] += progress_every

    summary = driver.run_cavity(case, cfg, setup, out, radius=0.2, progress=progress)
    traj = np.loadtxt(summary.files["centroid_trajectory.csv"], delimiter=",", skiprows=1)
    cmp = compare_paths(traj[:, 1:3], load_digitized())
    result = dict(N=N, method=method, T=summary.time, steps=summary.steps, wall_seconds=summary.wall_seconds,
                  steps_per_second=summary.steps / summary.wall_seconds,
                  final_centroid=[float(traj[-1, 1]), float(traj[-1, 2])],
                  phase_integral_ratio_range=[float(traj[:, 4].min()), float(traj[:, 4].max())],
                  path_vs_digitized=cmp)
    with open(os.path.join(out, "comparison.json"), "w") as f:
        json.dump(result, f, indent=1)
    return result


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--method", default="acdi")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    print(json.dumps(run(args.N, args.out, args.T, args.method)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
