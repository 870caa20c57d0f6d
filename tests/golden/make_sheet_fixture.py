"""Writes tests/golden/sheet_ap.npz with the FP64 oracle (run on the CPU).

Workload: the 2D Aliev-Panfilov sheet, 128 x 128 voxels at 1 mm (tool/create_geometry.py:41-59),
reference default parameters (k 8, a 0.15, eps0 0.002, mu1 0.2, mu2 0.3, c = 1.6 x 6),
dt = 0.5/6, 2000 steps, central 5 x 5 stimulus for 2 model-time units.

Stored: u after every step at 16 seeded probe voxels (2000 x 16, float64), the first
activation step of every voxel (int16), the 1 kHz sample times, and u at the last sample.

    python tests/golden/make_sheet_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sim_oracle  # noqa: E402
from pyheartsim_amd import configs  # noqa: E402


def main():
    w = configs.sheet_ap()
    P = configs.build_P_2d(w["geometry"], w["sim"], w["hm"])
    r = sim_oracle.simulate(P.shape[0], P, w["geometry"], w["sim"], w["arr"], probes=w["probes"])
    np.savez_compressed(os.path.join(HERE, "sheet_ap.npz"), probes=w["probes"], probe_u=r["probe_u"],
                        lat0=r["lat"][0].astype(np.int16), physical_time=r["physical_time"],
                        u_last=r["u"][-1])


if __name__ == "__main__":
    main()
