"""The C/OpenMP restatement (oracle/cpu_oracle.c) against the NumPy oracle."""
import numpy as np

from oracle import cpu_oracle
from tests.helpers import run_oracle, small_shell


def test_c_restatement_matches_numpy_oracle():
    w = small_shell(t_final=30.0, n_scar=1)
    ref = run_oracle(w, max_beats=7)
    tight = cpu_oracle.CpuSim(w["P_2d"], w["geometry"], w["sim"], w["arr"], rtol=1e-12, atol=0.0)
    tight.run(600, threads=4)
    assert np.max(np.abs(tight.u - ref["u_final"])) < 1e-10
    assert np.max(np.abs(tight.h - ref["h_final"])) < 1e-12
    loose = cpu_oracle.CpuSim(w["P_2d"], w["geometry"], w["sim"], w["arr"])
    assert loose.run(600, threads=4) == 0
    # the reference stop rule (||r|| < max(1e-5 ||b||, 1e-6)) is loose but bounded
    assert np.max(np.abs(loose.u - ref["u_final"])) < 1e-2
