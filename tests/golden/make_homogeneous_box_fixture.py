"""Pack the reference solver's seismograms of the 20^3 homogeneous box into a small fixture.

Source data: ``OUTPUT_FILES/DB.X??.HX?.semd`` written by SPECFEM3D Cartesian 4.1.1's CPU solver
(``xspecfem3D``, serial, NPROC=1) for the 20^3-element homogeneous box (DATA/Par_file: NSTEP=2000,
DT=0.01, Ricker f0=1 Hz CMT at (5000, 5000, -5000) m, 10 surface stations), once in the FP32
build and once in the FP64 (--enable-double-precision) build. Those runs were made in this
container before this package existed (directories /tmp/cfg1 and /tmp/cfg1_dp); this script
only converts their text output:

    python tests/golden/make_homogeneous_box_fixture.py /tmp/cfg1/OUTPUT_FILES /tmp/cfg1_dp/OUTPUT_FILES

Output: tests/golden/homogeneous_box_reference_seismograms.npz with float32 arrays ``fp32`` and ``fp64``
of shape (2000 steps, 10 receivers, 3 components x/y/z) and the time column ``t``.
"""
import os
import sys

import numpy as np


def load(d):
    out = np.zeros((2000, 10, 3), np.float64)
    t = None
    for r in range(10):
        for c, ch in enumerate("XYZ"):
            a = np.loadtxt(os.path.join(d, f"DB.X{r + 1:02d}.HX{ch}.semd"))
            out[:, r, c] = a[:, 1]
            t = a[:, 0]
    return out, t


if __name__ == "__main__":
    d32, d64 = sys.argv[1], sys.argv[2]
    s32, t = load(d32)
    s64, _ = load(d64)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "homogeneous_box_reference_seismograms.npz"), fp32=s32.astype(np.float32),
                        fp64=s64.astype(np.float32), t=t.astype(np.float32))
    print("wrote", os.path.join(here, "homogeneous_box_reference_seismograms.npz"))
