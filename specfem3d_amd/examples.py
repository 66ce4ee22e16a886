"""Reference examples packed as fixtures (inputs + reference seismograms)."""
from __future__ import annotations

import io
import os

import numpy as np

from . import sources as S
from .mesh.external import build_from_arrays

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def stacey5_example(path: str | None = None):
    """(db, sources, receivers, ref_seis, DT) of the reference example
    homogeneous_halfspace_HEX8_elastic_absorbing_Stacey_5sides (CMT with error-function STF)."""
    z = np.load(path or os.path.join(GOLDEN, "stacey5_example.npz"), allow_pickle=False)
    dt = float(z["DT"])
    db = build_from_arrays({k: z[k] for k in z.files}, dt)
    cmts = S.read_cmtsolution(io.StringIO(str(z["cmtsolution"])), dt)
    # this example's REF_SEIS come from an NPROC=4 run whose partition put the corner source
    # in the element below it
    src = S.setup_cmt_sources(db, cmts, dt, ricker=False, tie_break="lower")
    st = S.read_stations(io.StringIO(str(z["stations"])))
    rec = S.setup_receivers(db, st)
    return db, src, rec, z["ref_seis"], dt
