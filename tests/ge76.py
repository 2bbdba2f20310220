"""Loader of the Ge76 golden fixtures (tests/golden/ge76, written by
tests/golden/make_ge76_fixtures.py from the reference's data/Ge files)."""
import json
import os

import numpy as np

from sccm_amd.basis import make_basis
from sccm_amd.interaction import Interaction, list_indices, parse_spe
from sccm_amd.mesh import Mesh, gauleg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ge76")
ZVAL, NVAL, NCORE = 4, 16, 28


def load():
    d = np.load(os.path.join(HERE, "ge76_inputs.npz"))
    b = make_basis("f5pg9")
    idx = list_indices(b)
    t1b, e0 = parse_spe(str(d["spe"]), b)
    inter = Interaction(idx[:, 0].copy(), idx[:, 1].copy(), idx[:, 2].copy(), idx[:, 3].copy(),
                        d["me"].copy(), t1b, e0)
    kern = json.load(open(os.path.join(HERE, "ge76_kernels.json")))
    return b, inter, d["U"], d["V"], [float(x) for x in d["beta2"]], kern


def golden_mesh(jmax=6):
    """The mesh the Ge76 kernels were produced on (axial, 'kern.1D.05.01.16.01'):
    one alpha and gamma point, 16 Gauss-Legendre points in cos(beta) of which the
    8 with cos(beta) < 0 are used with doubled weight (the reference's
    reflection-symmetric beta loop, Ncosbet = 2 in Codes/PNAMP/PNAMP_Kernel1D.f90:64-66,195,498),
    5 Fomenko points per species, Pfaffian norm."""
    x, w = gauleg(-1.0, 1.0, 16)
    sel = x < 0
    return Mesh(np.zeros(1), np.ones(1), np.arccos(x[sel]), 2 * w[sel], np.zeros(1), np.ones(1),
                5, jmax, 0, 0)
