"""Loader of the Ti48 fixtures (tests/golden/ti48, written by
tests/golden/make_ti48_fixtures.py from the reference's data/Ti files):
eight axial KB3G HFB states of 48Ti in the pf space (n = 40, Z, N valence
2, 6 over the 40Ca core) with proton-neutron pairing."""
import os

import numpy as np

from sccm_amd.basis import make_basis
from sccm_amd.interaction import Interaction, list_indices, parse_spe
from sccm_amd.mesh import Mesh, gauleg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ti48")
ZVAL, NVAL = 2, 6


def load():
    d = np.load(os.path.join(HERE, "ti48_inputs.npz"))
    b = make_basis("pf")
    idx = list_indices(b)
    t1b, e0 = parse_spe(str(d["spe"]), b)
    inter = Interaction(idx[:, 0].copy(), idx[:, 1].copy(), idx[:, 2].copy(), idx[:, 3].copy(),
                        d["me"].copy(), t1b, e0)
    return b, inter, d["U"], d["V"], [float(x) for x in d["beta2"]]


def axial_mesh(nbet=16, nphi=5, jmax=6):
    """Axial projection mesh: one alpha and gamma point, nbet Gauss-Legendre
    points in cos(beta) over [-1, 1], Fomenko nphi per species (the run
    card's PNP = 5, Codes/exec/PNAMP.sh:25)."""
    x, w = gauleg(-1.0, 1.0, nbet)
    return Mesh(np.zeros(1), np.ones(1), np.arccos(x), w, np.zeros(1), np.ones(1), nphi, jmax, 0, 0)
