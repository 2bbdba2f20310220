"""Loader of the sdpf fixtures (tests/golden/sdpf, written by
tests/golden/make_sdpf_fixtures.py from the reference's data/Ca and data/Ne
files): the SDPFMU HFB states of 44Ca (Z, N valence 12, 16) and 20Ne (2, 2)
in the sd-pf space (n = 64, 16O core), both with proton-neutron pairing,
and the SDPFMU m-scheme list (51,048 records)."""
import os

import numpy as np

from sccm_amd.basis import make_basis
from sccm_amd.interaction import Interaction, list_indices, parse_spe

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sdpf")
NUCLEI = {"Ca44": (0, 12, 16), "Ne20": (1, 2, 2)}  # state index, Z, N valence


def load():
    d = np.load(os.path.join(HERE, "sdpf_inputs.npz"))
    b = make_basis("sdpf")
    idx = list_indices(b)
    t1b, e0 = parse_spe(str(d["spe"]), b)
    inter = Interaction(idx[:, 0].copy(), idx[:, 1].copy(), idx[:, 2].copy(), idx[:, 3].copy(),
                        d["me"].copy(), t1b, e0)
    return b, inter, d["U"], d["V"]
