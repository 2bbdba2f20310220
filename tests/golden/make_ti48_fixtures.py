"""Writes the Ti48 fixtures from the reference's own data files.

Inputs (read from the reference checkout, path given as argv[1], default
/root/reference):
  data/Ti/Ti48_HFB_KB3G_pf_hwHO11_s000_beta{-030,...,+040}gam00np00cf000.wf  (8 HFB states, pf, n = 40)
  Codes/int/KB3G/IMSRG_s000_KB3Gpf_hwHO11_m1m2m3m4.dat   (m-scheme V list, 14442 records)
  Codes/int/KB3G/SM_KB3G_ME1B.dat                        (single-particle energies)
Output (this directory): ti48/ti48_inputs.npz with U[8,40,40], V[8,40,40]
(as read, U[i,k]), ide[8,2] (the .wf header: 1 HFB, 0 HF per species),
beta2[8], me[14442] and the SPE file text.  The numbers are copied, not
recomputed.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__)) + "/ti48"
BETAS = [-0.3, -0.2, -0.1, 0.0, 0.1, 0.2, 0.3, 0.4]


def tag(b):
    return f"{'+' if b >= 0 else '-'}{int(round(abs(b) * 100)):03d}"


def read_wf(path, n=40):
    with open(path) as f:
        ide = [int(x) for x in f.readline().split()]
        d = np.loadtxt(f)
    return d[:, 0].reshape(n, n), d[:, 1].reshape(n, n), ide


def main():
    os.makedirs(OUT, exist_ok=True)
    Us, Vs, ides = [], [], []
    for b in BETAS:
        U, V, ide = read_wf(f"{REF}/data/Ti/Ti48_HFB_KB3G_pf_hwHO11_s000_beta{tag(b)}gam00np00cf000.wf")
        Us.append(U)
        Vs.append(V)
        ides.append(ide)
    raw = np.fromfile(f"{REF}/Codes/int/KB3G/IMSRG_s000_KB3Gpf_hwHO11_m1m2m3m4.dat",
                      dtype=np.dtype([("h", "<i4"), ("v", "<f8"), ("t", "<i4")]))
    assert np.all(raw["h"] == 8) and np.all(raw["t"] == 8)
    spe = open(f"{REF}/Codes/int/KB3G/SM_KB3G_ME1B.dat").read()
    np.savez_compressed(f"{OUT}/ti48_inputs.npz", U=np.array(Us), V=np.array(Vs), ide=np.array(ides),
                        beta2=np.array(BETAS), me=raw["v"].astype(np.float64), spe=np.array(spe))


if __name__ == "__main__":
    main()
