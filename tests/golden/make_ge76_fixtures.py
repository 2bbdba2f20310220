"""Writes the Ge76 golden fixtures from the reference's own data files.

Inputs (read from the reference checkout, path given as argv[1], default
/root/reference):
  data/Ge/Ge76_HFB_GCN2850_pf5g9_hwHO09_s000_beta{-020,-010,+000,+010,+020}gam00np00cf000.wf
  Codes/int/GCN2850/IMSRG_s000_GCN2850pf5g9_hwHO10_m1m2m3m4.dat   (m-scheme V list)
  Codes/int/GCN2850/SM_GCN2850_ME1B.dat                            (single-particle energies)
  data/Ge/kern.1D.05.01.16.01.GCN2850_pf5g9_hwHO09_s000_<q>_<q'>.elem  (expected kernels)
Outputs (this directory):
  ge76_inputs.npz  U[5,44,44], V[5,44,44] (as read, U[i,k]), beta2[5], me[10330], spe file lines
  ge76_kernels.json  {"<beta_q>,<beta_q'>": {"J": [N_re,N_im,H_re,H_im,nn_re,nn_im,zz_re,zz_im,J2_re,J2_im]}}
The numbers are copied, not recomputed: they are the reference run's inputs
and outputs (8 significant digits for the kernels).
"""
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__)) + "/ge76"
BETAS = [-0.2, -0.1, 0.0, 0.1, 0.2]
TAG = {-0.2: "-020", -0.1: "-010", 0.0: "+000", 0.1: "+010", 0.2: "+020"}
KTAG = {-0.2: "-02000", -0.1: "-01000", 0.0: "+00000", 0.1: "+01000", 0.2: "+02000"}


def read_wf(path, n=44):
    with open(path) as f:
        f.readline()
        d = np.loadtxt(f)
    return d[:, 0].reshape(n, n), d[:, 1].reshape(n, n)


def main():
    os.makedirs(OUT, exist_ok=True)
    Us, Vs = [], []
    for b in BETAS:
        U, V = read_wf(f"{REF}/data/Ge/Ge76_HFB_GCN2850_pf5g9_hwHO09_s000_beta{TAG[b]}gam00np00cf000.wf")
        Us.append(U)
        Vs.append(V)
    raw = np.fromfile(f"{REF}/Codes/int/GCN2850/IMSRG_s000_GCN2850pf5g9_hwHO10_m1m2m3m4.dat",
                      dtype=np.dtype([("h", "<i4"), ("v", "<f8"), ("t", "<i4")]))
    assert np.all(raw["h"] == 8) and np.all(raw["t"] == 8)
    spe = open(f"{REF}/Codes/int/GCN2850/SM_GCN2850_ME1B.dat").read()
    np.savez_compressed(f"{OUT}/ge76_inputs.npz", U=np.array(Us), V=np.array(Vs),
                        beta2=np.array(BETAS), me=raw["v"].astype(np.float64),
                        spe=np.array(spe))
    kern = {}
    for b0 in BETAS:
        for b1 in BETAS:
            path = (f"{REF}/data/Ge/kern.1D.05.01.16.01.GCN2850_pf5g9_hwHO09_s000_"
                    f"{KTAG[b0]}.000cf000_{KTAG[b1]}.000cf000.elem")
            lines = open(path).read().split("\n")
            rec = {}
            for blk in range(4):
                L = lines[blk * 10: blk * 10 + 10]
                J = int(L[0])
                vals = []
                for ln in (L[3], L[4]):
                    vals += [float(ln[i:i + 15]) for i in range(0, 60, 15)]
                vals += [float(L[5][i:i + 15]) for i in range(0, 30, 15)]
                rec[str(J)] = vals
            kern[f"{b0:+.1f},{b1:+.1f}"] = rec
    json.dump(kern, open(f"{OUT}/ge76_kernels.json", "w"), indent=1)


if __name__ == "__main__":
    main()
