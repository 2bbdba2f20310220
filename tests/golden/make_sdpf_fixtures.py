"""Writes the sdpf fixtures (n = 64) from the reference's own data files.

Inputs (read from the reference checkout, path given as argv[1], default
/root/reference):
  data/Ca/Ca44_HFB_SDPFMU_sdpf_hwHO11_s000_beta+000gam00np00cf000.wf  (44Ca, Z, N valence 12, 16)
  data/Ne/Ne20_HFB_SDPFMU_sdpf_hwHO15_s000_beta+000gam00np00cf000.wf  (20Ne, Z, N valence 2, 2)
  Codes/int/SDPFMU/IMSRG_s000_SDPFMUsdpf_hwHO11_m1m2m3m4.dat          (m-scheme V list, 51048 records)
  Codes/int/SDPFMU/SM_SDPFMU_ME1B.dat                                 (single-particle energies)
Output (this directory): sdpf/sdpf_inputs.npz with U[2,64,64], V[2,64,64]
(as read, U[i,k]; Ca44 first), ide[2,2] (the .wf headers), me[51048] and the
SPE file text.  The numbers are copied, not recomputed.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__)) + "/sdpf"
STATES = ["data/Ca/Ca44_HFB_SDPFMU_sdpf_hwHO11_s000_beta+000gam00np00cf000.wf",
          "data/Ne/Ne20_HFB_SDPFMU_sdpf_hwHO15_s000_beta+000gam00np00cf000.wf"]


def read_wf(path, n=64):
    with open(path) as f:
        ide = [int(x) for x in f.readline().split()]
        d = np.loadtxt(f)
    return d[:, 0].reshape(n, n), d[:, 1].reshape(n, n), ide


def main():
    os.makedirs(OUT, exist_ok=True)
    Us, Vs, ides = [], [], []
    for s in STATES:
        U, V, ide = read_wf(f"{REF}/{s}")
        Us.append(U)
        Vs.append(V)
        ides.append(ide)
    raw = np.fromfile(f"{REF}/Codes/int/SDPFMU/IMSRG_s000_SDPFMUsdpf_hwHO11_m1m2m3m4.dat",
                      dtype=np.dtype([("h", "<i4"), ("v", "<f8"), ("t", "<i4")]))
    assert np.all(raw["h"] == 8) and np.all(raw["t"] == 8)
    spe = open(f"{REF}/Codes/int/SDPFMU/SM_SDPFMU_ME1B.dat").read()
    np.savez_compressed(f"{OUT}/sdpf_inputs.npz", U=np.array(Us), V=np.array(Vs), ide=np.array(ides),
                        me=raw["v"].astype(np.float64), spe=np.array(spe))


if __name__ == "__main__":
    main()
