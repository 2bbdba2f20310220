"""Valence spaces and the m-scheme single-particle basis.

The ordering follows the reference exactly, because every array that crosses
the C-ABI (U, V, the interaction list, the one-body matrix) is indexed by it:

* orbits are taken in the order of the ``.val`` file
  (``Codes/PNAMP/PNAMP_Base.f90:37-60``);
* inside an orbit, m runs j, j-1, ..., -j
  (``Codes/GCM/HFBA_Hamiltonian.f90:44-68``: ``mjang = twoj - mjinc``);
* protons occupy indices 1..n/2 (t = -1) and neutrons n/2+1..n (t = +1)
  (``HFBA_Hamiltonian.f90:50-64``), which is what ``ZN_0``
  (``Codes/PNAMP/PNAMP_ME1B.f90:218-237``) and ``UVphiomega``
  (``Codes/PNAMP/PNAMP_Proj.f90:432-439``) assume.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# (n, l, 2j) per orbit in .val order; n is the radial quantum number as
# written in the .val files (0-based), the basis stores n+1 like tnljm%n.
SPACES: dict[str, list[tuple[int, int, int]]] = {
    # 0d5/2, 1s1/2, 0d3/2
    "sd": [(0, 2, 5), (1, 0, 1), (0, 2, 3)],
    # 0f7/2, 1p3/2, 0f5/2, 1p1/2  (Codes/val/pf.val)
    "pf": [(0, 3, 7), (1, 1, 3), (0, 3, 5), (1, 1, 1)],
    # 0f5/2, 1p3/2, 1p1/2, 0g9/2  (Codes/val/pf5g9.val)
    "f5pg9": [(0, 3, 5), (1, 1, 3), (1, 1, 1), (0, 4, 9)],
    # 0d5/2, 0d3/2, 1s1/2, 0f7/2, 0f5/2, 1p3/2, 1p1/2  (Codes/val/sdpf.val): n = 64
    "sdpf": [(0, 2, 5), (0, 2, 3), (1, 0, 1), (0, 3, 7), (0, 3, 5), (1, 1, 3), (1, 1, 1)],
    # tiny spaces used by the exact Fock-space checks
    "p": [(0, 1, 3), (0, 1, 1)],
    "p32": [(0, 1, 3)],
}
SPACES["pf5g9"] = SPACES["f5pg9"]


@dataclass
class Basis:
    """m-scheme basis of one valence space, both species.

    Arrays have length n (= NLEV), protons first.
    """

    space: str
    orbits: list[tuple[int, int, int]]
    t: np.ndarray = field(repr=False)      # -1 proton, +1 neutron
    nq: np.ndarray = field(repr=False)     # radial n+1 (tnljm%n)
    l: np.ndarray = field(repr=False)
    twoj: np.ndarray = field(repr=False)
    twom: np.ndarray = field(repr=False)
    orbit: np.ndarray = field(repr=False)  # orbit index 0..norb-1 (same for both species)

    @property
    def n(self) -> int:
        return int(self.t.size)

    @property
    def lj(self) -> np.ndarray:
        """tnljm%lj = (2l + 2j - 1)/2 with j doubled (HFBA_Hamiltonian.f90:88)."""
        return (2 * self.l + self.twoj - 1) // 2

    def blocks(self) -> list[tuple[int, int, int]]:
        """Contiguous rotation blocks (offset, size, 2j): R is non-zero only
        between m-states of the same (n, l, j, t) (PNAMP_Proj.f90:120-127)."""
        out = []
        k = 0
        while k < self.n:
            size = int(self.twoj[k]) + 1
            out.append((k, size, int(self.twoj[k])))
            k += size
        return out

    def as_int32(self):
        return tuple(np.ascontiguousarray(a, dtype=np.int32)
                     for a in (self.t, self.nq, self.l, self.twoj, self.twom))


def make_basis(space: str, orbits: list[tuple[int, int, int]] | None = None) -> Basis:
    """Basis of a named space, or of ``orbits`` ((n, l, 2j) per orbit in .val
    order) for any other valence space, which ``space`` then only names."""
    orbits = SPACES[space] if orbits is None else [tuple(int(x) for x in o) for o in orbits]
    t, nq, l, tj, tm, orb = [], [], [], [], [], []
    for species in (-1, 1):
        for io, (n, ll, twoj) in enumerate(orbits):
            for twom in range(twoj, -twoj - 1, -2):
                t.append(species)
                nq.append(n + 1)
                l.append(ll)
                tj.append(twoj)
                tm.append(twom)
                orb.append(io)
    a = lambda x: np.asarray(x, dtype=np.int32)
    return Basis(space, list(orbits), a(t), a(nq), a(l), a(tj), a(tm), a(orb))
