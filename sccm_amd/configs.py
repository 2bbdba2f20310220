"""Seeded synthetic workloads of the projected-kernel path.

sd_plain   sd space (n=24), Z,N valence 2,2, one HFB state (q = q'),
           Euler mesh alpha=gamma 8 GL points, beta 16, gauge 7 per species, J <= 8:
           1024 Euler x 49 gauge = 50,176 mesh points.
pf_gcm     pf space (n=40), Z,N valence 2,4, 20 HFB states on a (beta2, gamma)
           grid, all 210 pairs q <= q', alpha=gamma 10, beta 20, gauge 9, J <= 8,
           Hill-Wheeler norm cutoff 1e-6: 210 x 2000 x 81 = 34,020,000 mesh points.

Interaction: J-coupled TBMEs N(0, sigma) (1.5 MeV sd, 1.2 MeV pf) uncoupled to
the m-scheme list, single-particle energies uniform in [-5, 0] MeV.
States: expm of a seeded species-block-diagonal real Bogoliubov generator
(entries N(0, 0.3)) applied to a Slater state; the GCM states share a base
generator and differ by a smooth (beta2, gamma) offset.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from .basis import Basis, make_basis
from .interaction import Interaction, synthetic_interaction
from .mesh import Mesh, make_mesh
from .states import synthetic_state


@dataclass
class Workload:
    name: str
    basis: Basis
    inter: Interaction
    mesh: Mesh
    zval: int
    nval: int
    U: np.ndarray          # [nstates, n, n]
    V: np.ndarray
    pairs: np.ndarray      # [npairs, 2] (q, q') 0-based
    coords: np.ndarray     # [nstates, 2] (beta2, gamma_deg)
    hw_eps: float = 1e-6

    @property
    def n(self) -> int:
        return self.basis.n

    @property
    def mesh_points(self) -> int:
        return len(self.pairs) * self.mesh.n_points


@lru_cache(maxsize=None)
def _interaction(space: str, sigma: float, seed_tbme: int, seed_spe: int) -> Interaction:
    return synthetic_interaction(make_basis(space), sigma, seed_tbme, seed_spe)


def gcm_grid(nstates: int) -> np.ndarray:
    """(beta2, gamma) grid over beta2 in [0, 0.5], gamma in [0, 60 deg]."""
    if nstates == 1:
        return np.array([[0.25, 20.0]])
    nb = 5 if nstates % 5 == 0 else nstates
    ng = nstates // nb
    b2 = np.linspace(0.0, 0.5, nb)
    ga = np.linspace(0.0, 60.0, ng) if ng > 1 else np.array([20.0])
    return np.array([(b, g) for g in ga for b in b2])


def make_states(basis: Basis, zval: int, nval: int, coords: np.ndarray, seed: int = 2000,
                sigma: float = 0.3) -> tuple[np.ndarray, np.ndarray]:
    Us, Vs = [], []
    for q, (b2, g) in enumerate(coords):
        U, V = synthetic_state(basis, zval, nval, seed=seed, sigma=sigma, beta2=float(b2),
                               gamma_deg=float(g), offset_seed=seed + 1, noise=0.02 if q else 0.0)
        Us.append(U)
        Vs.append(V)
    return np.array(Us), np.array(Vs)


def sd_plain(jmax: int = 8, nalp: int = 8, nbet: int = 16, nphi: int = 7) -> Workload:
    b = make_basis("sd")
    inter = _interaction("sd", 1.5, 1001, 1002)
    mesh = make_mesh(nalp, nbet, nalp, nphi, jmax)
    coords = gcm_grid(1)
    U, V = make_states(b, 2, 2, coords)
    return Workload("sd_plain", b, inter, mesh, 2, 2, U, V, np.array([[0, 0]]), coords)


def pf_gcm(nstates: int = 20, jmax: int = 8, nalp: int = 10, nbet: int = 20, nphi: int = 9) -> Workload:
    b = make_basis("pf")
    inter = _interaction("pf", 1.2, 1011, 1012)
    mesh = make_mesh(nalp, nbet, nalp, nphi, jmax)
    coords = gcm_grid(nstates)
    U, V = make_states(b, 2, 4, coords)
    pairs = np.array([(q, q1) for q in range(nstates) for q1 in range(q, nstates)], dtype=np.int32)
    return Workload("pf_gcm", b, inter, mesh, 2, 4, U, V, pairs, coords)
