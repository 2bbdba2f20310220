"""Euler-angle and gauge-angle meshes.

Follows ``Gaussmesh`` (``Codes/GCM/PNAMP_Gaus.f90:27-46``): Gauss-Legendre
alpha, gamma on [0, 2 pi]; cos(beta) Gauss-Legendre on [-1, 1] with
beta = acos(x); a one-point direction is the angle 0 with weight 1.
The Euler points are enumerated beta outer, alpha middle, gamma inner, the
loop order of ``Kernel4Triaxial`` (``Codes/PNAMP/PNAMP_Kernel3D.f90:200-217``):
iang = (ib * nalp + ia) * ngam + ig.
Gauge angles are Fomenko points phi_k = 2 pi k / nphi, k = 0..nphi-1
(``PNAMP_Kernel3D.f90:262-276``).

``fold_mesh`` is the reduced mesh of the literal 3D mirror fold
(``PNAMP_Kernel3D.f90:415-526``): alpha in [0, pi], beta in [0, pi/2],
gamma in [0, 2 pi], the ranges the reference's run card states for the
library ``Gaussmesh`` it links (``Codes/exec/PNAMP.sh:26-28``; that library
file is not shipped).  For real D2-symmetric states the fold on this mesh
equals 2 / (d0_K1 d0_K2) times the plain projection over the full mesh
(d0_K = 2 for K = 0, else 1): an overall factor and a diagonal rescaling of
the K >= 0 basis, which leave the Hill-Wheeler energies unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def gauleg(a: float, b: float, n: int) -> tuple[np.ndarray, np.ndarray]:
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (b - a) * x + 0.5 * (b + a)
    return x, 0.5 * (b - a) * w


@dataclass
class Mesh:
    alp: np.ndarray
    walp: np.ndarray
    bet: np.ndarray
    wbet: np.ndarray
    gam: np.ndarray
    wgam: np.ndarray
    nphi: int
    jmax: int
    fold: int = 0          # 0 plain projection, 1 literal reference fold
    norm_method: int = 0   # 0 Pfaffian, 1 Onishi
    cscale: float = 1.0
    iscale: int = 0

    @property
    def n_euler(self) -> int:
        return self.alp.size * self.bet.size * self.gam.size

    @property
    def n_gauge(self) -> int:
        return self.nphi * self.nphi

    @property
    def n_points(self) -> int:
        return self.n_euler * self.n_gauge


def make_mesh(nalp: int, nbet: int, ngam: int, nphi: int, jmax: int, fold: int = 0,
              norm_method: int = 0, cscale: float = 1.0, iscale: int = 0) -> Mesh:
    alp, walp = gauleg(0.0, 2 * np.pi, nalp)
    x, wbet = gauleg(-1.0, 1.0, nbet)
    bet = np.arccos(x)
    gam, wgam = gauleg(0.0, 2 * np.pi, ngam)
    if nalp == 1:
        alp, walp = np.zeros(1), np.ones(1)
    if nbet == 1:
        bet, wbet = np.zeros(1), np.ones(1)
    if ngam == 1:
        gam, wgam = np.zeros(1), np.ones(1)
    return Mesh(alp, walp, bet, wbet, gam, wgam, nphi, jmax, fold, norm_method, cscale, iscale)


def fold_mesh(nalp: int, nbet: int, ngam: int, nphi: int, jmax: int, cscale: float = 1.0,
              iscale: int = 0) -> Mesh:
    """Gauss-Legendre alpha on [0, pi], cos(beta) on [0, 1], gamma on
    [0, 2 pi], for the literal fold (fold = 1)."""
    alp, walp = gauleg(0.0, np.pi, nalp)
    x, wbet = gauleg(0.0, 1.0, nbet)
    gam, wgam = gauleg(0.0, 2 * np.pi, ngam)
    return Mesh(alp, walp, np.arccos(x), wbet, gam, wgam, nphi, jmax, 1, 0, cscale, iscale)
