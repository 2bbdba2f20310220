"""One-body operators of the E2 observables.

``q2_matrices`` builds the m-scheme quadrupole matrices Q_{2mu} (mu = -2..2)
the way ``q2mume`` does (``Codes/GCM/VAPHFB_Q2.f90:2-106``, called from
``SingleParticleME``, ``Codes/PNAMP/PNAMP_ME1B.f90:138-216``):

    Q_{2mu}(a, b) = delta(t_a, t_b) <n_a l_a | r^2 | n_b l_b> <l_a j_a m_a| C^(2)_mu |l_b j_b m_b>,

C^(2) = sqrt(4 pi / 5) Y_2 (the reference divides r^2 Y_2mu by sqrt(5 / 4 pi)),
the angular part summed over the spin projections with Clebsch-Gordan
coefficients (the CJJ calls of q2mume), the radial part the harmonic-
oscillator integral (``rnla``, ``Codes/HFB/MATH_LIB.f90:85``) times b^2, here by
generalized Gauss-Laguerre quadrature, exact for these polynomial integrands.
"""
from __future__ import annotations

from math import pi, sqrt

import numpy as np
from scipy.special import eval_genlaguerre, roots_genlaguerre

from .basis import Basis
from .interaction import cg


def ho_radial_r2(n1: int, l1: int, n2: int, l2: int) -> float:
    """<n1 l1 | (r/b)^2 | n2 l2> for harmonic-oscillator radial functions
    R_nl(r) = N r^l L_n^(l+1/2)(r^2) e^(-r^2/2), n = 0, 1, ... (b = 1)."""
    def norm(n, l):
        x, w = roots_genlaguerre(n + 4, l + 0.5)
        return 1.0 / sqrt(0.5 * np.sum(w * eval_genlaguerre(n, l + 0.5, x) ** 2))
    # r^2 dr r^2 R1 R2 with x = r^2: (1/2) x^((l1+l2+3)/2) L1(x) L2(x) e^-x dx
    alpha = 0.5 * (l1 + l2 + 3)
    x, w = roots_genlaguerre(n1 + n2 + 4, alpha)
    val = 0.5 * np.sum(w * eval_genlaguerre(n1, l1 + 0.5, x) * eval_genlaguerre(n2, l2 + 0.5, x))
    return norm(n1, l1) * norm(n2, l2) * val


def q2_matrices(b: Basis, b_osc: float = 1.0) -> np.ndarray:
    """[5, n, n] real: Q_{2mu}(a, b) for mu = -2..2 (index mu + 2)."""
    n = b.n
    Q = np.zeros((5, n, n))
    for ia in range(n):
        la, ja, ma = int(b.l[ia]), int(b.twoj[ia]), int(b.twom[ia])
        for ib in range(n):
            if b.t[ia] != b.t[ib]:
                continue
            lb, jb, mb = int(b.l[ib]), int(b.twoj[ib]), int(b.twom[ib])
            mu2 = ma - mb  # 2 mu
            if abs(mu2) > 4:
                continue
            common = cg(4, 0, 2 * lb, 0, 2 * la, 0)
            if common == 0.0:
                continue
            ang = 0.0
            for ms in (-1, 1):
                mla, mlb = ma - ms, mb - ms
                if abs(mla) > 2 * la or abs(mlb) > 2 * lb:
                    continue
                ang += cg(2 * la, mla, 1, ms, ja, ma) * cg(2 * lb, mlb, 1, ms, jb, mb) * cg(4, mu2, 2 * lb, mlb, 2 * la, mla)
            factor = sqrt((2 * lb + 1) / (2 * la + 1))  # <la||C2||lb> part with the 1/sqrt(5/4pi) applied
            rad = b_osc ** 2 * ho_radial_r2(int(b.nq[ia]) - 1, la, int(b.nq[ib]) - 1, lb)
            Q[mu2 // 2 + 2, ia, ib] = factor * common * rad * ang
    return Q
