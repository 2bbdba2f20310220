"""The oracle against exact Fock-space brute force (tests/fock.py) on the
p shell (0p3/2, 0p1/2; n = 12 m-states), Z = N = 2.

Pins what the Ge76 goldens cannot: the full 3D rotation convention, the sign
of the Pfaffian norm, the transition densities, both energy formulas and the
plain J-projection for all (J, K, K')."""
import numpy as np
import pytest

from fock import Fock
from sccm_amd.basis import make_basis
from sccm_amd.interaction import synthetic_interaction
from sccm_amd.mesh import Mesh, make_mesh
from sccm_amd.states import synthetic_state

POINTS = [(0.0, 0.0, 0.0, 0.0, 0.0), (0.3, 0.7, 1.1, 0.4, 2.0), (2.0, 2.5, -0.3, 1.0, 0.2),
          (4.0, 1.9, 5.5, 3.0, 4.4)]


@pytest.fixture(scope="module")
def setup():
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    U0, V0 = synthetic_state(b, 2, 2, seed=5, sigma=0.4, sigma_pair=0.3)
    U1, V1 = synthetic_state(b, 2, 2, seed=6, sigma=0.4, sigma_pair=0.3)
    F = Fock(b)
    v0, v1 = F.hfb_state(U0, V0), F.hfb_state(U1, V1)
    return b, inter, (U0, V0, U1, V1), F, v0, v1


@pytest.mark.parametrize("pt", POINTS)
def test_point_quantities(oracle_mod, setup, pt):
    b, inter, uv, F, v0, v1 = setup
    alp, bet, gam, pp, pn = pt
    o = oracle_mod.Oracle(b, inter, make_mesh(1, 1, 1, 1, 0), 2, 2)
    r = o.point(*uv, alp, bet, gam, pp, pn)
    w1 = F.rotate(v1, alp, bet, gam, pp, pn)
    exact = np.vdot(v0, w1) * np.exp(-1j * (2 * pp + 2 * pn))
    assert abs(r["norm"] - exact) < 1e-13 * max(1.0, abs(exact)) + 1e-15
    assert abs(r["det0"] - abs(v0[0])) < 1e-14
    # the Onishi formula fixes the magnitude but not the sign
    assert abs(abs(r["onishi"]) - abs(np.vdot(v0, w1))) < 1e-12
    rho, k10, k01 = F.densities(v0, w1)
    scale = 1e-12 / max(abs(np.vdot(v0, w1)), 1e-3)
    assert np.abs(rho - r["rho"]).max() < scale * max(1, np.abs(rho).max())
    assert np.abs(k10 - r["k10"]).max() < scale * max(1, np.abs(k10).max())
    assert np.abs(k01 - r["k01"]).max() < scale * max(1, np.abs(k01).max())
    H = F.hamiltonian(inter)
    E = np.vdot(v0, H @ w1) / np.vdot(v0, w1)
    assert abs(r["ehfb"] - E) < 1e-11 * max(1, abs(E)) / max(abs(np.vdot(v0, w1)), 1e-3)
    assert abs(r["ehfb_field"] - r["ehfb"]) < 1e-11 * max(1, abs(E))


def test_plain_projection_exact(oracle_mod, setup):
    """N^J_{KK'} = <Phi0|P^J_{KK'} P^Z P^N|Phi1> and H^J likewise, exactly:
    uniform alpha/gamma meshes (exact for the trigonometric polynomials
    involved), Gauss-Legendre cos(beta) with enough points, 7 Fomenko points
    (exact for 0..6 particles per species)."""
    b, inter, uv, F, v0, v1 = setup
    na, jmax = 12, 4
    alp = np.arange(na) * 2 * np.pi / na
    w = np.full(na, 2 * np.pi / na)
    x, wb = np.polynomial.legendre.leggauss(12)
    m = Mesh(alp, w, np.arccos(x), wb, alp.copy(), w.copy(), 7, jmax, 0, 0)
    o = oracle_mod.Oracle(b, inter, m, 2, 2)
    _, kern = o.kernel_pair(*uv)
    sec, P = F.projectors(2, 2, jmax)
    H = F.hamiltonian(inter)[sec][:, sec].toarray()
    a0, a1 = v0[sec], v1[sec]
    nmax = max(abs(np.vdot(a0, Pm @ a1)) for Pm in P.values())
    assert nmax > 1e-3
    seen = set()
    for (J, K1, K2), Pm in P.items():
        seen.add((J, K1, K2))
        assert abs(np.vdot(a0, Pm @ a1) - kern[0, J, K1 + jmax, K2 + jmax]) < 1e-13
        assert abs(np.vdot(a0, H @ (Pm @ a1)) - kern[4, J, K1 + jmax, K2 + jmax]) < 1e-12
    # J values with no states in the sector project to zero
    for J in range(jmax + 1):
        for K1 in range(-J, J + 1):
            for K2 in range(-J, J + 1):
                if (J, K1, K2) not in seen:
                    assert abs(kern[0, J, K1 + jmax, K2 + jmax]) < 1e-13


def test_j2_kernel_is_j_j_plus_1(oracle_mod, setup):
    """<J^2> projected onto J is J(J+1) times the norm kernel (the reference's
    own check, Codes/GCM/GCM_IO.f90:233-237), exactly on an exact mesh."""
    b, inter, uv, F, v0, v1 = setup
    na, jmax = 12, 4
    alp = np.arange(na) * 2 * np.pi / na
    w = np.full(na, 2 * np.pi / na)
    x, wb = np.polynomial.legendre.leggauss(12)
    m = Mesh(alp, w, np.arccos(x), wb, alp.copy(), w.copy(), 7, jmax, 0, 0)
    _, kern = oracle_mod.Oracle(b, inter, m, 2, 2).kernel_pair(*uv)
    for J in range(jmax + 1):
        assert np.abs(kern[5, J] - J * (J + 1) * kern[0, J]).max() < 1e-12
