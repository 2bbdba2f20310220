"""E2 kernels and the one-body transition density (iE2 = 1 of Kernel4Triaxial,
Codes/PNAMP/PNAMP_Kernel3D.f90:331-360,527-710; ZRho1B_Angles/Integral,
Codes/PNAMP/PNAMP_Rho1B.f90:2-62), pinned by mathematics:

* the Q_{2mu} one-body matrices (q2mume restated in sccm_amd.operators):
  harmonic-oscillator <r^2> = (2n + l + 3/2) b^2, hermiticity
  Q_mu^T = (-1)^mu Q_{-mu}, Wigner-Eckart structure in m;
* the reduced matrix elements: by 3j orthogonality the reference's formula is
  the Wigner-Eckart reduced element between projected states,
  <Phi0| P^{J1}_{K1 M1} Q_{2mu} P^{J2}_{M2 K2} |Phi1>
      = (-1)^(J1 - M1) (J1 2 J2; -M1 mu M2) <J1 K1 || Q2 || J2 K2>,
  checked against exact Fock-space projectors on the p shell, and zero for
  J1 above the largest J of the space;
* the J = 0 one-body density against <Phi0| sum_m c+_{b m} c_{a m} P^0_00 |Phi1> / sqrt(2 j + 1).
"""
import numpy as np
import pytest

from sccm_amd.basis import make_basis
from sccm_amd.interaction import synthetic_interaction
from sccm_amd.mesh import Mesh
from sccm_amd.operators import ho_radial_r2, q2_matrices
from sccm_amd.states import synthetic_state

REL = 1e-10


def full_mesh(nang, nbet, nphi, jmax):
    a = 2 * np.pi * np.arange(nang) / nang
    w = np.full(nang, 2 * np.pi / nang)
    x, wb = np.polynomial.legendre.leggauss(nbet)
    return Mesh(a, w, np.arccos(x), wb, a.copy(), w.copy(), nphi, jmax, 0)


def qidx(jm, dJ, J2, K1, K2):
    return dJ + 2, J2, K1 + jm + 2, K2 + jm


def test_q2_matrices_closed_forms(oracle_mod):
    for (n, l) in ((0, 0), (0, 3), (1, 1), (2, 2)):
        assert abs(ho_radial_r2(n, l, n, l) - (2 * n + l + 1.5)) < 1e-12
    b = make_basis("pf")
    Q = q2_matrices(b, b_osc=1.7)
    for mu in range(-2, 3):
        assert np.abs(Q[mu + 2].T - (-1) ** mu * Q[-mu + 2]).max() < 1e-14
    # Wigner-Eckart: Q_mu(a, b) = (-1)^(ja - ma) (ja 2 jb; -ma mu mb) <a||Q||b>, one
    # reduced element per orbit pair (both species equal)
    red = {}
    for a in range(b.n):
        for c in range(b.n):
            if b.t[a] != b.t[c]:
                assert np.all(Q[:, a, c] == 0)
                continue
            mu2 = int(b.twom[a] - b.twom[c])
            tj = oracle_mod.threej(int(b.twoj[a]), 4, int(b.twoj[c]), -int(b.twom[a]), mu2, int(b.twom[c]))
            if abs(tj) < 1e-12:
                assert abs(Q[:, a, c]).max() < 1e-13
                continue
            ph = (-1) ** ((int(b.twoj[a]) - int(b.twom[a])) // 2)
            r = Q[mu2 // 2 + 2, a, c] / (ph * tj)
            key = (int(b.orbit[a]), int(b.orbit[c]))
            red.setdefault(key, r)
            assert abs(r - red[key]) < 1e-12 * max(1.0, abs(r))
    # f7/2 diagonal: <f7/2 m=7/2| Q_20 |f7/2 m=7/2> = <r^2> <j m|C_20|j m> = 4.5 b^2 (-1/3)
    k = int(np.flatnonzero((b.twoj == 7) & (b.twom == 7) & (b.t < 0))[0])
    assert abs(Q[2, k, k] - 1.7 ** 2 * 4.5 * (-1.0 / 3.0)) < 1e-12


def _setup(oracle_mod, jmax=4):
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    Q = q2_matrices(b)
    U, V = zip(*[synthetic_state(b, 2, 2, seed=s, sigma=0.4, sigma_pair=0.3) for s in (5, 6)])
    o = oracle_mod.Oracle(b, inter, full_mesh(14, 12, 5, jmax), 2, 2, q2=Q)
    return b, inter, Q, U, V, o


def test_e2_reduced_elements_match_fock_wigner_eckart(oracle_mod):
    from fock import Fock
    jmax = 4
    b, inter, Q, U, V, o = _setup(oracle_mod, jmax)
    _, kern, ze2, zrho, qred, rho1b = o.kernel_pair_e2(U[0], V[0], U[1], V[1])
    F = Fock(b)
    v0, v1 = F.hfb_state(U[0], V[0]), F.hfb_state(U[1], V[1])
    sec, P = F.projectors(2, 2, jmax + 2)
    h = b.n // 2
    scale = np.abs(qred).max()
    checked = 0
    for sp in range(2):
        inside = (np.arange(b.n) >= h) == bool(sp)  # m-states of species sp
        Qs = [F.one_body(np.where(np.outer(inside, inside), Q[m], 0.0))[sec][:, sec].toarray() for m in range(5)]
        for J2 in range(jmax + 1):
            for dJ in range(-2, 3):
                J1 = J2 + dJ
                if J1 < 0:
                    continue
                for K1 in range(-J1, J1 + 1):
                    for K2 in range(-J2, J2 + 1):
                        red = qred[(sp,) + qidx(jmax, dJ, J2, K1, K2)]
                        if (J1, K1, K1) not in P or (J2, K2, K2) not in P:
                            assert abs(red) < 1e-10 * scale, (sp, J1, K1, J2, K2)
                            continue
                        for mu in range(-2, 3):
                            for M2 in range(-J2, J2 + 1):
                                M1 = M2 + mu
                                if abs(M1) > J1:
                                    continue
                                tj = oracle_mod.threej(2 * J1, 4, 2 * J2, -2 * M1, 2 * mu, 2 * M2)
                                if abs(tj) < 1e-12:
                                    continue
                                G = np.vdot(v0[sec], P[(J1, K1, M1)] @ (Qs[mu + 2] @ (P[(J2, M2, K2)] @ v1[sec])))
                                assert abs(G - (-1) ** (J1 - M1) * tj * red) < 1e-10 * scale, (sp, J1, K1, J2, K2, mu, M2)
                                checked += 1
    assert checked > 1000


def test_rho1b_matches_fock(oracle_mod):
    from fock import Fock
    b, inter, Q, U, V, o = _setup(oracle_mod)
    _, kern, ze2, zrho, qred, rho1b = o.kernel_pair_e2(U[0], V[0], U[1], V[1])
    F = Fock(b)
    v0, v1 = F.hfb_state(U[0], V[0]), F.hfb_state(U[1], V[1])
    sec, P = F.projectors(2, 2, 0)
    h = b.n // 2
    first = [0, 4]  # p3/2 (4 m-states), p1/2 (2) in each species half
    tj = [3, 1]
    for sp in range(2):
        for a in range(2):
            for c in range(2):
                if tj[a] != tj[c]:
                    assert abs(rho1b[sp, a, c]) < 1e-14
                    continue
                E = np.zeros((b.n, b.n))
                for m in range(tj[a] + 1):
                    E[sp * h + first[c] + m, sp * h + first[a] + m] = 1.0  # c+_{c m} c_{a m}
                O = F.one_body(E)[sec][:, sec].toarray()
                ref = np.vdot(v0[sec], O @ (P[(0, 0, 0)] @ v1[sec])) / np.sqrt(tj[a] + 1.0)
                assert abs(rho1b[sp, a, c] - ref) < 1e-12, (sp, a, c)
    # sum over orbits of sqrt(2j+1) rho1b = <Z> or <N> times N^0
    for sp, nval in ((0, 2), (1, 2)):
        tot = sum(np.sqrt(tj[a] + 1.0) * rho1b[sp, a, a] for a in range(2))
        assert abs(tot - nval * kern[0, 0, 4, 4]) < 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# f-p orbit sets between the named spaces: with f5pg9 they give the plain
# size classes 16, 32, 48 with E2 (n = 16, 28, 44) and the factored form at
# h = 8, 14
FP_ORBITS = {"f7": [(0, 3, 7)], "f7p": [(0, 3, 7), (1, 1, 3), (1, 1, 1)]}


@pytest.mark.gpu
@pytest.mark.parametrize("space,zn,form", [("p", (2, 2), 1), ("sd", (2, 2), 0), ("sd", (2, 2), 1), ("pf", (2, 4), 0),
                                           ("pf", (2, 4), 1), ("f5pg9", (4, 2), 0), ("f5pg9", (4, 2), 1),
                                           *[(s, (2, 4), f) for s in FP_ORBITS for f in (0, 1)]])
def test_e2_kernels_on_gpu_match_oracle(oracle_mod, space, zn, form):
    from sccm_amd.mesh import make_mesh
    from sccm_amd.pnamp import PNAMP
    b = make_basis(space, FP_ORBITS.get(space))
    inter = synthetic_interaction(b, 1.2, 201, 202)
    Q = q2_matrices(b, b_osc=1.9)
    mesh = make_mesh(3, 4, 3, 5, 4)
    U, V = zip(*[synthetic_state(b, *zn, seed=s) for s in (21, 22)])
    g = PNAMP(b, inter, mesh)
    g.set_states(np.array(U), np.array(V))
    g.set_q2(Q)
    g.set_form(form)
    pairs = [(0, 1), (1, 1)]
    zr_g, k_g, ze2_g, qred_g, rho_g = g.kernels_e2(pairs, *zn)
    o = oracle_mod.Oracle(b, inter, mesh, *zn, q2=Q)
    for i, (p, q) in enumerate(pairs):
        zr_o, k_o, ze2_o, zrho_o, qred_o, rho_o = o.kernel_pair_e2(U[p], V[p], U[q], V[q])
        assert _rel(zr_g[i], zr_o) < REL
        assert _rel(ze2_g[i], ze2_o) < REL, (space, form, "ze2", _rel(ze2_g[i], ze2_o))
        assert _rel(qred_g[i], qred_o) < REL, (space, form, "qred", _rel(qred_g[i], qred_o))
        assert _rel(rho_g[i], rho_o) < REL, (space, form, "rho1b", _rel(rho_g[i], rho_o))
    g.close()


@pytest.mark.gpu
def test_e2_kernels_on_gpu_pn_mixed_ti48(oracle_mod):
    """48Ti (proton-neutron pairing, plain form only) through the E2 path."""
    import ti48
    from sccm_amd.mesh import make_mesh
    from sccm_amd.pnamp import PNAMP
    b, inter, U, V, betas = ti48.load()
    Q = q2_matrices(b, b_osc=2.0)
    mesh = make_mesh(2, 4, 2, 5, 4)
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    g.set_q2(Q)
    zr_g, k_g, ze2_g, qred_g, rho_g = g.kernels_e2([(3, 5)], ti48.ZVAL, ti48.NVAL)
    o = oracle_mod.Oracle(b, inter, mesh, ti48.ZVAL, ti48.NVAL, q2=Q)
    zr_o, k_o, ze2_o, zrho_o, qred_o, rho_o = o.kernel_pair_e2(U[3], V[3], U[5], V[5])
    assert _rel(ze2_g[0], ze2_o) < REL and _rel(qred_g[0], qred_o) < REL and _rel(rho_g[0], rho_o) < REL
    g.close()


@pytest.mark.gpu
def test_e2_refuses_the_literal_fold():
    from sccm_amd.mesh import fold_mesh
    from sccm_amd.pnamp import PNAMP, PnampError
    b = make_basis("sd")
    g = PNAMP(b, synthetic_interaction(b, 1.2, 201, 202), fold_mesh(2, 2, 2, 3, 2))
    U, V = synthetic_state(b, 2, 2, seed=21)
    g.set_states(U, V)
    g.set_q2(q2_matrices(b))
    with pytest.raises(PnampError):
        g.kernels_e2([(0, 0)], 2, 2)
    g.close()
