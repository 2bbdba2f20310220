"""The literal 3D mirror fold of the J-projection (fold = 1,
Codes/PNAMP/PNAMP_Kernel3D.f90:415-526) pinned by mathematics.

The fold sums four mirrored terms over the reduced mesh alpha in [0, pi],
beta in [0, pi/2], gamma in [0, 2 pi] (the ranges of the reference's run card,
Codes/exec/PNAMP.sh:26-28) with weight 2 (2J + 1) / (8 pi^2) and 1 / (d0_K1 d0_K2).
For real states with D2 symmetry (invariant under rotations by pi about x,
y, z: the symmetry the reference's GCM assumes when it keeps even K >= 0,
Codes/GCM/GCM_HWG.f90:166-170) it equals

    fold^J_{K1 K2} = 2 / (d0_K1 d0_K2) * N^J_{K1 K2}(plain, full mesh),

d0_K = 2 for K = 0, else 1, for every even J, every K1, K2 and every
observable; the plain projection is itself pinned by exact Fock-space
projection.  The factor is an overall scale times a diagonal rescaling of the
K >= 0 basis, so the Hill-Wheeler energies from the folded kernels in the
even K >= 0 basis equal those of the plain kernels in the full K basis.  A
state without the symmetry breaks the identity (negative check).
"""
import numpy as np
import pytest

from sccm_amd.basis import make_basis
from sccm_amd.driver import hw_matrices
from sccm_amd.interaction import synthetic_interaction
from sccm_amd.mesh import Mesh, fold_mesh
from sccm_amd.states import synthetic_state

REL = 1e-10


def plain_mesh(nang: int, nbet: int, nphi: int, jmax: int) -> Mesh:
    """Full Euler mesh that integrates the kernels of small spaces exactly:
    trapezoid rule in alpha, gamma (exact for |frequency| < nang),
    Gauss-Legendre in cos(beta)."""
    a = 2 * np.pi * np.arange(nang) / nang
    w = np.full(nang, 2 * np.pi / nang)
    x, wb = np.polynomial.legendre.leggauss(nbet)
    return Mesh(a, w, np.arccos(x), wb, a.copy(), w.copy(), nphi, jmax, 0)


def fold_factor(jmax: int) -> np.ndarray:
    d0 = np.array([2.0 if K == 0 else 1.0 for K in range(-jmax, jmax + 1)])
    return 2.0 / np.outer(d0, d0)


def fold_error(k_fold, k_plain, jmax, obs=range(6)):
    """max over observables of max_{even J, K1, K2} |fold - 2/(d0 d0) plain| /
    max |plain| (the observable's largest even-J kernel entry)"""
    f = fold_factor(jmax)
    worst = 0.0
    for io in obs:
        scale = max(np.abs(k_plain[io, J]).max() for J in range(0, jmax + 1, 2))
        for J in range(0, jmax + 1, 2):
            worst = max(worst, np.abs(k_fold[io, J] - f * k_plain[io, J]).max() / scale)
    return worst


# Gauss-Legendre over a full period of gamma needs about four points per unit
# of the largest frequency in the integrand (p shell, Z = N = 2, J <= 4: 8);
# alpha over half a period and cos(beta) converge with fewer
FOLD_P = (16, 12, 32)


def states(b, zn, seeds, d2=True, **kw):
    U, V = zip(*[synthetic_state(b, *zn, seed=s, sigma=0.4, sigma_pair=0.2, d2=d2, **kw) for s in seeds])
    return np.array(U), np.array(V)


def test_fold_equals_scaled_plain_for_d2_states(oracle_mod):
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    jmax = 4
    U, V = states(b, (2, 2), (5, 6))
    _, kp = oracle_mod.Oracle(b, inter, plain_mesh(12, 12, 5, jmax), 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    _, kf = oracle_mod.Oracle(b, inter, fold_mesh(*FOLD_P, 5, jmax), 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    for J in range(jmax + 1):  # D2: odd K components vanish
        for K1 in range(-J, J + 1, 1):
            if K1 % 2:
                assert np.abs(kp[:, J, K1 + jmax, :]).max() < 1e-14 * np.abs(kp[0]).max()
    e = fold_error(kf, kp, jmax)
    print(f"fold vs 2/(d0 d0) plain, all observables, even J <= {jmax}: {e:.2e}")
    assert e < REL


def test_fold_differs_without_d2_symmetry(oracle_mod):
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    jmax = 4
    U, V = states(b, (2, 2), (5, 6), d2=False)
    _, kp = oracle_mod.Oracle(b, inter, plain_mesh(12, 12, 5, jmax), 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    _, kf = oracle_mod.Oracle(b, inter, fold_mesh(*FOLD_P, 5, jmax), 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    assert fold_error(kf, kp, jmax, obs=(0,)) > 1e-2


def test_fold_matches_exact_fock_projection(oracle_mod):
    """fold = 2/(d0 d0) <Phi0| P^J_{K1K2} P^Z P^N |Phi1> (and H P^J P^Z P^N) from
    the exact Fock-space projectors."""
    from fock import Fock
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    jmax = 4
    U, V = states(b, (2, 2), (5, 6))
    _, kf = oracle_mod.Oracle(b, inter, fold_mesh(*FOLD_P, 5, jmax), 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    F = Fock(b)
    v0, v1 = F.hfb_state(U[0], V[0]), F.hfb_state(U[1], V[1])
    sec, P = F.projectors(2, 2, jmax)
    H = F.hamiltonian(inter)[sec][:, sec].toarray()
    f = fold_factor(jmax)
    scale = max(abs(np.vdot(v0[sec], Pm @ v1[sec])) for Pm in P.values())
    for (J, K1, K2), Pm in P.items():
        if J % 2:
            continue
        fk = f[K1 + jmax, K2 + jmax]
        assert abs(kf[0, J, K1 + jmax, K2 + jmax] - fk * np.vdot(v0[sec], Pm @ v1[sec])) < 1e-12 * scale
        assert abs(kf[4, J, K1 + jmax, K2 + jmax] - fk * np.vdot(v0[sec], H @ (Pm @ v1[sec]))) < 1e-10 * scale


def test_fold_gcm_energies_equal_plain(oracle_mod):
    """Hill-Wheeler energies of a 3-state D2 GCM: folded kernels in the even
    K >= 0 basis (what the reference's GCM reads) equal the plain kernels in
    the full K basis."""
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    jmax = 4
    coords = [(0.0, 0.0), (0.3, 20.0), (0.5, 45.0)]
    Us, Vs = zip(*[synthetic_state(b, 2, 2, seed=9, sigma=0.4, sigma_pair=0.2, d2=True, beta2=b2, gamma_deg=g,
                                   offset_seed=10) for b2, g in coords])
    pairs = np.array([(p, q) for p in range(3) for q in range(p, 3)])
    op = oracle_mod.Oracle(b, inter, plain_mesh(12, 12, 5, jmax), 2, 2)
    of = oracle_mod.Oracle(b, inter, fold_mesh(*FOLD_P, 5, jmax), 2, 2)
    kp = np.array([op.kernel_pair(Us[p], Vs[p], Us[q], Vs[q])[1] for p, q in pairs])
    kf = np.array([of.kernel_pair(Us[p], Vs[p], Us[q], Vs[q])[1] for p, q in pairs])
    for J in (0, 2, 4):
        Np, Hp, _, _ = hw_matrices(kp, pairs, 3, J, jmax)
        Nf, Hf, _, _ = hw_matrices(kf, pairs, 3, J, jmax, kvals=list(range(0, J + 1, 2)))
        rp = oracle_mod.hndiag(Np, Hp, 1e-8)
        rf = oracle_mod.hndiag(Nf, Hf, 1e-8)
        assert rp["m"] == rf["m"] > 0
        print(f"J={J}: {rf['m']} natural states, E {np.round(rf['E'], 6)}")
        assert np.abs(rp["E"] - rf["E"]).max() < 1e-8 * np.abs(rp["E"]).max()


@pytest.mark.gpu
def test_fold_on_gpu_equals_scaled_plain_and_oracle(oracle_mod):
    """sd space (n = 24), D2 states: GPU fold = 2/(d0 d0) GPU plain, and GPU fold
    = oracle fold on the same reduced mesh."""
    from sccm_amd.pnamp import PNAMP
    b = make_basis("sd")
    inter = synthetic_interaction(b, 1.5, 1001, 1002)
    jmax = 4
    U, V = states(b, (2, 2), (5, 6))
    gp = PNAMP(b, inter, plain_mesh(24, 24, 5, jmax))
    gp.set_states(U, V)
    kp = gp.kernel_pair(0, 1, 2, 2)
    gp.close()
    fm = fold_mesh(24, 16, 48, 5, jmax)
    gf = PNAMP(b, inter, fm)
    gf.set_states(U, V)
    kf = gf.kernel_pair(0, 1, 2, 2)
    gf.close()
    e = fold_error(kf, kp, jmax)
    print(f"GPU fold vs 2/(d0 d0) GPU plain (sd): {e:.2e}")
    assert e < REL
    sub = fold_mesh(6, 5, 6, 5, jmax)
    g = PNAMP(b, inter, sub)
    g.set_states(U, V)
    kg = g.kernel_pair(0, 1, 2, 2)
    g.close()
    _, ko = oracle_mod.Oracle(b, inter, sub, 2, 2).kernel_pair(U[0], V[0], U[1], V[1])
    for io in range(6):
        for J in range(0, jmax + 1, 2):
            assert np.abs(kg[io, J] - ko[io, J]).max() < REL * np.abs(ko[io, J]).max()
