"""Natural-state selection of PreSolve_HWG (Codes/GCM/GCM_HWG.f90:2-100,
mean_variance :205-254): the oracle restatement pinned by mathematics, and
the device scan (pnamp_hw_presolve) against it."""
import numpy as np
import pytest

from sccm_amd import configs
from sccm_amd.driver import hw_matrices


def _pd_problem(dim, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
    N = (Q * np.geomspace(1.0, 1.0 / cond, dim)) @ Q.T
    Hs = rng.normal(size=(dim, dim))
    return N, Hs + Hs.T


def test_presolve_exact_eigenpairs_have_zero_variance(oracle_mod):
    """With a well-conditioned N every natural state is kept, and with all of
    them (NOS = dim) the HNDIAG amplitudes solve H f = E N f exactly: the
    variance vanishes to rounding and is the smallest, so every level picks
    NOS = dim and its energy is the generalized eigenvalue."""
    import scipy.linalg
    N, H = _pd_problem(12, 3)
    r = oracle_mod.presolve_hwg(N, H, 1e-8, kmax=5)
    assert r["m"] == 12
    assert (r["nos"] == 12).all()
    ev = scipy.linalg.eigh(H, N, eigvals_only=True)
    assert np.abs(r["e_best"] - ev[:5]).max() < 1e-10 * np.abs(ev).max()
    assert r["var"][11, :5].max() < 1e-20
    assert r["var"][:11].max() > 1e-6  # truncated bases leave a residual


def test_presolve_literal_equals_nested_blocks(oracle_mod):
    """The reference's loop (one HNDIAG per NOS) and the nested-block form
    give the same levels, variances and choices."""
    N, H = _pd_problem(18, 5, cond=1e8)
    a = oracle_mod.presolve_hwg(N, H, 1e-6, kmax=4, literal=True)
    b = oracle_mod.presolve_hwg(N, H, 1e-6, kmax=4, literal=False)
    assert a["m"] == b["m"]
    assert np.abs(a["E"] - b["E"]).max() < 1e-9 * np.abs(a["E"]).max()
    assert np.abs(a["var"] - b["var"]).max() < 1e-8 * max(a["var"].max(), 1.0)
    assert (a["nos"] == b["nos"]).all()


@pytest.mark.gpu
def test_presolve_on_gpu_matches_oracle(oracle_mod):
    """Reduced pf GCM (5 states, 15 pairs): GPU kernels and pnamp_hw_presolve
    against the oracle kernels and the literal PreSolve_HWG restatement."""
    from sccm_amd.pnamp import FORM_FACTORED, PNAMP
    w = configs.pf_gcm(nstates=5, nalp=4, nbet=6, nphi=5, jmax=4)
    o = oracle_mod.Oracle(w.basis, w.inter, w.mesh, w.zval, w.nval)
    k_o = np.array([o.kernel_pair(w.U[p], w.V[p], w.U[q], w.V[q])[1] for p, q in w.pairs])
    g = PNAMP(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(FORM_FACTORED)
    k_g = g.kernels_batched(w.pairs, w.zval, w.nval)
    for J in range(w.mesh.jmax + 1):
        No, Ho, _, _ = hw_matrices(k_o, w.pairs, 5, J, w.mesh.jmax)
        Ng, Hg, _, _ = hw_matrices(k_g, w.pairs, 5, J, w.mesh.jmax)
        ro = oracle_mod.presolve_hwg(No, Ho, w.hw_eps, kmax=3)
        rg = g.hw_presolve(Ng, Hg, w.hw_eps, kmax=3)
        assert rg["m"] == ro["m"], J
        assert np.abs(rg["E"] - ro["E"]).max() < 1e-10 * np.abs(ro["E"]).max(), J
        assert np.abs(rg["var"] - ro["var"]).max() < 1e-8 * max(ro["var"].max(), 1.0), J
        assert (rg["nos"] == ro["nos"]).all(), (J, rg["nos"], ro["nos"])
        print(f"J={J}: M={ro['m']} NOS per level {list(ro['nos'])} E {np.round(ro['e_best'], 4)}")
    g.close()


@pytest.mark.gpu
def test_presolve_on_gpu_dimension_340(oracle_mod):
    from sccm_amd.basis import make_basis
    from sccm_amd.interaction import synthetic_interaction
    from sccm_amd.mesh import make_mesh
    from sccm_amd.pnamp import PNAMP
    rng = np.random.default_rng(341)
    dim = 340
    Q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
    en = np.concatenate([np.geomspace(1.0, 1e-4, 120), np.geomspace(1e-9, 1e-13, dim - 120)])
    N = (Q * en) @ Q.T
    Hs = rng.normal(size=(dim, dim))
    H = N @ (Hs + Hs.T) @ N / 10.0 + 1e-7 * (Hs + Hs.T)
    b = make_basis("sd")
    g = PNAMP(b, synthetic_interaction(b, 1.0, 1, 2), make_mesh(1, 1, 1, 1, 0))
    rg = g.hw_presolve(N, H, 1e-6, kmax=10)
    ro = oracle_mod.presolve_hwg(N, H, 1e-6, kmax=10, literal=False)
    assert rg["m"] == ro["m"] == 120
    assert np.abs(rg["E"] - ro["E"]).max() < 1e-10 * np.abs(ro["E"]).max()
    assert (rg["nos"] == ro["nos"]).all(), (rg["nos"], ro["nos"])
    g.close()
