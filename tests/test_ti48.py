"""The reference's 48Ti states (data/Ti: KB3G, pf space, n = 40, with
proton-neutron pairing) through the plain FP64 path against the oracle."""
import numpy as np
import pytest

import ti48
from sccm_amd.driver import hw_matrices
from sccm_amd.mesh import make_mesh
from sccm_amd.states import check_state

REL = 1e-10


def test_ti48_inputs_are_bogoliubov_states_with_pn_mixing(oracle_mod):
    b, inter, U, V, betas = ti48.load()
    assert U.shape == (8, 40, 40) and inter.nme == 14442
    o = oracle_mod.Oracle(b, inter, make_mesh(1, 1, 1, 1, 0), ti48.ZVAL, ti48.NVAL)
    for q in range(8):
        unit, skew = check_state(U[q], V[q])
        # states with an HF-like species have a nearly singular U (condition
        # number up to 6e4 for beta2 = -0.3): Z = V U^-1 is skew relative to |Z|
        zmax = np.abs(V[q] @ np.linalg.inv(U[q])).max()
        assert unit < 1e-9 and skew < 1e-9 * zmax
        r = o.point(U[q], V[q], U[q], V[q], 0, 0, 0, 0, 0)
        assert abs(np.trace(r["rho"][:20, :20]).real - ti48.ZVAL) < 1e-3
        assert abs(np.trace(r["rho"][20:, 20:]).real - ti48.NVAL) < 1e-3
    Z = V[3] @ np.linalg.inv(U[3])
    assert np.abs(Z[:20, 20:]).max() > 1e-3  # proton-neutron pairing: only the plain form applies


@pytest.mark.gpu
def test_ti48_pairs_match_oracle(oracle_mod):
    from sccm_amd.pnamp import FORM_FACTORED, PNAMP, PnampError
    b, inter, U, V, betas = ti48.load()
    mesh = make_mesh(2, 4, 2, 5, 4)
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    assert not g.separable
    o = oracle_mod.Oracle(b, inter, mesh, ti48.ZVAL, ti48.NVAL)
    for p, q in ((3, 3), (3, 5), (0, 7)):
        zr_g, k_g = g.kernel_pair(p, q, ti48.ZVAL, ti48.NVAL, want_zres=True)
        zr_o, k_o = o.kernel_pair(U[p], V[p], U[q], V[q])
        for io in range(6):
            # relative to the observable's scale, at least the norm's (<J^2> of the
            # spherical beta2 = 0 state vanishes: its own maximum is rounding)
            sz = max(np.abs(zr_o[:, io]).max(), np.abs(zr_o[:, 0]).max())
            assert np.abs(zr_g[:, io] - zr_o[:, io]).max() < REL * sz, (p, q, io)
            sk = max(np.abs(k_o[io]).max(), np.abs(k_o[0]).max())
            assert np.abs(k_g[io] - k_o[io]).max() < REL * sk, (p, q, io)
    g.set_form(FORM_FACTORED)
    with pytest.raises(PnampError):
        g.kernel_pair(0, 1, ti48.ZVAL, ti48.NVAL)
    g.close()


@pytest.mark.gpu
def test_ti48_axial_gcm_matches_oracle_chain(oracle_mod):
    """The eight axial states as a GCM in beta2 (K = 0, J = 0, 2, 4, 6 on the
    axial mesh): GPU kernels and pnamp_hw_solve against the oracle kernels and
    HNDIAG; projected <Z>, <N> of the diagonal kernels."""
    from sccm_amd.pnamp import PNAMP
    b, inter, U, V, betas = ti48.load()
    mesh = ti48.axial_mesh()
    pairs = np.array([(p, q) for p in range(8) for q in range(p, 8)])
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    k_g = g.kernels_batched(pairs, ti48.ZVAL, ti48.NVAL)
    o = oracle_mod.Oracle(b, inter, mesh, ti48.ZVAL, ti48.NVAL)
    k_o = np.array([o.kernel_pair(U[p], V[p], U[q], V[q])[1] for p, q in pairs])
    jm = mesh.jmax
    for i in range(len(pairs)):
        for io in range(6):  # scale at least the norm's, as above
            sk = max(np.abs(k_o[i, io]).max(), np.abs(k_o[i, 0]).max())
            for J in (0, 2, 4, 6):
                ref = k_o[i, io, J, jm, jm]
                assert abs(k_g[i, io, J, jm, jm] - ref) <= REL * sk, (i, io, J)
    for J in (0, 2, 4, 6):
        Ng, Hg, _, _ = hw_matrices(k_g, pairs, 8, J, jm, kvals=[0])
        No, Ho, _, _ = hw_matrices(k_o, pairs, 8, J, jm, kvals=[0])
        rg = g.hw_solve(Ng, Hg, 1e-6)
        ro = oracle_mod.hndiag(No, Ho, 1e-6)
        assert rg["m"] == ro["m"]
        assert np.abs(rg["E"] - ro["E"]).max() < REL * np.abs(ro["E"]).max()
        d = [i for i, (p, q) in enumerate(pairs) if p == q]
        zz = [(k_g[i, 1, J, jm, jm] / k_g[i, 0, J, jm, jm]).real for i in d]
        nn = [(k_g[i, 2, J, jm, jm] / k_g[i, 0, J, jm, jm]).real for i in d]
        print(f"Ti48 J={J}: {rg['m']} natural states, E {np.round(rg['E'][:3], 4)} MeV, "
              f"<Z> {min(zz):.6f}..{max(zz):.6f}, <N> {min(nn):.6f}..{max(nn):.6f}")
    g.close()
