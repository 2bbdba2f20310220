"""The reference's sd-pf states (data/Ca, data/Ne: SDPFMU, n = 64, with
proton-neutron pairing) through the plain FP64 path against the oracle: the
largest model space the reference ships, sized by NLEV in its kernels
(Codes/PNAMP/PNAMP_Proj.f90:142-160)."""
import numpy as np
import pytest

import sdpf
from sccm_amd.mesh import make_mesh
from sccm_amd.operators import q2_matrices
from sccm_amd.states import check_state

REL = 1e-10


def test_sdpf_inputs_are_bogoliubov_states_with_pn_mixing(oracle_mod):
    from sccm_amd.interaction import list_indices
    b, inter, U, V = sdpf.load()
    assert b.n == 64 and U.shape == (2, 64, 64)
    # the binary list holds exactly the records of the reference's nested loop order
    assert inter.nme == 51048 == len(list_indices(b))
    for name, (q, zv, nv) in sdpf.NUCLEI.items():
        unit, skew = check_state(U[q], V[q])
        zmax = np.abs(V[q] @ np.linalg.inv(U[q])).max()
        assert unit < 1e-9 and skew < 1e-9 * zmax, name
        rho = V[q] @ V[q].T
        assert abs(np.trace(rho[:32, :32]) - zv) < 1e-3 and abs(np.trace(rho[32:, 32:]) - nv) < 1e-3, name
        Z = V[q] @ np.linalg.inv(U[q])
        assert np.abs(Z[:32, 32:]).max() > 1e-2, name  # proton-neutron pairing: plain form only


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["Ca44", "Ne20"])
def test_sdpf_pairs_match_oracle(oracle_mod, name):
    from sccm_amd.pnamp import PNAMP
    b, inter, U, V = sdpf.load()
    q, zv, nv = sdpf.NUCLEI[name]
    mesh = make_mesh(2, 3, 2, 5, 4)
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    assert not g.separable
    print(f"sdpf {name}: n = 64, {mesh.n_points} mesh points", flush=True)
    zr_g, k_g = g.kernel_pair(q, q, zv, nv, want_zres=True)
    zr_o, k_o = oracle_mod.Oracle(b, inter, mesh, zv, nv).kernel_pair(U[q], V[q], U[q], V[q])
    for io in range(6):
        sz = max(np.abs(zr_o[:, io]).max(), np.abs(zr_o[:, 0]).max())
        assert np.abs(zr_g[:, io] - zr_o[:, io]).max() < REL * sz, (name, io)
        sk = max(np.abs(k_o[io]).max(), np.abs(k_o[0]).max())
        assert np.abs(k_g[io] - k_o[io]).max() < REL * sk, (name, io)
    g.close()


@pytest.mark.gpu
def test_sdpf_e2_kernels_match_oracle(oracle_mod):
    """E2 traces, reduced E2 elements and rho1b at n = 64 (7 orbits per species)."""
    from sccm_amd.pnamp import PNAMP
    b, inter, U, V = sdpf.load()
    q, zv, nv = sdpf.NUCLEI["Ne20"]
    Q = q2_matrices(b, b_osc=1.8)
    mesh = make_mesh(2, 3, 2, 3, 4)
    g = PNAMP(b, inter, mesh)
    g.set_states(U, V)
    g.set_q2(Q)
    zr_g, k_g, ze2_g, qred_g, rho_g = g.kernels_e2([(q, q)], zv, nv)
    zr_o, k_o, ze2_o, zrho_o, qred_o, rho_o = oracle_mod.Oracle(b, inter, mesh, zv, nv, q2=Q).kernel_pair_e2(
        U[q], V[q], U[q], V[q])
    rel = lambda a, r: float(np.abs(a - r).max() / max(np.abs(r).max(), 1e-300))
    assert rel(zr_g[0], zr_o) < REL and rel(ze2_g[0], ze2_o) < REL
    assert rel(qred_g[0], qred_o) < REL and rel(rho_g[0], rho_o) < REL
    g.close()
