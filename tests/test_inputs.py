"""Host-side inputs: basis order, interaction list rule, synthetic states."""
import numpy as np

from fock import Fock
from sccm_amd import configs
from sccm_amd.basis import make_basis
from sccm_amd.interaction import list_indices, synthetic_interaction
from sccm_amd.states import check_state, synthetic_state


def test_list_counts_match_reference_binaries():
    # record counts of Codes/int/{KB3G,GCN2850}/IMSRG_*_m1m2m3m4.dat and the sd value
    assert len(list_indices(make_basis("pf"))) == 14442
    assert len(list_indices(make_basis("f5pg9"))) == 10330
    assert len(list_indices(make_basis("sd"))) == 2570


def test_basis_order():
    b = make_basis("pf")
    assert b.n == 40
    assert list(b.twom[:8]) == [7, 5, 3, 1, -1, -3, -5, -7]
    assert (b.t[:20] == -1).all() and (b.t[20:] == 1).all()
    assert [s for (_, s, _) in b.blocks()] == [8, 4, 6, 2] * 2


def test_synthetic_interaction_is_rotation_invariant():
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 3, 4)
    F = Fock(b)
    H = F.hamiltonian(inter)
    for J in (F.Jy, F.Jp):
        C = H @ J - J @ H
        assert abs(C).max() < 1e-12
    assert abs(H - H.conj().T).max() < 1e-12


def test_synthetic_states_are_bogoliubov():
    b = make_basis("pf")
    for seed in (1, 2):
        U, V = synthetic_state(b, 2, 4, seed=seed)
        unit, skew = check_state(U, V)
        assert unit < 1e-13 and skew < 1e-10
        rho = V @ V.T
        assert 1.5 < np.trace(rho[:20, :20]) < 4 and 3 < np.trace(rho[20:, 20:]) < 6


def test_workload_shapes():
    w = configs.sd_plain()
    assert w.mesh_points == 50176 and w.n == 24
    w1 = configs.pf_gcm(nstates=20)
    assert len(w1.pairs) == 210 and w1.mesh.n_points == 162000
    assert w1.mesh_points == 34020000


def test_converter_reproduces_reference_mscheme_list():
    """Known answer: uncoupling the reference's J-scheme GCN2850 file
    (Codes/int/GCN2850/SM_GCN2850_ME2B.dat) reproduces every record of its
    binary m-scheme list IMSRG_s000_GCN2850pf5g9_hwHO10_m1m2m3m4.dat."""
    import os

    import ge76
    from sccm_amd.interaction import parse_jscheme_tbme, uncouple
    b = make_basis("f5pg9")
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gcn2850")
    table = parse_jscheme_tbme(open(os.path.join(here, "SM_GCN2850_ME2B.dat")).read(), b)
    me = uncouple(b, table)
    _, inter, _, _, _, _ = ge76.load()
    assert me.size == inter.me.size == 10330
    assert np.abs(me - inter.me).max() < 1e-9


def test_onishi_norm_is_refused_before_any_device_call():
    """The HIP path's norm is the Pfaffian overlap (the reference's branch for
    shell-model interactions); a mesh asking for the Onishi norm is refused
    with an error, not computed with the other formula."""
    import dataclasses
    import pytest
    from sccm_amd.basis import make_basis
    from sccm_amd.interaction import synthetic_interaction
    from sccm_amd.mesh import make_mesh
    from sccm_amd.pnamp import PNAMP, PnampError
    b = make_basis("p")
    mesh = dataclasses.replace(make_mesh(1, 1, 1, 1, 0), norm_method=1)
    with pytest.raises(PnampError, match="Onishi"):
        PNAMP(b, synthetic_interaction(b, 1.0, 1, 2), mesh)
