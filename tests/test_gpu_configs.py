"""HIP path against the oracle at the benchmark workloads' own sizes.

* C0 (sd, n = 24, one state, alpha = gamma 8 x beta 16 x 7^2 gauge, J <= 8,
  50,176 mesh points) in full, plain and factored forms.
* C1 (pf, n = 40): one full pair of the 20-state GCM (alpha = gamma 10 x
  beta 20 x 9^2 gauge = 162,000 mesh points, J <= 8), plain and factored.
* C2: the mixed forms (FP32, split BF16 V contraction) against the oracle
  with the field-form energy (Method I, HFB_FIELD_COMPLEX + HFB_ENER_comp),
  per kernel and per Hill-Wheeler level of a reduced pf GCM.
* C4: one seeded pair of one instance of each model space (sd, pf, f5pg9) at
  the batch mesh, ZRES on a strided subset of beta nodes against the oracle.
* The GCM chain: a reduced pf GCM (5 states, 15 pairs) through the GPU
  kernels and pnamp_hw_solve against the oracle kernels and its HNDIAG.
* pnamp_hw_solve at dimension 340 (20 states x 17 K values, J = 8).

Tolerance 1e-10 relative (normwise per observable block) for every FP64
form; the mixed forms carry the stated MFMA rounding of the energy only.
"""
import numpy as np
import pytest

from sccm_amd import configs
from sccm_amd.driver import hw_matrices
from sccm_amd.mesh import Mesh, make_mesh

pytestmark = pytest.mark.gpu
REL = 1e-10


def _gpu(*a, **k):
    from sccm_amd.pnamp import PNAMP
    return PNAMP(*a, **k)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _blockwise(k_gpu, k_or, obs=None):
    worst = 0.0
    for io in (range(k_or.shape[0]) if obs is None else obs):
        for J in range(k_or.shape[1]):
            ref = k_or[io, J]
            nr = np.abs(ref).max()
            if nr == 0:
                continue
            worst = max(worst, np.abs(k_gpu[io, J] - ref).max() / nr)
    return worst


def _check_pair(zr_g, k_g, zr_o, k_o, tag):
    for io in range(zr_o.shape[1]):
        e = _rel(zr_g[:, io], zr_o[:, io])
        assert e < REL, (tag, "zres", io, e)
    e = _blockwise(k_g, k_o)
    assert e < REL, (tag, "kern", e)
    return e


@pytest.mark.parametrize("form", [0, 1])
def test_c0_full_mesh_matches_oracle(oracle_mod, form):
    w = configs.sd_plain()
    assert w.mesh.n_points == 50176
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(form)
    zr_g, k_g = g.kernel_pair(0, 0, w.zval, w.nval, want_zres=True)
    o = oracle_mod.Oracle(w.basis, w.inter, w.mesh, w.zval, w.nval)
    zr_o, k_o = o.kernel_pair(w.U[0], w.V[0], w.U[0], w.V[0])
    e = _check_pair(zr_g, k_g, zr_o, k_o, ("C0", form))
    print(f"C0 form {form}: 50176 mesh points, kernels vs oracle {e:.2e}")
    g.close()


def test_c1_full_pair_matches_oracle(oracle_mod):
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    w = configs.pf_gcm()
    p, q = 3, 11
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    o = oracle_mod.Oracle(w.basis, w.inter, w.mesh, w.zval, w.nval)
    zr_o, k_o = o.kernel_pair(w.U[p], w.V[p], w.U[q], w.V[q])
    for form in (FORM_PLAIN, FORM_FACTORED):
        g.set_form(form)
        zr_g, k_g = g.kernel_pair(p, q, w.zval, w.nval, want_zres=True)
        e = _check_pair(zr_g, k_g, zr_o, k_o, ("C1", form))
        print(f"C1 pair ({p},{q}) form {form}: 162000 mesh points, kernels vs oracle {e:.2e}")
    g.close()


def _reduced_gcm():
    return configs.pf_gcm(nstates=5, nalp=4, nbet=6, nphi=5, jmax=4)


def _oracle_kernels(oracle_mod, w, energy_method=0):
    o = oracle_mod.Oracle(w.basis, w.inter, w.mesh, w.zval, w.nval, energy_method=energy_method)
    return np.array([o.kernel_pair(w.U[p], w.V[p], w.U[q], w.V[q])[1] for p, q in w.pairs])


def _levels_gpu(g, kern, w, J):
    N, H, _, _ = hw_matrices(kern, w.pairs, len(w.U), J, w.mesh.jmax)
    return g.hw_solve(N, H, w.hw_eps)


def test_gcm_energies_match_oracle_chain(oracle_mod):
    """Kernels of all pairs and the Hill-Wheeler solve of every J: GPU
    (plain and factored kernels, pnamp_hw_solve) against the oracle (kernels,
    HNDIAG)."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    w = _reduced_gcm()
    k_o = _oracle_kernels(oracle_mod, w)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    for form in (FORM_PLAIN, FORM_FACTORED):
        g.set_form(form)
        k_g = g.kernels_batched(w.pairs, w.zval, w.nval)
        for i in range(len(w.pairs)):
            assert _blockwise(k_g[i], k_o[i]) < REL, (form, i)
        for J in range(w.mesh.jmax + 1):
            N, H, _, _ = hw_matrices(k_o, w.pairs, len(w.U), J, w.mesh.jmax)
            ro = oracle_mod.hndiag(N, H, w.hw_eps)
            rg = _levels_gpu(g, k_g, w, J)
            assert rg["ifl"] == ro["ifl"] and rg["m"] == ro["m"], (form, J)
            dE = np.abs(rg["E"] - ro["E"])
            assert (dE <= REL * np.maximum(np.abs(ro["E"]), 1.0)).all(), (form, J, dE.max())
            print(f"form {form} J={J}: {ro['m']} natural states, max |dE| {dE.max():.2e} MeV")
    g.close()


def test_hw_solve_dimension_340(oracle_mod):
    """HNDIAG at J = 8 of the 20-state GCM: dimension 20 x 17 = 340."""
    from sccm_amd.basis import make_basis
    from sccm_amd.interaction import synthetic_interaction
    rng = np.random.default_rng(340)
    dim = 340
    Q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
    en = np.concatenate([np.geomspace(1.0, 1e-4, 150), np.geomspace(1e-9, 1e-13, dim - 150)])
    N = (Q * en) @ Q.T
    Hs = rng.normal(size=(dim, dim))
    H = N @ (Hs + Hs.T) @ N / 10.0  # a Hamiltonian that lives in the range of N
    b = make_basis("sd")
    g = _gpu(b, synthetic_interaction(b, 1.0, 1, 2), make_mesh(1, 1, 1, 1, 0))
    r = g.hw_solve(N, H, 1e-6)
    ro = oracle_mod.hndiag(N, H, 1e-6)
    assert r["ifl"] == 0 and r["m"] == ro["m"] == 150
    assert np.abs(r["EN"][:150] - ro["EN"][:150]).max() < REL * ro["EN"][0]
    assert np.abs(r["E"] - ro["E"]).max() < REL * np.abs(ro["E"]).max()
    print(f"HNDIAG dim {dim}: {r['m']} kept, max |dE| {np.abs(r['E'] - ro['E']).max():.2e}")
    g.close()


# C2: the FP32 contraction keeps H^J within 1e-6 of the FP64 oracle and the
# levels within 1 keV; split BF16 within 2e-5 relative on H^J.
MIX_H_TOL = {2: 1e-6, 3: 2e-5}


@pytest.mark.parametrize("form", [2, 3])
def test_mixed_forms_vs_oracle_per_kernel_and_level(oracle_mod, form):
    w = _reduced_gcm()
    k_o = _oracle_kernels(oracle_mod, w, energy_method=1)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(form)
    k_m = g.kernels_batched(w.pairs, w.zval, w.nval)
    worst_n = max(_blockwise(k_m[i], k_o[i], obs=(0, 1, 2, 3, 5)) for i in range(len(w.pairs)))
    assert worst_n < REL
    for J in range(w.mesh.jmax + 1):
        eh = max(_blockwise(k_m[i][:, J:J + 1], k_o[i][:, J:J + 1], obs=(4,)) for i in range(len(w.pairs)))
        N, H, _, _ = hw_matrices(k_o, w.pairs, len(w.U), J, w.mesh.jmax)
        ro = oracle_mod.hndiag(N, H, w.hw_eps)
        rm = _levels_gpu(g, k_m, w, J)
        k = min(3, ro["m"], rm["m"])
        dE = 1e3 * np.abs(rm["E"][:k] - ro["E"][:k])
        print(f"form {form} J={J}: H^J rel err {eh:.2e}, |dE| of the lowest {k} levels (keV) {np.round(dE, 5)}")
        assert eh < MIX_H_TOL[form], (J, eh)
        if form == 2:
            assert (dE < 1.0).all(), (J, dE)
    g.close()


@pytest.mark.parametrize("space", ["sd", "pf", "f5pg9"])
def test_c4_sampled_pair_per_space(oracle_mod, space):
    """One seeded pair of the first batch instance of each model space, at
    the batch mesh (alpha = gamma 10, beta 20, 9 gauge, J <= 8) in the
    factored form; ZRES of beta nodes 0, 6, 13, 19 against the oracle."""
    from sccm_amd import batch
    from sccm_amd.basis import make_basis
    from sccm_amd.interaction import synthetic_interaction
    from sccm_amd.pnamp import FORM_FACTORED
    inst = next(i for i in batch.draw_instances() if i.space == space)
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.5, *batch.SEEDS[space])
    mesh = make_mesh(10, 20, 10, 9, 8)
    U, V, pairs = batch.instance_states(inst, b)
    p, q = pairs[np.random.default_rng(7).integers(len(pairs))]
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    g.set_form(FORM_FACTORED)
    zr_g, _ = g.kernel_pair(p, q, inst.zval, inst.nval, want_zres=True)
    nodes = [0, 6, 13, 19]
    sub = Mesh(mesh.alp, mesh.walp, mesh.bet[nodes], mesh.wbet[nodes], mesh.gam, mesh.wgam, mesh.nphi, mesh.jmax)
    o = oracle_mod.Oracle(b, inter, sub, inst.zval, inst.nval)
    zr_o, _ = o.kernel_pair(U[p], V[p], U[q], V[q])
    rows = np.concatenate([np.arange(ib * 100, (ib + 1) * 100) for ib in nodes])
    for io in range(zr_o.shape[1]):
        e = _rel(zr_g[rows, io], zr_o[:, io])
        assert e < REL, (space, io, e)
    print(f"C4 {space} Z={inst.zval} N={inst.nval} pair ({p},{q}): 400 Euler x 81 gauge points match the oracle")
    g.close()
