"""HIP path (through the C-ABI) against the oracle and the reference goldens.

Tolerances: the plain FP64 path must match the oracle to 1e-10 relative
(normwise per observable block), the Ge76 reference kernels to their 8 printed
digits, and exact Fock-space projections to 1e-12."""
import numpy as np
import pytest

import ge76
from sccm_amd import configs
from sccm_amd.basis import make_basis
from sccm_amd.interaction import synthetic_interaction
from sccm_amd.mesh import Mesh, make_mesh
from sccm_amd.states import synthetic_state

pytestmark = pytest.mark.gpu
REL = 1e-10


def _gpu(*a, **k):
    from sccm_amd.pnamp import PNAMP
    return PNAMP(*a, **k)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _blockwise(k_gpu, k_or):
    """max over (obs, J) of ||gpu - oracle|| / ||oracle|| on each J block"""
    worst = 0.0
    for io in range(k_or.shape[0]):
        for J in range(k_or.shape[1]):
            ref = k_or[io, J]
            nr = np.abs(ref).max()
            if nr == 0:
                assert np.abs(k_gpu[io, J]).max() == 0
                continue
            worst = max(worst, np.abs(k_gpu[io, J] - ref).max() / nr)
    return worst


@pytest.mark.parametrize("space,zn,mesh_args,fold", [
    ("sd", (2, 2), (3, 4, 3, 3, 4), 0),
    ("sd", (2, 2), (4, 4, 4, 3, 4), 1),
    ("pf", (2, 4), (2, 3, 2, 3, 6), 0),
    ("f5pg9", (4, 2), (2, 2, 3, 2, 8), 0),
])
def test_pair_matches_oracle(oracle_mod, space, zn, mesh_args, fold):
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.2, 101, 102)
    mesh = make_mesh(*mesh_args, fold=fold)
    Us, Vs = [], []
    for s in (11, 12):
        U, V = synthetic_state(b, *zn, seed=s)
        Us.append(U)
        Vs.append(V)
    g = _gpu(b, inter, mesh)
    g.set_states(np.array(Us), np.array(Vs))
    o = oracle_mod.Oracle(b, inter, mesh, *zn)
    for (p, q) in ((0, 1), (1, 1)):
        zr_g, k_g = g.kernel_pair(p, q, *zn, want_zres=True)
        zr_o, k_o = o.kernel_pair(Us[p], Vs[p], Us[q], Vs[q])
        for io in range(zr_o.shape[1]):
            assert _rel(zr_g[:, io], zr_o[:, io]) < REL, (space, p, q, io)
        assert _blockwise(k_g, k_o) < REL
    g.close()


def test_batched_equals_single(oracle_mod):
    w = configs.pf_gcm(nstates=4, nalp=2, nbet=2, nphi=3, jmax=4)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    kb = g.kernels_batched(w.pairs, w.zval, w.nval)
    for i, (p, q) in enumerate(w.pairs):
        ks = g.kernel_pair(p, q, w.zval, w.nval)
        assert np.array_equal(ks, kb[i])
    g.close()


def test_ge76_goldens_on_gpu(oracle_mod):
    b, inter, U, V, betas, kern = ge76.load()
    mesh = ge76.golden_mesh()
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    dets = g.state_overlaps()
    for q in range(len(betas)):
        _, d = oracle_mod.zbareover(U[q], V[q])
        assert abs(dets[q] / d - 1) < 1e-12
    strict = [i for i, x in enumerate(betas) if x != -0.2]
    pairs = [(p, q) for p in strict for q in strict]
    kb = g.kernels_batched(pairs, ge76.ZVAL, ge76.NVAL)
    jm = mesh.jmax
    for (p, q), k in zip(pairs, kb):
        gold = kern[f"{betas[p]:+.1f},{betas[q]:+.1f}"]
        n0, h0 = gold["0"][0], gold["0"][2]
        for J in (0, 2, 4, 6):
            N, H = k[0, J, jm, jm].real, k[4, J, jm, jm].real
            assert abs(N - gold[str(J)][0]) <= 2e-7 * max(abs(gold[str(J)][0]), 1e-6 * abs(n0))
            assert abs(H - gold[str(J)][2]) <= 2e-7 * max(abs(gold[str(J)][2]), 1e-6 * abs(h0))
            j2 = (k[5, J, jm, jm] / k[0, J, jm, jm]).real
            assert abs(j2 - gold[str(J)][8]) <= 2e-7 * max(abs(gold[str(J)][8]), 10.0 * abs(n0 / N))
    o = oracle_mod.Oracle(b, inter, mesh, ge76.ZVAL, ge76.NVAL)
    _, ko = o.kernel_pair(U[1], V[1], U[3], V[3])
    assert _blockwise(kb[pairs.index((1, 3))], ko) < REL
    g.close()


@pytest.mark.parametrize("form", [0, 1])
def test_exact_projection_on_gpu(form):
    from fock import Fock
    b = make_basis("p")
    inter = synthetic_interaction(b, 1.0, 11, 12)
    U0, V0 = synthetic_state(b, 2, 2, seed=5, sigma=0.4, sigma_pair=0.3)
    U1, V1 = synthetic_state(b, 2, 2, seed=6, sigma=0.4, sigma_pair=0.3)
    na, jmax = 12, 4
    alp = np.arange(na) * 2 * np.pi / na
    w = np.full(na, 2 * np.pi / na)
    x, wb = np.polynomial.legendre.leggauss(12)
    m = Mesh(alp, w, np.arccos(x), wb, alp.copy(), w.copy(), 7, jmax, 0, 0)
    g = _gpu(b, inter, m)
    g.set_states(np.array([U0, U1]), np.array([V0, V1]))
    g.set_form(form)
    kern = g.kernel_pair(0, 1, 2, 2)
    F = Fock(b)
    v0, v1 = F.hfb_state(U0, V0), F.hfb_state(U1, V1)
    sec, P = F.projectors(2, 2, jmax)
    H = F.hamiltonian(inter)[sec][:, sec].toarray()
    for (J, K1, K2), Pm in P.items():
        assert abs(np.vdot(v0[sec], Pm @ v1[sec]) - kern[0, J, K1 + jmax, K2 + jmax]) < 1e-12
        assert abs(np.vdot(v0[sec], H @ (Pm @ v1[sec])) - kern[4, J, K1 + jmax, K2 + jmax]) < 1e-11
    for J in range(jmax + 1):  # <J^2> = J(J+1) on the projected kernels
        assert np.abs(kern[5, J] - J * (J + 1) * kern[0, J]).max() < 1e-11
    g.close()


def test_hw_solve_matches_oracle(oracle_mod):
    rng = np.random.default_rng(4)
    dim = 30
    G = rng.normal(size=(dim, 12))
    N = G @ G.T + 1e-10 * np.eye(dim)
    Hs = rng.normal(size=(dim, dim))
    H = Hs + Hs.T
    b = make_basis("sd")
    g = _gpu(b, synthetic_interaction(b, 1.0, 1, 2), make_mesh(1, 1, 1, 1, 0))
    r = g.hw_solve(N, H, 1e-6)
    ro = oracle_mod.hndiag(N, H, 1e-6)
    assert r["ifl"] == 0 and r["m"] == ro["m"] == 12
    assert np.abs(r["E"] - ro["E"]).max() < 1e-10 * np.abs(ro["E"]).max()
    assert np.abs(r["EN"] - ro["EN"]).max() < 1e-10 * ro["EN"][0]
    for k in range(r["m"]):  # amplitudes up to sign
        a, c = r["F"][:, k], ro["F"][:, k]
        s = np.sign(a @ c)
        assert np.abs(a - s * c).max() < 1e-8 * np.abs(c).max()
    g.close()


def test_singular_state_is_reported():
    from sccm_amd.pnamp import PnampError
    b = make_basis("sd")
    U, V = synthetic_state(b, 2, 2, seed=3)
    U[:, 5] = 0.0
    g = _gpu(b, synthetic_interaction(b, 1.0, 1, 2), make_mesh(1, 1, 1, 1, 0))
    with pytest.raises(PnampError) as e:
        g.set_states(U, V)
    assert e.value.code > 0
    g.close()


@pytest.mark.parametrize("space,zn,mesh_args", [
    ("p32", (2, 2), (3, 4, 3, 3, 4)),
    ("p", (2, 2), (3, 4, 3, 5, 4)),
    ("sd", (2, 2), (3, 4, 3, 5, 6)),
    ("pf", (2, 4), (2, 3, 3, 9, 8)),
    ("f5pg9", (4, 2), (2, 2, 2, 3, 4)),
])
def test_factored_form_matches_plain_and_oracle(oracle_mod, space, zn, mesh_args):
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(*mesh_args)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (21, 22, 23)])
    U, V = np.array(Us), np.array(Vs)
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    assert g.separable
    pairs = [(0, 1), (1, 2), (2, 2), (0, 0)]
    g.set_form(FORM_PLAIN)
    zp, kp = g.kernels_batched(pairs, *zn, want_zres=True)
    g.set_form(FORM_FACTORED)
    zf, kf = g.kernels_batched(pairs, *zn, want_zres=True)
    for i in range(len(pairs)):
        for io in range(zp.shape[2]):
            assert _rel(zf[i][:, io], zp[i][:, io]) < REL, (space, pairs[i], io)
        assert _blockwise(kf[i], kp[i]) < REL
    o = oracle_mod.Oracle(b, inter, mesh, *zn)
    _, ko = o.kernel_pair(U[0], V[0], U[1], V[1])
    assert _blockwise(kf[0], ko) < REL
    g.close()


@pytest.mark.parametrize("space,zn,nphi", [("sd", (2, 2), 2), ("sd", (2, 2), 4), ("sd", (2, 2), 12),
                                            ("sd", (2, 2), 16), ("f5pg9", (4, 2), 16)])
def test_factored_form_at_every_gauge_count_class(space, zn, nphi):
    """The combine kernel is instantiated per gauge count (1..16; more than
    nine angles run at two waves per SIMD): even counts and the largest,
    against the plain form."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(2, 3, 2, nphi, 4)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (31, 32)])
    g = _gpu(b, inter, mesh)
    g.set_states(np.array(Us), np.array(Vs))
    pairs = [(0, 1), (1, 1)]
    g.set_form(FORM_PLAIN)
    zp, kp = g.kernels_batched(pairs, *zn, want_zres=True)
    g.set_form(FORM_FACTORED)
    zf, kf = g.kernels_batched(pairs, *zn, want_zres=True)
    for i in range(len(pairs)):
        for io in range(zp.shape[2]):
            assert _rel(zf[i][:, io], zp[i][:, io]) < REL, (space, nphi, pairs[i], io)
        assert _blockwise(kf[i], kp[i]) < REL
    g.close()


# f-p spaces of every species size the factored form and the plain size
# classes 16..40 have between the named spaces (h = 8, 10, 14, 16, 18; the
# Pfaffian's 2D layout at T = 2..5 slot rows, plain classes 16, 24, 32, 40)
FP_SPACES = {
    "f7": [(0, 3, 7)],
    "f7p1": [(0, 3, 7), (1, 1, 1)],
    "f7p": [(0, 3, 7), (1, 1, 3), (1, 1, 1)],
    "f7f5p1": [(0, 3, 7), (0, 3, 5), (1, 1, 1)],
    "f7f5p3": [(0, 3, 7), (0, 3, 5), (1, 1, 3)],
}


@pytest.mark.parametrize("space", list(FP_SPACES))
def test_every_species_size_matches_oracle(oracle_mod, space):
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis(space, FP_SPACES[space])
    zn = (2, 4)
    inter = synthetic_interaction(b, 1.2, 211, 212)
    mesh = make_mesh(2, 3, 2, 5, 4)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (51, 52)])
    U, V = np.array(Us), np.array(Vs)
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    pairs = [(0, 1), (1, 1)]
    g.set_form(FORM_PLAIN)
    zp, kp = g.kernels_batched(pairs, *zn, want_zres=True)
    g.set_form(FORM_FACTORED)
    zf, kf = g.kernels_batched(pairs, *zn, want_zres=True)
    o = oracle_mod.Oracle(b, inter, mesh, *zn)
    for i, (p, q) in enumerate(pairs):
        zo, ko = o.kernel_pair(U[p], V[p], U[q], V[q])
        for io in range(zo.shape[1]):
            assert _rel(zp[i][:, io], zo[:, io]) < REL, (space, "plain", pairs[i], io)
            assert _rel(zf[i][:, io], zo[:, io]) < REL, (space, "factored", pairs[i], io)
        assert _blockwise(kp[i], ko) < REL
        assert _blockwise(kf[i], ko) < REL
    g.close()


@pytest.mark.parametrize("cscale,iscale", [(10.0, 1), (0.5, 0), (1.0, -1)])
def test_overflow_scaling_matches_oracle(oracle_mod, cscale, iscale):
    """Input%Cscale multiplies the Pfaffian's matrix (Codes/PNAMP/PNAMP_Proj.f90:206)
    and Input%iscale the vacuum normalisations by 10^(-n iscale / 2)
    (Codes/PNAMP/PNAMP_Kernel3D.f90:147-151, 291-294): both forms against the
    oracle with the same scaling."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis("sd")
    zn = (2, 2)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(2, 3, 2, 5, 4, cscale=cscale, iscale=iscale)
    # well-conditioned states (cond(U) about 40): the vacuum normalisation
    # det(1 + Z^+ Z), Z = V U^-1, carries cond(U)^2 eps, so two correct
    # evaluations of it differ by 6e-10 already at cond(U) = 2.4e3 (seed 62)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (21, 22)])
    U, V = np.array(Us), np.array(Vs)
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    o = oracle_mod.Oracle(b, inter, mesh, *zn)
    zo, ko = o.kernel_pair(U[0], V[0], U[1], V[1])
    for form in (FORM_PLAIN, FORM_FACTORED):
        g.set_form(form)
        zr, k = g.kernel_pair(0, 1, *zn, want_zres=True)
        for io in range(zo.shape[1]):
            assert _rel(zr[:, io], zo[:, io]) < REL, (cscale, iscale, form, io)
        assert _blockwise(k, ko) < REL
    g.close()


@pytest.mark.parametrize("jmax", [9, 10, 12])
def test_high_spin_projection_matches_oracle(oracle_mod, jmax):
    """J up to the library's limit of 12: the separable projection kernel up
    to jmax = 9, the general one above, both forms against the oracle."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis("f5pg9")
    zn = (4, 2)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(3, 5, 3, 3, jmax)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (21, 22)])
    U, V = np.array(Us), np.array(Vs)
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    o = oracle_mod.Oracle(b, inter, mesh, *zn)
    zo, ko = o.kernel_pair(U[0], V[0], U[1], V[1])
    assert ko.shape[1] == jmax + 1
    for form in (FORM_PLAIN, FORM_FACTORED):
        g.set_form(form)
        zr, k = g.kernel_pair(0, 1, *zn, want_zres=True)
        for io in range(zo.shape[1]):
            assert _rel(zr[:, io], zo[:, io]) < REL, (jmax, form, io)
        assert _blockwise(k, ko) < REL, (jmax, form)
    g.close()


def test_factored_form_rejects_mixed_states():
    from sccm_amd.pnamp import FORM_FACTORED, PnampError
    b, inter, U, V, betas, kern = ge76.load()
    g = _gpu(b, inter, ge76.golden_mesh())
    g.set_states(U[:2], V[:2])
    assert not g.separable
    g.set_form(FORM_FACTORED)
    with pytest.raises(PnampError):
        g.kernel_pair(0, 1, ge76.ZVAL, ge76.NVAL)
    g.close()


def _modified_sd_interaction(kind):
    """sd interaction with one list entry changed: "unordered" swaps a and b
    of a proton-proton entry (a > b) and negates its value (the same
    antisymmetrized element), "neutron_first" writes a proton-neutron entry
    as (b, a, d, c), "pn_m" moves d of a proton-neutron entry to a state of
    another m (the entry no longer conserves M)."""
    import dataclasses
    b = make_basis("sd")
    inter = synthetic_interaction(b, 1.5, 1001, 1002)
    ka, kb, kc, kd = (np.array(x, copy=True) for x in (inter.ka, inter.kb, inter.kc, inter.kd))
    h = b.n // 2
    me = np.array(inter.me, copy=True)
    if kind == "unordered":
        q = int(np.flatnonzero((ka <= h) & (kb <= h) & (kc <= h) & (kd <= h))[0])
        ka[q], kb[q] = kb[q], ka[q]
        me[q] = -me[q]
    elif kind == "neutron_first":
        for q in np.flatnonzero((ka <= h) & (kb > h) & (kc <= h) & (kd > h))[:50]:
            ka[q], kb[q], kc[q], kd[q] = kb[q], ka[q], kd[q], kc[q]
    else:
        q = int(np.flatnonzero((ka <= h) & (kb > h) & (kc <= h) & (kd > h))[0])
        d = int(kd[q]) - 1
        other = [s for s in range(h, b.n) if b.twom[s] != b.twom[d]]
        kd[q] = other[0] + 1
    return b, dataclasses.replace(inter, ka=ka, kb=kb, kc=kc, kd=kd, me=me)


@pytest.mark.parametrize("kind", ["unordered", "neutron_first"])
def test_reordered_list_entries_give_the_same_kernels(kind):
    """Entries written in another pair order (with the antisymmetry sign) are
    the same interaction: the plain form takes them as they are, the factored
    form brings them to canonical order; both equal the canonical list."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b, inter = _modified_sd_interaction(kind)
    ref = synthetic_interaction(b, 1.5, 1001, 1002)
    mesh = make_mesh(2, 3, 2, 3, 2)
    U, V = configs.make_states(b, 2, 2, configs.gcm_grid(3))
    out = {}
    for name, it in (("mod", inter), ("ref", ref)):
        g = _gpu(b, it, mesh)
        g.set_states(U, V)
        for form in (FORM_PLAIN, FORM_FACTORED):
            g.set_form(form)
            out[(name, form)] = g.kernel_pair(0, 1, 2, 2)
        g.close()
    r = out[("ref", FORM_PLAIN)]
    for key, k in out.items():
        assert _blockwise(k, r) < REL, key


def test_separability_does_not_depend_on_the_quasiparticle_basis():
    """U -> U C, V -> V C with C orthogonal and mixing proton and neutron
    columns leaves the state (and Z = V U^-1) unchanged: it is found
    separable, and the factored kernels equal the plain ones and those of the
    unrotated states."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    b = make_basis("pf")
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(2, 3, 2, 5, 4)
    U, V = configs.make_states(b, 2, 4, configs.gcm_grid(3))
    C, _ = np.linalg.qr(np.random.default_rng(9).normal(size=(b.n, b.n)))
    Ur, Vr = U @ C, V @ C
    h = b.n // 2  # some quasiparticle columns now have proton and neutron components
    assert ((np.abs(Ur[0][:h]).max(0) > 1e-3) & (np.abs(Ur[0][h:]).max(0) > 1e-3)).any()
    g = _gpu(b, inter, mesh)
    g.set_states(U, V)
    g.set_form(FORM_FACTORED)
    k0 = g.kernels_batched([(0, 1), (2, 2)], 2, 4)
    g.set_states(Ur, Vr)
    assert g.separable
    kf = g.kernels_batched([(0, 1), (2, 2)], 2, 4)
    g.set_form(FORM_PLAIN)
    kp = g.kernels_batched([(0, 1), (2, 2)], 2, 4)
    for i in range(2):
        assert _blockwise(kf[i], kp[i]) < REL and _blockwise(kf[i], k0[i]) < REL
    g.close()


@pytest.mark.parametrize("kind", ["pn_m"])
def test_factored_form_refuses_lists_it_cannot_split(kind):
    """The factored forms need proton-neutron entries that conserve M; other
    lists run in the plain form and are refused by the factored one with an
    error, never computed wrongly."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN, PnampError
    b, inter = _modified_sd_interaction(kind)
    g = _gpu(b, inter, make_mesh(2, 3, 2, 3, 2))
    U, V = configs.make_states(b, 2, 2, configs.gcm_grid(3))
    g.set_states(U, V)
    assert g.separable
    g.set_form(FORM_PLAIN)
    zr, kern = g.kernel_pair(0, 1, 2, 2, want_zres=True)
    assert np.all(np.isfinite(zr))
    g.set_form(FORM_FACTORED)
    with pytest.raises(PnampError):
        g.kernel_pair(0, 1, 2, 2)
    g.close()


def test_factored_graph_replays_match():
    """The factored form replays a captured hipGraph on calls of the same
    shape; every replay, with eager work of the plain form in between, must
    equal the first (capturing) call bit for bit."""
    from sccm_amd.pnamp import FORM_FACTORED, FORM_PLAIN
    w = configs.pf_gcm(nstates=4, nalp=2, nbet=3, nphi=5, jmax=4)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(FORM_FACTORED)
    first = g.kernels_batched(w.pairs, w.zval, w.nval)
    for rep in range(3):
        g.set_form(FORM_PLAIN)
        g.kernels_batched(w.pairs[:3], w.zval, w.nval)
        g.set_form(FORM_FACTORED)
        again = g.kernels_batched(w.pairs, w.zval, w.nval)
        assert np.array_equal(again, first), rep
    g.close()


def test_factored_graph_survives_set_states():
    """New states of the same count are written into the same buffers: the
    captured graph is replayed, not captured again, and gives the kernels of
    a fresh context with those states."""
    from sccm_amd.pnamp import FORM_FACTORED
    w = configs.pf_gcm(nstates=4, nalp=2, nbet=3, nphi=5, jmax=4)
    U2, V2 = configs.make_states(w.basis, w.zval, w.nval, w.coords, seed=777)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(FORM_FACTORED)
    g.kernels_batched(w.pairs, w.zval, w.nval)
    g.set_states(U2, V2)
    k2 = g.kernels_batched(w.pairs, w.zval, w.nval)
    assert g.graph_info()[:2] == (1, 2)
    f = _gpu(w.basis, w.inter, w.mesh)
    f.set_states(U2, V2)
    f.set_form(FORM_FACTORED)
    assert np.array_equal(f.kernels_batched(w.pairs, w.zval, w.nval), k2)
    g.close()
    f.close()


def test_autotune_keeps_the_kernels():
    """The autotuner times every launch candidate on real pairs and keeps the
    fastest; the candidates differ only in occupancy and chunking, so the
    kernels after tuning equal those before."""
    from sccm_amd.pnamp import FORM_FACTORED
    w = configs.pf_gcm(nstates=4, nalp=2, nbet=3, nphi=5, jmax=4)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(FORM_FACTORED)
    before = g.kernels_batched(w.pairs, w.zval, w.nval)
    t = g.autotune(w.pairs, w.zval, w.nval)
    # h = 20: the spill-free Pfaffian variants (3 and 4 waves; 5 spills) x chunk divisors 1, 2, 4
    assert [(c["pf_waves"], c["chunk_div"]) for c in t["candidates"]] == [(p, d) for p in (3, 4) for d in (1, 2, 4)]
    assert all(c["ms"] > 0 and c["rest_waves"] == 3 for c in t["candidates"])
    best = min(t["candidates"], key=lambda c: c["ms"])
    assert (best["pf_waves"], best["chunk_div"]) == (t["pf_waves"], t["chunk_div"])
    after = g.kernels_batched(w.pairs, w.zval, w.nval)
    assert np.abs(after - before).max() <= 1e-14 * np.abs(before).max()
    # with the Schur-complement Pfaffian allowed it is candidate 0 as well
    g.set_schur_pfaffian(True)
    t = g.autotune(w.pairs, w.zval, w.nval)
    assert [c["pf_waves"] for c in t["candidates"]][:3] == [0, 0, 0]
    g.close()


@pytest.mark.parametrize("space,zn", [("sd", (2, 2)), ("f5pg9", (4, 2))])
def test_every_launch_variant_gives_the_same_kernels(space, zn):
    """Each launch configuration pnamp_set_tune accepts (Pfaffian at 3, 4, 5
    waves per SIMD, densities at 4 where h <= 16, chunk divisors 1, 2, 4)
    computes the same kernels as the default: the variants differ only in
    register budget and chunking."""
    from sccm_amd.pnamp import FORM_FACTORED
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (41, 42, 43)])
    g = _gpu(b, inter, make_mesh(2, 3, 2, 5, 4))
    g.set_states(np.array(Us), np.array(Vs))
    g.set_form(FORM_FACTORED)
    pairs = [(0, 1), (1, 2), (2, 2)]
    ref = g.kernels_batched(pairs, *zn)
    for pf in (3, 4, 5):
        for rest in (0, 4):
            for div in (1, 2, 4):
                g.set_tune(pf, rest, div)
                k = g.kernels_batched(pairs, *zn)
                assert np.abs(k - ref).max() <= 1e-13 * np.abs(ref).max(), (space, pf, rest, div)
    g.close()


def _states_from_z(b, zs, seed):
    """A separable HFB state built from canonical pair amplitudes: per species
    Z_s = O diag([[0, z_k], [-z_k, 0]]) O^T (O random orthogonal), U_s = (1 +
    Z_s^T Z_s)^(-1/2), V_s = Z_s U_s; cond(Z_s) = max z / min z."""
    rng = np.random.default_rng(seed)
    h = b.n // 2
    U = np.zeros((b.n, b.n))
    V = np.zeros((b.n, b.n))
    for s in range(2):
        O, _ = np.linalg.qr(rng.normal(size=(h, h)))
        D = np.zeros((h, h))
        for k, z in enumerate(zs):
            D[2 * k, 2 * k + 1], D[2 * k + 1, 2 * k] = z, -z
        Z = O @ D @ O.T
        w, Q = np.linalg.eigh(np.eye(h) + Z.T @ Z)
        Us = Q @ np.diag(w ** -0.5) @ Q.T
        U[s * h:(s + 1) * h, s * h:(s + 1) * h] = Us
        V[s * h:(s + 1) * h, s * h:(s + 1) * h] = Z @ Us
    return U, V


@pytest.mark.parametrize("space,zn", [("sd", (2, 2)), ("pf", (2, 4)), ("f5pg9", (4, 2))])
def test_schur_pfaffian_route(space, zn):
    """The opt-in h x h Schur-complement Pfaffian: off by default (the kernels
    of a context that never enabled it are the 2h x 2h route's), on for
    well-conditioned states within its stated rounding of the 2h x 2h route
    and of the oracle chain's inputs, off again (exact) for states with
    cond_1(Z_s) above 1e6."""
    from sccm_amd.pnamp import FORM_FACTORED
    b = make_basis(space)
    inter = synthetic_interaction(b, 1.2, 201, 202)
    g = _gpu(b, inter, make_mesh(2, 3, 2, 5, 4))
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (41, 42, 43)])
    g.set_states(np.array(Us), np.array(Vs))
    g.set_form(FORM_FACTORED)
    pairs = [(0, 1), (1, 2), (2, 2)]
    full = g.kernels_batched(pairs, *zn)
    assert g.schur_pfaffian_active == 0
    g.set_schur_pfaffian(True)
    assert g.schur_pfaffian_active == 2
    sch = g.kernels_batched(pairs, *zn)
    d = np.abs(sch - full).max() / np.abs(full).max()
    assert 0.0 < d < 1e-8, d
    for i in range(len(pairs)):
        assert _blockwise(sch[i], full[i]) < 2e-8
    g.set_schur_pfaffian(False)
    assert np.array_equal(g.kernels_batched(pairs, *zn), full)
    # ill-conditioned Z_s (pair amplitudes 1e-7 .. 1): the 2h x 2h route stays
    h = b.n // 2
    zs = np.geomspace(1e-7, 1.0, h // 2)
    U2, V2 = zip(*[_states_from_z(b, zs, seed=s) for s in (7, 8)])
    g.set_schur_pfaffian(True)
    g.set_states(np.array(U2), np.array(V2))
    assert g.separable and g.schur_pfaffian_active == 0
    k_on = g.kernels_batched([(0, 1)], 2, 2)
    g.set_schur_pfaffian(False)
    assert np.array_equal(g.kernels_batched([(0, 1)], 2, 2), k_on)
    # a mixed set: pairs whose bra state passes take the Schur route, the
    # others keep the 2h x 2h Pfaffian (both kernels over one grid)
    g.set_states(np.array([Us[0], U2[0], Us[1]]), np.array([Vs[0], V2[0], Vs[1]]))
    mixed_pairs = [(0, 2), (1, 0), (2, 1), (1, 1)]
    k_full = g.kernels_batched(mixed_pairs, *zn)
    g.set_schur_pfaffian(True)
    assert g.schur_pfaffian_active == 1
    k_mix = g.kernels_batched(mixed_pairs, *zn)
    for i, (p, q) in enumerate(mixed_pairs):
        if p == 1:  # ill-conditioned bra: the 2h x 2h route, bit for bit
            assert np.array_equal(k_mix[i], k_full[i]), (p, q)
        else:
            assert 0.0 < _blockwise(k_mix[i], k_full[i]) < 2e-8, (p, q)
    g.close()


# Mixed forms: the V contraction on MFMA in FP32 or split BF16.  Observables
# without the two-body contraction (norm, Z, N, E_kin, <J^2>) stay FP64 and
# match the factored form to 1e-10; the energy kernel carries the GEMM
# rounding: FP32 operands (2^-24) and split BF16 (about 2^-17 per operand)
# with FP32 accumulation over h^2 terms.
MIX_TOL = {2: 2e-5, 3: 5e-4}


@pytest.mark.parametrize("form", [2, 3])
@pytest.mark.parametrize("space,zn,mesh_args", [
    ("p32", (2, 2), (3, 4, 3, 3, 4)),
    ("p", (2, 2), (3, 4, 3, 5, 4)),
    *[(s, (2, 4), (2, 3, 2, 5, 4)) for s in FP_SPACES],
    ("sd", (2, 2), (3, 4, 3, 5, 6)),
    ("pf", (2, 4), (2, 3, 3, 9, 8)),
    ("f5pg9", (4, 2), (2, 2, 2, 3, 4)),
])
def test_mixed_forms_match_factored(space, zn, mesh_args, form):
    from sccm_amd.pnamp import FORM_FACTORED
    b = make_basis(space, FP_SPACES.get(space))
    inter = synthetic_interaction(b, 1.2, 201, 202)
    mesh = make_mesh(*mesh_args)
    Us, Vs = zip(*[synthetic_state(b, *zn, seed=s) for s in (21, 22, 23)])
    g = _gpu(b, inter, mesh)
    g.set_states(np.array(Us), np.array(Vs))
    pairs = [(0, 1), (1, 2), (2, 2), (0, 0)]
    g.set_form(FORM_FACTORED)
    zf, kf = g.kernels_batched(pairs, *zn, want_zres=True)
    g.set_form(form)
    zm, km = g.kernels_batched(pairs, *zn, want_zres=True)
    worst_e = 0.0
    for i in range(len(pairs)):
        for io in (0, 1, 2, 3, 5):
            assert _rel(zm[i][:, io], zf[i][:, io]) < REL, (space, pairs[i], io)
        e = _rel(zm[i][:, 4], zf[i][:, 4])
        worst_e = max(worst_e, e)
        assert e < MIX_TOL[form], (space, pairs[i], e)
        for io in (0, 1, 2, 3, 5):
            assert _blockwise(km[i][io:io + 1], kf[i][io:io + 1]) < REL
        assert _blockwise(km[i][4:5], kf[i][4:5]) < MIX_TOL[form]
    print(f"{space} form {form}: max rel error of the energy ZRES {worst_e:.3e}")
    g.close()


@pytest.mark.parametrize("form", [2, 3])
def test_mixed_graph_replays_match(form):
    from sccm_amd.pnamp import FORM_PLAIN
    w = configs.pf_gcm(nstates=4, nalp=2, nbet=3, nphi=5, jmax=4)
    g = _gpu(w.basis, w.inter, w.mesh)
    g.set_states(w.U, w.V)
    g.set_form(form)
    first = g.kernels_batched(w.pairs, w.zval, w.nval)
    for rep in range(2):
        g.set_form(FORM_PLAIN)
        g.kernels_batched(w.pairs[:2], w.zval, w.nval)
        g.set_form(form)
        again = g.kernels_batched(w.pairs, w.zval, w.nval)
        assert np.array_equal(again, first), rep
    g.close()
