"""Host-side pieces around the library: the reference's PNAMP output layout,
the Hill-Wheeler (GCM) assembly, and the case file of the Fortran host.

* ``chop_kernels`` applies the CHOP rule of ``Kernel4Triaxial``
  (``Codes/PNAMP/PNAMP_Kernel3D.f90:729-737``) to the library's projected sums
  and returns the ``Kernel%njkk/hjkk/zz/nn`` quantities.
* ``hw_matrices`` assembles the (q, K) norm and norm-weighted Hamiltonian
  matrices of one J from the pair kernels (q <= q' computed, q > q' by
  hermiticity, as the reference's symmetric-pair mode ``ICons = 0`` implies,
  ``Codes/PNAMP/PNAMP_Main.f90:61-66``); ``gcm_energies`` solves them with the
  library's HNDIAG (``Codes/GCM/GCM_HNDiag.f90``).
* ``gcm_levels`` adds the natural-state selection of ``PreSolve_HWG``
  (``Codes/GCM/GCM_HWG.f90:2-100``): per J and level the number of natural
  states of least variance, the level there, the plateau around it and the
  norm spectrum; ``levels_at_nos`` evaluates another kernel set at those NOS.
* ``write_case`` / ``read_elem`` serve the Fortran host ``fortran/pnamp_driver``.
"""
from __future__ import annotations

import numpy as np

CHOP = 1e-8


def chop_kernels(kern: np.ndarray, ncore_p: int = 0, ncore_n: int = 0) -> dict:
    """kern [5, J+1, 2J+1, 2J+1] -> njkk, hjkk, zz, nn (zero where |N| < CHOP)."""
    N = kern[0]
    ok = np.abs(N) >= CHOP
    safe = np.where(ok, N, 1.0)
    return dict(njkk=np.where(ok, N, 0), hjkk=np.where(ok, kern[4], 0),
                zz=np.where(ok, kern[1] / safe + ncore_p, 0), nn=np.where(ok, kern[2] / safe + ncore_n, 0))


def hw_matrices(kern_pairs: np.ndarray, pairs: np.ndarray, nstates: int, J: int, jmax: int, kvals=None):
    """Real symmetric (N, H) over the basis (q, K), K in ``kvals`` (default -J..J)."""
    if kvals is None:
        kvals = list(range(-J, J + 1))
    nk = len(kvals)
    dim = nstates * nk
    N = np.zeros((dim, dim), dtype=complex)
    H = np.zeros((dim, dim), dtype=complex)
    for (q0, q1), k in zip(pairs, kern_pairs):
        for a, K1 in enumerate(kvals):
            for c, K2 in enumerate(kvals):
                nv = k[0, J, K1 + jmax, K2 + jmax]
                hv = k[4, J, K1 + jmax, K2 + jmax]
                N[q0 * nk + a, q1 * nk + c] = nv
                H[q0 * nk + a, q1 * nk + c] = hv
                if q0 != q1:
                    N[q1 * nk + c, q0 * nk + a] = np.conj(nv)
                    H[q1 * nk + c, q0 * nk + a] = np.conj(hv)
    N = 0.5 * (N + N.conj().T)
    H = 0.5 * (H + H.conj().T)
    return N.real.copy(), H.real.copy(), float(np.abs(N.imag).max()), float(np.abs(H.imag).max())


def gcm_energies(ctx, kern_pairs, pairs, nstates: int, jmax: int, eps: float = 1e-6, jlist=None,
                 nlevels: int = 5) -> dict:
    """{J: lowest ``nlevels`` Hill-Wheeler energies} (empty if no norm state)."""
    out = {}
    for J in (range(jmax + 1) if jlist is None else jlist):
        N, H, _, _ = hw_matrices(kern_pairs, pairs, nstates, J, jmax)
        if np.abs(N).max() < 1e-12:
            out[J] = np.zeros(0)
            continue
        r = ctx.hw_solve(N, H, eps, want_f=False)
        out[J] = r["E"][:nlevels] if r["ifl"] == 0 else np.zeros(0)
    return out


def gcm_levels(ctx, kern_pairs, pairs, nstates: int, jmax: int, eps: float = 1e-6, kmax: int = 3,
               jlist=None) -> dict:
    """Per J: the norm spectrum (natural states kept at ``eps``, the most
    negative EN(i)/EN(1), a sign of an unconverged projection), the levels of
    HNDIAG with every natural state above eps, and PreSolve_HWG's choice: per
    level k < kmax the NOS of least variance, the energy there and the largest
    change of that energy for NOS - 1 and NOS + 1 (the plateau)."""
    out = {}
    for J in (range(jmax + 1) if jlist is None else jlist):
        N, H, _, _ = hw_matrices(kern_pairs, pairs, nstates, J, jmax)
        if np.abs(N).max() < 1e-12:
            continue
        r = ctx.hw_presolve(N, H, eps, kmax)
        if r["ifl"]:
            continue
        M, Et = r["m"], r["E"]
        plateau = []
        for k in range(kmax):
            n = int(r["nos"][k])
            if n <= 0:
                plateau.append(None)
                continue
            nb = [Et[i - 1, k] for i in (n - 1, n + 1) if k + 1 <= i <= M]
            plateau.append(float(max(abs(x - Et[n - 1, k]) for x in nb)) if nb else 0.0)
        out[J] = dict(dim=N.shape[0], m=M, en_min_rel=float(r["EN"].min() / r["EN"][0]),
                      e_eps=[float(x) for x in Et[M - 1, :min(kmax, M)]],
                      nos=[int(x) for x in r["nos"]], e=[float(x) for x in r["e_best"]], plateau=plateau)
    return out


def levels_at_nos(ctx, kern_pairs, pairs, nstates: int, jmax: int, levels: dict, eps: float = 1e-6) -> dict:
    """Energies of ``kern_pairs`` at the NOS that ``levels`` (gcm_levels of
    another kernel set) chose per J and level, by HNDIAG with that NOS."""
    out = {}
    for J, lv in levels.items():
        N, H, _, _ = hw_matrices(kern_pairs, pairs, nstates, J, jmax)
        e = []
        for k, n in enumerate(lv["nos"]):
            if n <= 0:
                e.append(None)
                continue
            r = ctx.hw_solve(N, H, eps, nos=n, want_f=False)
            e.append(float(r["E"][k]) if r["ifl"] == 0 and r["m"] > k else None)
        out[J] = e
    return out


def write_case(path, basis, inter, mesh, U, V, pairs, coords, zval, nval, ncore=(0, 0), q2=None,
               echarge=(1.0, 0.0)):
    """Unformatted stream input of fortran/pnamp_driver; with ``q2`` (Q_{2mu}
    matrices [5, n, n]) and the effective charges (e_p, e_n) the host also
    computes the E2 kernels."""
    n = basis.n
    nst = U.shape[0]
    hdr = np.array([n, inter.nme, mesh.alp.size, mesh.bet.size, mesh.gam.size, mesh.nphi, mesh.jmax,
                    mesh.fold, mesh.iscale, nst, len(pairs), zval, nval, ncore[0], ncore[1],
                    0 if q2 is None else 1], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([mesh.cscale, inter.e0], dtype=np.float64).tobytes())
        for a in (basis.t, basis.nq, basis.l, basis.twoj, basis.twom, inter.ka, inter.kb, inter.kc, inter.kd,
                  np.asarray(pairs)[:, 0], np.asarray(pairs)[:, 1]):
            f.write(np.ascontiguousarray(a, dtype=np.int32).tobytes())
        f.write(np.asarray(inter.me, dtype=np.float64).tobytes())
        f.write(np.asfortranarray(inter.t1b, dtype=np.float64).tobytes(order="F"))
        for a in (mesh.alp, mesh.walp, mesh.bet, mesh.wbet, mesh.gam, mesh.wgam):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(np.transpose(U, (0, 2, 1)), dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(np.transpose(V, (0, 2, 1)), dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(coords, dtype=np.float64).tobytes())
        if q2 is not None:
            f.write(np.asarray(echarge, dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(np.transpose(np.asarray(q2, dtype=np.float64), (0, 2, 1))).tobytes())


def read_elem(path: str) -> dict:
    """post_general file (``Codes/PNAMP/PNAMP_IO.f90:39-94``) -> {J: {"coords": (b1, g1, b2, g2),
    "k": {(K1, K2): (N, H, nn, zz, J2, qpred0, qcpred0, q0p, q0pc)},
    "dJ1": {(K1, K2): (qpred1, qcpred1)}, "dJ2": {(K1, K2): (qpred2, qcpred2)}}}.

    Follows the record structure line by line: (10i5) J; (8e15.8) the
    deformations; per K2, K1 in [-k2max, k2max] ((10i5) K1 K2 and five (8e15.8)
    lines); then per K2 the K1 in [-k2max-1, k2max+1] and [-k2max-2, k2max+2]
    lines of the dJ = 1, 2 reduced Q2 elements; k2max = J, or |J - 1| for odd J.
    Raises ValueError on any line that breaks the layout."""
    lines = open(path).read().split("\n")
    while lines and not lines[-1].strip():
        lines.pop()
    out, i = {}, 0

    def nums(s, count):
        if len(s) != 15 * count:
            raise ValueError(f"{path}:{i + 1}: expected {count} fields of 15 characters")
        return [float(s[15 * k:15 * k + 15]) for k in range(count)]

    def cpx(s, count):
        v = nums(s, 2 * count)
        return [complex(v[2 * k], v[2 * k + 1]) for k in range(count)]

    def ints(s, count):
        if len(s) != 5 * count:
            raise ValueError(f"{path}:{i + 1}: expected {count} fields of 5 characters")
        return [int(s[5 * k:5 * k + 5]) for k in range(count)]

    while i < len(lines):
        J = ints(lines[i], 1)[0]
        coords = tuple(nums(lines[i + 1], 4))
        i += 2
        k2max = J if J % 2 == 0 else abs(J - 1)
        blk = {}
        for k2 in range(-k2max, k2max + 1):
            for k1 in range(-k2max, k2max + 1):
                if ints(lines[i], 2) != [k1, k2]:
                    raise ValueError(f"{path}:{i + 1}: expected K1 K2 = {k1} {k2}")
                n_h, nn_zz = cpx(lines[i + 1], 2), cpx(lines[i + 2], 2)
                j2, q0, q0p = cpx(lines[i + 3], 1), cpx(lines[i + 4], 2), cpx(lines[i + 5], 2)
                blk[(k1, k2)] = (n_h[0], n_h[1], nn_zz[0], nn_zz[1], j2[0], q0[0], q0[1], q0p[0], q0p[1])
                i += 6
        d1, d2 = {}, {}
        for k2 in range(-k2max, k2max + 1):
            for k1 in range(-k2max - 1, k2max + 2):
                d1[(k1, k2)] = tuple(cpx(lines[i], 2))
                i += 1
            for k1 in range(-k2max - 2, k2max + 3):
                d2[(k1, k2)] = tuple(cpx(lines[i], 2))
                i += 1
        out[J] = {"coords": coords, "k": blk, "dJ1": d1, "dJ2": d2}
    return out
