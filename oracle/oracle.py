"""ctypes front end of the CPU oracle (``oracle/pnamp_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``bench.py``'s CPU-baseline
leg and ``__graft_entry__.smoke()`` as the checker.  The product path
(``sccm_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpnamp_oracle.so")
NOBS = 6

_i32p = C.POINTER(C.c_int)
_f64p = C.POINTER(C.c_double)


class OrCfg(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("t", _i32p), ("nq", _i32p), ("l", _i32p), ("twoj", _i32p), ("twom", _i32p),
        ("nme", C.c_int64),
        ("ka", _i32p), ("kb", _i32p), ("kc", _i32p), ("kd", _i32p),
        ("me", _f64p), ("t1b", _f64p), ("e0", C.c_double),
        ("nalp", C.c_int), ("nbet", C.c_int), ("ngam", C.c_int),
        ("alp", _f64p), ("walp", _f64p), ("bet", _f64p), ("wbet", _f64p),
        ("gam", _f64p), ("wgam", _f64p),
        ("nphi", C.c_int), ("jmax", C.c_int), ("fold", C.c_int), ("norm_method", C.c_int),
        ("cscale", C.c_double), ("iscale", C.c_int),
        ("zval", C.c_int), ("nval", C.c_int), ("nthreads", C.c_int), ("energy_method", C.c_int),
        ("q2", _f64p),
    ]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "pnamp_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-B" if force else "-s", "libpnamp_oracle.so"], cwd=HERE, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.or_wigner_d.restype = C.c_double
        L.or_wigner_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        L.or_kernel_pair.restype = C.c_int
        L.or_point.restype = C.c_int
        L.or_zbareover.restype = C.c_int
        L.or_hndiag.restype = C.c_int
        L.or_threej.restype = C.c_double
        L.or_threej.argtypes = [C.c_int] * 6
        L.or_norb.restype = C.c_int
        L.or_kernel_pair_e2.restype = C.c_int
        _lib = L
    return _lib


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    """Holds one (basis, interaction, mesh, Z, N) configuration; keeps the
    numpy buffers alive for the C side."""

    def __init__(self, basis, inter, mesh, zval: int, nval: int, nthreads: int = 0, energy_method: int = 0,
                 q2=None):
        """energy_method 0: the list sum of HFB_Energy_COMPLEX (Method II, the
        reference's path); 1: the field form (Method I).  q2: the one-body
        Q_{2mu} matrices [5, n, n] (mu = -2..2) for the E2 observables."""
        self._keep = []
        k = self._keep

        def i32(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            k.append(a)
            return _p(a, C.c_int)

        def f64(x):
            a = np.ascontiguousarray(x, dtype=np.float64)
            k.append(a)
            return _p(a)

        self.basis, self.inter, self.mesh = basis, inter, mesh
        t1b = np.asfortranarray(inter.t1b, dtype=np.float64)
        k.append(t1b)
        self.cfg = OrCfg(
            basis.n, i32(basis.t), i32(basis.nq), i32(basis.l), i32(basis.twoj), i32(basis.twom),
            inter.nme, i32(inter.ka), i32(inter.kb), i32(inter.kc), i32(inter.kd),
            f64(inter.me), _p(t1b), float(inter.e0),
            mesh.alp.size, mesh.bet.size, mesh.gam.size,
            f64(mesh.alp), f64(mesh.walp), f64(mesh.bet), f64(mesh.wbet), f64(mesh.gam), f64(mesh.wgam),
            mesh.nphi, mesh.jmax, mesh.fold, mesh.norm_method, float(mesh.cscale), int(mesh.iscale),
            int(zval), int(nval), int(nthreads), int(energy_method),
            None if q2 is None else _p(self._q2(q2)))

    def _q2(self, q2):
        a = np.ascontiguousarray(np.transpose(np.asarray(q2, dtype=np.float64), (0, 2, 1)))  # column-major
        self._keep.append(a)
        return a

    @property
    def n(self):
        return self.basis.n

    @property
    def norb(self):
        return lib().or_norb(C.byref(self.cfg))

    def kernel_pair_e2(self, U0, V0, U1, V1):
        """(zres, kern, ze2 [NE, 2, 5], zrho [NE, 2, norb, norb],
        qred [2, 5, jmax+1, 2jmax+5, 2jmax+1], rho1b [2, norb, norb])."""
        m = self.mesh
        NE, J2m, jm, nb = m.n_euler, 2 * m.jmax + 1, m.jmax, self.norb
        zres = np.zeros((NE, NOBS), dtype=np.complex128)
        kern = np.zeros((NOBS, jm + 1, J2m, J2m), dtype=np.complex128)
        ze2 = np.zeros((NE, 2, 5), dtype=np.complex128)
        zrho = np.zeros((NE, 2, nb, nb), dtype=np.complex128)
        qred = np.zeros((2, 5, jm + 1, J2m + 4, J2m), dtype=np.complex128)
        rho1b = np.zeros((2, nb, nb), dtype=np.complex128)
        args = [np.asfortranarray(x, dtype=np.float64) for x in (U0, V0, U1, V1)]
        info = lib().or_kernel_pair_e2(C.byref(self.cfg), *[_p(a) for a in args], _p(zres), _p(kern), _p(ze2),
                                       _p(zrho), _p(qred), _p(rho1b))
        if info:
            raise RuntimeError(f"oracle kernel_pair_e2 failed: info={info}")
        return zres, kern, ze2, zrho, qred, rho1b

    def kernel_pair(self, U0, V0, U1, V1):
        """Returns (zres [NE, 5], kern [5, jmax+1, 2jmax+1, 2jmax+1]) complex."""
        m = self.mesh
        NE = m.n_euler
        J2m = 2 * m.jmax + 1
        zres = np.zeros((NE, NOBS), dtype=np.complex128)
        kern = np.zeros((NOBS, m.jmax + 1, J2m, J2m), dtype=np.complex128)
        args = [np.asfortranarray(x, dtype=np.float64) for x in (U0, V0, U1, V1)]
        info = lib().or_kernel_pair(C.byref(self.cfg), *[_p(a) for a in args],
                                    _p(zres), _p(kern))
        if info:
            raise RuntimeError(f"oracle kernel_pair failed: info={info}")
        return zres, kern

    def point(self, U0, V0, U1, V1, alp, bet, gam, phip, phin):
        n = self.n
        norm = np.zeros(4, dtype=np.complex128)
        rho = np.zeros((n, n), dtype=np.complex128, order="F")
        k10 = np.zeros_like(rho)
        k01 = np.zeros_like(rho)
        e = np.zeros(4, dtype=np.complex128)
        args = [np.asfortranarray(x, dtype=np.float64) for x in (U0, V0, U1, V1)]
        L = lib()
        info = L.or_point(C.byref(self.cfg), C.c_double(alp), C.c_double(bet), C.c_double(gam),
                          C.c_double(phip), C.c_double(phin), *[_p(a) for a in args],
                          _p(norm), _p(rho), _p(k10), _p(k01),
                          _p(e[0:]), _p(e[1:]), _p(e[2:]), _p(e[3:]))
        if info:
            raise RuntimeError(f"oracle point failed: info={info}")
        return dict(norm=norm[0], onishi=norm[1], det0=norm[2].real, det1=norm[3].real,
                    rho=rho, k10=k10, k01=k01, ekin=e[0], ehfb=e[1], ekin_field=e[2], ehfb_field=e[3])

    def rotation(self, alp, bet, gam):
        n = self.n
        R = np.zeros((n, n), dtype=np.complex128, order="F")
        lib().or_rotation_mat(C.byref(self.cfg), C.c_double(alp), C.c_double(bet), C.c_double(gam), _p(R))
        return R


def wigner_d(twoj: int, twom1: int, twom2: int, beta: float) -> float:
    return lib().or_wigner_d(twoj, twom1, twom2, beta)


def threej(tj1, tj2, tj3, tm1, tm2, tm3) -> float:
    """Wigner 3j symbol, arguments doubled."""
    return lib().or_threej(tj1, tj2, tj3, tm1, tm2, tm3)


def pfaffian(S: np.ndarray) -> complex:
    """ZPfaffianF of a skew matrix (upper triangle used)."""
    A = np.asfortranarray(S, dtype=np.complex128).copy(order="F")
    lib().or_pfaffian(C.c_int(A.shape[0]), _p(A))
    return complex(A[0, 0])


def zbareover(U, V):
    n = U.shape[0]
    Zc = np.zeros((n, n), dtype=np.complex128, order="F")
    det = C.c_double()
    info = lib().or_zbareover(C.c_int(n), _p(np.asfortranarray(U, dtype=np.float64)),
                              _p(np.asfortranarray(V, dtype=np.float64)), _p(Zc), C.byref(det))
    if info:
        raise RuntimeError(f"zbareover info={info}")
    return Zc, det.value


def hndiag(NN, HH, eps: float, nos: int = 0):
    dim = NN.shape[0]
    EN = np.zeros(dim)
    E = np.zeros(dim)
    F = np.zeros((dim, dim), order="F")
    m = C.c_int()
    ifl = lib().or_hndiag(C.c_int(dim), _p(np.asfortranarray(NN, dtype=np.float64)),
                          _p(np.asfortranarray(HH, dtype=np.float64)), C.c_double(eps),
                          C.c_int(nos), _p(EN), C.byref(m), _p(E), _p(F))
    M = m.value
    return dict(ifl=ifl, m=M, EN=EN, E=E[:M], F=np.array(F[:, :M]))


def presolve_hwg(NN, HH, eps: float, kmax: int = 10, literal: bool = True):
    """PreSolve_HWG (Codes/GCM/GCM_HWG.f90:2-100) with mean_variance (:205-254).

    For NOS = 1..M (M = norm eigenvalues kept by HNDIAG's cutoff) the levels
    E_k of HNDIAG with NOS natural states and the variance
    q2 = sum_i [sum_j (H_ij - E_k N_ij) f_j]^2 of their amplitudes f = FF(:, k),
    k < min(kmax, NOS); per level the NOS of least variance, starting from
    the reference's 100 and updating on a strictly smaller |q2|.  NOS > M
    repeat NOS = M and never win.  ``literal``: one or_hndiag call per NOS, as
    the reference loops; otherwise the same quantities from one numpy
    eigendecomposition of N and the nested leading blocks of R^T H R (for
    large dimensions)."""
    dim = NN.shape[0]
    r0 = hndiag(NN, HH, eps)
    M = r0["m"]
    E = np.zeros((max(M, 0), kmax))
    var = np.zeros((max(M, 0), kmax))
    best = np.full(kmax, 100.0)
    nos_best = np.zeros(kmax, dtype=int)
    e_best = np.zeros(kmax)
    if M <= 0:
        return dict(m=0, EN=r0["EN"], E=E, var=var, nos=nos_best, e_best=e_best)
    if not literal:
        w, v = np.linalg.eigh(-NN)
        order = np.argsort(w, kind="stable")
        R = v[:, order[:M]] / np.sqrt(-w[order[:M]])
        G = R.T @ HH @ R
    for nos in range(1, M + 1):
        if literal:
            r = hndiag(NN, HH, eps, nos)
            ek, F = r["E"], r["F"]
        else:
            ek, g = np.linalg.eigh(G[:nos, :nos])
            F = R[:, :nos] @ g
        for k in range(min(kmax, nos)):
            z = (HH - ek[k] * NN) @ F[:, k]
            q2 = float(z @ z)
            E[nos - 1, k] = ek[k]
            var[nos - 1, k] = q2
            if abs(best[k]) > abs(q2):
                best[k] = q2
                nos_best[k] = nos
                e_best[k] = ek[k]
    return dict(m=M, EN=r0["EN"], E=E, var=var, nos=nos_best, e_best=e_best)
