"""Host front end of the HIP library ``libpnamp_gpu.so`` (C-ABI in
``include/pnamp_gpu.h``).

Mirrors the reference's per-pair kernel step: ``Kernel4Triaxial``
(``Codes/PNAMP/PNAMP_Kernel3D.f90:1``) for one pair, the pair loop of
``PNAMP_Main`` (``Codes/PNAMP/PNAMP_Main.f90:79-121``) as one batched call, and
``HNDIAG`` (``Codes/GCM/GCM_HNDiag.f90:2``).  There is no CPU path: if the
library is missing or fails to load, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpnamp_gpu.so")
NOBS = 6
OBS = ("norm", "Z", "N", "Ekin", "E", "J2")
FORM_PLAIN, FORM_FACTORED, FORM_MIXED_FP32, FORM_MIXED_BF16 = 0, 1, 2, 3

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class _Basis(C.Structure):
    _fields_ = [("n", C.c_int), ("t", _i32p), ("nq", _i32p), ("l", _i32p), ("twoj", _i32p), ("twom", _i32p)]


class _Inter(C.Structure):
    _fields_ = [("nme", C.c_int64), ("ka", _i32p), ("kb", _i32p), ("kc", _i32p), ("kd", _i32p),
                ("me2b", _f64p), ("t1b", _f64p), ("e0", C.c_double)]


class _Mesh(C.Structure):
    _fields_ = [("nalp", C.c_int), ("nbet", C.c_int), ("ngam", C.c_int),
                ("alp", _f64p), ("walp", _f64p), ("bet", _f64p), ("wbet", _f64p), ("gam", _f64p), ("wgam", _f64p),
                ("nphi", C.c_int), ("jmax", C.c_int), ("fold", C.c_int), ("cscale", C.c_double), ("iscale", C.c_int)]


class PnampError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pnamp error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.pnamp_create.argtypes = [C.POINTER(_Basis), C.POINTER(_Inter), C.POINTER(_Mesh), C.c_int, C.POINTER(vp)]
    L.pnamp_create.restype = C.c_int
    L.pnamp_set_states.argtypes = [vp, C.c_int, _f64p, _f64p]
    L.pnamp_set_states.restype = C.c_int
    L.pnamp_set_form.argtypes = [vp, C.c_int]
    L.pnamp_set_form.restype = C.c_int
    L.pnamp_states_separable.argtypes = [vp]
    L.pnamp_states_separable.restype = C.c_int
    L.pnamp_state_overlaps.argtypes = [vp, _f64p]
    L.pnamp_state_overlaps.restype = C.c_int
    L.pnamp_kernel_pair.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _f64p]
    L.pnamp_kernel_pair.restype = C.c_int
    L.pnamp_kernels_batched.argtypes = [vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _f64p, _f64p]
    L.pnamp_kernels_batched.restype = C.c_int
    L.pnamp_kernels_batched_zn.argtypes = [vp, C.c_int, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p]
    L.pnamp_kernels_batched_zn.restype = C.c_int
    L.pnamp_kernels_batched_noalloc.argtypes = [vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _f64p, _f64p]
    L.pnamp_kernels_batched_noalloc.restype = C.c_int
    L.pnamp_set_q2.argtypes = [vp, _f64p]
    L.pnamp_set_q2.restype = C.c_int
    L.pnamp_norb.argtypes = [vp]
    L.pnamp_norb.restype = C.c_int
    L.pnamp_kernels_e2.argtypes = [vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.pnamp_kernels_e2.restype = C.c_int
    L.pnamp_mesh_samples.argtypes = [vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _f64p]
    L.pnamp_mesh_samples.restype = C.c_int
    L.pnamp_project_zres.argtypes = [vp, C.c_int, _f64p, _f64p]
    L.pnamp_project_zres.restype = C.c_int
    L.pnamp_hw_solve.argtypes = [vp, C.c_int, _f64p, _f64p, C.c_double, C.c_int, _f64p, _f64p, _f64p,
                                 C.POINTER(C.c_int)]
    L.pnamp_hw_solve.restype = C.c_int
    L.pnamp_hw_presolve.argtypes = [vp, C.c_int, _f64p, _f64p, C.c_double, C.c_int, _f64p, C.POINTER(C.c_int),
                                    _f64p, _f64p, _i32p, _f64p]
    L.pnamp_hw_presolve.restype = C.c_int
    L.pnamp_last_timing.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.pnamp_last_timing.restype = C.c_int
    L.pnamp_graph_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pnamp_graph_info.restype = C.c_int
    L.pnamp_autotune.argtypes = [vp, C.c_int, _i32p, _i32p, C.c_int, C.c_int, _i32p, _i32p, C.POINTER(C.c_float),
                                 C.POINTER(C.c_int)]
    L.pnamp_autotune.restype = C.c_int
    L.pnamp_set_tune.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.pnamp_set_tune.restype = C.c_int
    L.pnamp_set_schur_pfaffian.argtypes = [vp, C.c_int]
    L.pnamp_set_schur_pfaffian.restype = C.c_int
    L.pnamp_schur_pfaffian_active.argtypes = [vp]
    L.pnamp_schur_pfaffian_active.restype = C.c_int
    L.pnamp_n_euler.argtypes = [vp]
    L.pnamp_n_euler.restype = C.c_int
    L.pnamp_last_error.argtypes = [vp]
    L.pnamp_last_error.restype = C.c_char_p
    L.pnamp_destroy.argtypes = [vp]
    L.pnamp_destroy.restype = None
    _lib = L
    return L


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class PNAMP:
    """One device context: basis, interaction, mesh and HFB states live in HBM."""

    def __init__(self, basis, inter, mesh, device: int = 0):
        # the norm is the Pfaffian overlap, the branch Kernel4Triaxial takes for
        # shell-model interactions (Codes/PNAMP/PNAMP_Kernel3D.f90:140-143,
        # 291-294); its Onishi branch (:286-289, HF states of other
        # interactions) has no HIP form
        if getattr(mesh, "norm_method", 0) != 0:
            raise PnampError(-1, "norm_method 1 (Onishi) is not a HIP form: the norm is the Pfaffian overlap")
        L = load_library()
        self._L = L
        self._keep = []

        def i32(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            self._keep.append(a)
            return _p(a, C.c_int32)

        def f64(x, order="C"):
            a = np.asarray(x, dtype=np.float64, order=order).copy(order=order)
            self._keep.append(a)
            return _p(a)

        self.basis, self.inter, self.mesh = basis, inter, mesh
        b = _Basis(basis.n, i32(basis.t), i32(basis.nq), i32(basis.l), i32(basis.twoj), i32(basis.twom))
        it = _Inter(inter.nme, i32(inter.ka), i32(inter.kb), i32(inter.kc), i32(inter.kd), f64(inter.me),
                    f64(inter.t1b, "F"), float(inter.e0))
        m = _Mesh(mesh.alp.size, mesh.bet.size, mesh.gam.size, f64(mesh.alp), f64(mesh.walp), f64(mesh.bet),
                  f64(mesh.wbet), f64(mesh.gam), f64(mesh.wgam), mesh.nphi, mesh.jmax, mesh.fold,
                  float(mesh.cscale), int(mesh.iscale))
        h = C.c_void_p()
        rc = L.pnamp_create(C.byref(b), C.byref(it), C.byref(m), device, C.byref(h))
        self._h = h
        if rc != 0:
            msg = L.pnamp_last_error(h).decode() if h.value else "create failed"
            self.close()
            raise PnampError(rc, msg)
        self.nstates = 0

    def _check(self, rc):
        if rc != 0:
            raise PnampError(rc, self._L.pnamp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pnamp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_euler(self) -> int:
        return self._L.pnamp_n_euler(self._h)

    def set_states(self, U: np.ndarray, V: np.ndarray):
        """U, V: [nstates, n, n] with U[q][i, k] (single-particle i, quasiparticle k)."""
        U = np.asarray(U, dtype=np.float64)
        V = np.asarray(V, dtype=np.float64)
        if U.ndim == 2:
            U, V = U[None], V[None]
        n = self.basis.n
        if U.shape[1:] != (n, n) or V.shape != U.shape:
            raise ValueError("U, V must be [nstates, n, n]")
        Uf = np.ascontiguousarray(np.transpose(U, (0, 2, 1)))  # per-state column-major
        Vf = np.ascontiguousarray(np.transpose(V, (0, 2, 1)))
        self._check(self._L.pnamp_set_states(self._h, U.shape[0], _p(Uf), _p(Vf)))
        self.nstates = U.shape[0]

    def set_form(self, form: int):
        """FORM_PLAIN (general), FORM_FACTORED (states without p-n mixing, FP64) or
        FORM_MIXED_FP32 / FORM_MIXED_BF16 (factored, V contraction on MFMA)."""
        self._check(self._L.pnamp_set_form(self._h, int(form)))

    @property
    def separable(self) -> bool:
        return self._L.pnamp_states_separable(self._h) == 1

    def state_overlaps(self) -> np.ndarray:
        out = np.zeros(self.nstates)
        self._check(self._L.pnamp_state_overlaps(self._h, _p(out)))
        return out

    def _kshape(self):
        j = self.mesh.jmax
        return (NOBS, j + 1, 2 * j + 1, 2 * j + 1)

    def kernel_pair(self, q0: int, q1: int, zval: int, nval: int, want_zres: bool = False):
        kern = np.zeros(self._kshape(), dtype=np.complex128)
        zres = np.zeros((self.n_euler, NOBS), dtype=np.complex128) if want_zres else None
        self._check(self._L.pnamp_kernel_pair(self._h, int(q0), int(q1), int(zval), int(nval),
                                              _p(zres) if zres is not None else None, _p(kern)))
        return (zres, kern) if want_zres else kern

    def kernels_batched(self, pairs, zval: int, nval: int, want_zres: bool = False, want_kern: bool = True,
                        noalloc: bool = False):
        """Kernels of many pairs.  ``noalloc`` selects the entry point that
        allocates no device memory (plain form, static scratch array)."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        q0 = np.ascontiguousarray(pairs[:, 0])
        q1 = np.ascontiguousarray(pairs[:, 1])
        npairs = len(pairs)
        kern = np.zeros((npairs,) + self._kshape(), dtype=np.complex128) if want_kern else None
        zres = np.zeros((npairs, self.n_euler, NOBS), dtype=np.complex128) if want_zres else None
        fn = self._L.pnamp_kernels_batched_noalloc if noalloc else self._L.pnamp_kernels_batched
        self._check(fn(self._h, npairs, _p(q0, C.c_int32), _p(q1, C.c_int32), int(zval), int(nval),
                       _p(zres) if zres is not None else None, _p(kern) if kern is not None else None))
        return (zres, kern) if want_zres else kern

    def kernels_batched_zn(self, pairs, zval, nval, want_zres: bool = False):
        """Kernels of many pairs with valence numbers per pair (zval[i], nval[i])."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        q0 = np.ascontiguousarray(pairs[:, 0])
        q1 = np.ascontiguousarray(pairs[:, 1])
        z = np.ascontiguousarray(zval, dtype=np.int32)
        nv = np.ascontiguousarray(nval, dtype=np.int32)
        npairs = len(pairs)
        if z.shape != (npairs,) or nv.shape != (npairs,):
            raise ValueError("zval, nval need one entry per pair")
        kern = np.zeros((npairs,) + self._kshape(), dtype=np.complex128)
        zres = np.zeros((npairs, self.n_euler, NOBS), dtype=np.complex128) if want_zres else None
        self._check(self._L.pnamp_kernels_batched_zn(self._h, npairs, _p(q0, C.c_int32), _p(q1, C.c_int32),
                                                     _p(z, C.c_int32), _p(nv, C.c_int32),
                                                     _p(zres) if zres is not None else None, _p(kern)))
        return (zres, kern) if want_zres else kern

    def set_q2(self, q2: np.ndarray):
        """Q_{2mu} one-body matrices [5, n, n] (mu = -2..2) for the E2 kernels."""
        a = np.ascontiguousarray(np.transpose(np.asarray(q2, dtype=np.float64), (0, 2, 1)))  # column-major
        self._check(self._L.pnamp_set_q2(self._h, _p(a)))

    @property
    def norb(self) -> int:
        return self._L.pnamp_norb(self._h)

    def kernels_e2(self, pairs, zval: int, nval: int):
        """(zres, kern, ze2 [P, NE, 2, 5], qred [P, 2, 5, jmax+1, 2jmax+5, 2jmax+1],
        rho1b [P, 2, norb, norb]) of many pairs (see pnamp_kernels_e2)."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        q0 = np.ascontiguousarray(pairs[:, 0])
        q1 = np.ascontiguousarray(pairs[:, 1])
        P, jm, nb = len(pairs), self.mesh.jmax, self.norb
        kern = np.zeros((P,) + self._kshape(), dtype=np.complex128)
        zres = np.zeros((P, self.n_euler, NOBS), dtype=np.complex128)
        ze2 = np.zeros((P, self.n_euler, 2, 5), dtype=np.complex128)
        qred = np.zeros((P, 2, 5, jm + 1, 2 * jm + 5, 2 * jm + 1), dtype=np.complex128)
        rho1b = np.zeros((P, 2, nb, nb), dtype=np.complex128)
        self._check(self._L.pnamp_kernels_e2(self._h, P, _p(q0, C.c_int32), _p(q1, C.c_int32), int(zval), int(nval),
                                             _p(zres), _p(kern), _p(ze2), _p(qred), _p(rho1b)))
        return zres, kern, ze2, qred, rho1b

    def mesh_samples(self, pairs, zval: int, nval: int) -> np.ndarray:
        """Norm and energy of every gauge point: [npairs, NE, nphi, nphi, 2]
        complex (factored and mixed forms)."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        q0 = np.ascontiguousarray(pairs[:, 0])
        q1 = np.ascontiguousarray(pairs[:, 1])
        f = self.mesh.nphi
        out = np.zeros((len(pairs), self.n_euler, f, f, 2), dtype=np.complex128)
        self._check(self._L.pnamp_mesh_samples(self._h, len(pairs), _p(q0, C.c_int32), _p(q1, C.c_int32),
                                               int(zval), int(nval), _p(out)))
        return out

    def project_zres(self, zres: np.ndarray) -> np.ndarray:
        """J-projection of ZRES [npairs, NE, NOBS] complex into kernels."""
        z = np.ascontiguousarray(zres, dtype=np.complex128)
        kern = np.zeros((z.shape[0],) + self._kshape(), dtype=np.complex128)
        self._check(self._L.pnamp_project_zres(self._h, z.shape[0], _p(z), _p(kern)))
        return kern

    def timing(self) -> tuple[float, float]:
        a, b = C.c_float(), C.c_float()
        self._check(self._L.pnamp_last_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, pairs, zval: int, nval: int) -> dict:
        """Times the FP64 factored form's launch candidates on ``pairs`` and
        keeps the fastest: {"pf_waves", "rest_waves", "chunk_div",
        "candidates": [{"pf_waves", "rest_waves", "chunk_div", "ms"}]}."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        q0 = np.ascontiguousarray(pairs[:, 0])
        q1 = np.ascontiguousarray(pairs[:, 1])
        ch = np.zeros(3, dtype=np.int32)
        cand = np.zeros(3 * 32, dtype=np.int32)
        ms = np.zeros(32, dtype=np.float32)
        nc = C.c_int()
        self._check(self._L.pnamp_autotune(self._h, len(q0), q0.ctypes.data_as(_i32p), q1.ctypes.data_as(_i32p),
                                           zval, nval, ch.ctypes.data_as(_i32p), cand.ctypes.data_as(_i32p),
                                           ms.ctypes.data_as(C.POINTER(C.c_float)), C.byref(nc)))
        return {"pf_waves": int(ch[0]), "rest_waves": int(ch[1]), "chunk_div": int(ch[2]),
                "candidates": [{"pf_waves": int(cand[3 * i]), "rest_waves": int(cand[3 * i + 1]),
                                "chunk_div": int(cand[3 * i + 2]), "ms": float(ms[i])} for i in range(nc.value)]}

    def set_schur_pfaffian(self, enable: bool):
        """Allows the factored form's h x h Schur-complement Pfaffian (states
        with cond_1(Z_s) <= 1e6; rounding up to about cond(Z_s) eps)."""
        self._check(self._L.pnamp_set_schur_pfaffian(self._h, 1 if enable else 0))

    @property
    def schur_pfaffian_active(self) -> int:
        """2: every state takes the Schur-complement Pfaffian, 1: the pairs
        of some bra states do, 0: none (not allowed, or no state passes)."""
        r = self._L.pnamp_schur_pfaffian_active(self._h)
        self._check(min(r, 0))
        return r

    def set_tune(self, pf_waves: int, rest_waves: int, chunk_div: int):
        """Launch configuration of the factored form (a kept autotune choice)."""
        self._check(self._L.pnamp_set_tune(self._h, pf_waves, rest_waves, chunk_div))

    def graph_info(self) -> tuple[int, int, int]:
        """(graphs captured, replays launched, kernel nodes of the last graph)
        of the factored and mixed forms."""
        a, b, n = C.c_int(), C.c_int(), C.c_int()
        self._check(self._L.pnamp_graph_info(self._h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def hw_solve(self, N: np.ndarray, H: np.ndarray, eps: float, nos: int = 0, want_f: bool = True):
        dim = N.shape[0]
        Nf = np.asfortranarray(N, dtype=np.float64)
        Hf = np.asfortranarray(H, dtype=np.float64)
        EN = np.zeros(dim)
        E = np.zeros(dim)
        F = np.zeros((dim, dim), order="F") if want_f else None
        m = C.c_int()
        rc = self._L.pnamp_hw_solve(self._h, dim, _p(Nf), _p(Hf), float(eps), int(nos), _p(EN), _p(E),
                                    _p(F) if F is not None else None, C.byref(m))
        if rc not in (0, 1):
            self._check(rc)
        M = m.value
        return dict(ifl=rc, m=M, EN=EN, E=E[:M], F=None if F is None else np.array(F[:, :M]))

    def hw_presolve(self, N: np.ndarray, H: np.ndarray, eps: float, kmax: int = 10) -> dict:
        """PreSolve_HWG natural-state selection of one J block: levels and
        variances for every NOS, and per level the NOS of least variance."""
        dim = N.shape[0]
        Nf = np.asfortranarray(N, dtype=np.float64)
        Hf = np.asfortranarray(H, dtype=np.float64)
        EN = np.zeros(dim)
        E = np.zeros((dim, kmax))
        var = np.zeros((dim, kmax))
        nos = np.zeros(kmax, dtype=np.int32)
        eb = np.zeros(kmax)
        m = C.c_int()
        rc = self._L.pnamp_hw_presolve(self._h, dim, _p(Nf), _p(Hf), float(eps), int(kmax), _p(EN), C.byref(m),
                                       _p(E), _p(var), _p(nos, C.c_int32), _p(eb))
        if rc not in (0, 1):
            self._check(rc)
        M = m.value
        return dict(ifl=rc, m=M, EN=EN, E=E[:M], var=var[:M], nos=nos, e_best=eb)
