"""Learned surrogate of the projected-kernel mesh on the HIP network kernels
(C-ABI include/pnamp_mlp.h, sccm_amd/csrc/mlp_gpu.hip).

The network maps a mesh point (pair q, q', Euler alpha, beta, gamma, gauge
phi_p, phi_n) to the norm and energy kernels there: outputs log|N|,
cos arg N, sin arg N, Re E, Im E (E = H/N + E0 as summed into ZRES slot 7 of
Kernel4Triaxial, PNAMP_Kernel3D.f90:296-324).  arg N is learned through its
cosine and sine so that the target has no 2 pi jump.

Inputs (14): each state of the pair by its deformation (beta2 cos gamma,
beta2 sin gamma), cos and sin of alpha, beta, gamma, and of phi_p, phi_n.
alpha and gamma are inputs: the synthetic states are triaxial, so the
kernels depend on all three Euler angles.

Training data come from the exact factored path (pnamp_mesh_samples): a seeded
random subset of all mesh points of the GCM, disjoint seeded train and test
sets.  The surrogate kernels replace the exact value at points where the
predicted |N| >= a threshold and keep the exact FP64 value elsewhere; the
kernels are then projected on the GPU (pnamp_project_zres) and solved
(Hill-Wheeler) as the exact ones, and the levels compared at PreSolve_HWG's
natural-state numbers of the exact kernels.

``screen_bound`` measures what any screen or surrogate could gain at all on
a GCM, from the exact ZRES alone: the |dE| of the levels when the Euler
points whose number-projected norm |ZRES_0| lies below tau times the pair's
largest are dropped (an ideal screen, which knows the exact values), and when
every Euler point's observables carry a relative error eps (a surrogate of
that accuracy).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .pnamp import NOBS, load_library

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
NIN, NOUT = 14, 5
_bound = False


def load_mlp():
    """The HIP library with the pnamp_mlp_* entry points bound."""
    global _bound
    L = load_library()
    if not _bound:
        vp = C.c_void_p
        L.pnamp_mlp_create.argtypes = [C.c_int, _i32p, C.c_uint32, C.c_int, C.POINTER(vp)]
        L.pnamp_mlp_create.restype = C.c_int
        L.pnamp_mlp_n_params.argtypes = [vp]
        L.pnamp_mlp_n_params.restype = C.c_int64
        L.pnamp_mlp_get_params.argtypes = [vp, _f32p]
        L.pnamp_mlp_get_params.restype = C.c_int
        L.pnamp_mlp_set_params.argtypes = [vp, _f32p]
        L.pnamp_mlp_set_params.restype = C.c_int
        L.pnamp_mlp_set_data.argtypes = [vp, C.c_int, _f32p, _f32p, _f32p]
        L.pnamp_mlp_set_data.restype = C.c_int
        L.pnamp_mlp_loss_grad.argtypes = [vp, C.c_int, _i32p, _f32p, _f32p]
        L.pnamp_mlp_loss_grad.restype = C.c_int
        L.pnamp_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint32, _f32p]
        L.pnamp_mlp_train.restype = C.c_int
        L.pnamp_mlp_predict.argtypes = [vp, C.c_int, _f32p, _f32p, C.c_int]
        L.pnamp_mlp_predict.restype = C.c_int
        L.pnamp_mlp_last_ms.argtypes = [vp]
        L.pnamp_mlp_last_ms.restype = C.c_float
        L.pnamp_mlp_last_error.argtypes = [vp]
        L.pnamp_mlp_last_error.restype = C.c_char_p
        L.pnamp_mlp_destroy.argtypes = [vp]
        L.pnamp_mlp_destroy.restype = None
        _bound = True
    return L


def _f(a):
    return a.ctypes.data_as(_f32p)


class MLP:
    """Fully connected network on the GPU: sizes [n_in, width, ..., n_out]."""

    def __init__(self, sizes, seed: int = 0, device: int = 0):
        self._L = load_mlp()
        self.sizes = list(sizes)
        sz = np.ascontiguousarray(self.sizes, dtype=np.int32)
        h = C.c_void_p()
        rc = self._L.pnamp_mlp_create(len(sizes) - 1, sz.ctypes.data_as(_i32p), seed, device, C.byref(h))
        self._h = h
        if rc != 0:
            msg = self._L.pnamp_mlp_last_error(h).decode() if h.value else "create failed"
            self.close()
            raise RuntimeError(f"pnamp_mlp_create: {rc} {msg}")

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"pnamp_mlp error {rc}: {self._L.pnamp_mlp_last_error(self._h).decode()}")

    @property
    def n_params(self) -> int:
        return int(self._L.pnamp_mlp_n_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.pnamp_mlp_get_params(self._h, _f(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self._L.pnamp_mlp_set_params(self._h, _f(p)))

    def set_data(self, X: np.ndarray, Y: np.ndarray, wout=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        w = None if wout is None else np.ascontiguousarray(wout, dtype=np.float32)
        self._check(self._L.pnamp_mlp_set_data(self._h, X.shape[0], _f(X), _f(Y), _f(w) if w is not None else None))

    def loss_grad(self, idx) -> tuple[float, np.ndarray]:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_float()
        g = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.pnamp_mlp_loss_grad(self._h, len(idx), idx.ctypes.data_as(_i32p), C.byref(loss), _f(g)))
        return float(loss.value), g

    def train(self, epochs: int, batch: int, lr: float, decay: float = 1.0, seed: int = 0) -> np.ndarray:
        el = np.zeros(epochs, dtype=np.float32)
        self._check(self._L.pnamp_mlp_train(self._h, epochs, batch, lr, decay, seed, _f(el)))
        return el

    def predict(self, X: np.ndarray, bf16: bool = False) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.sizes[-1]), dtype=np.float32)
        self._check(self._L.pnamp_mlp_predict(self._h, X.shape[0], _f(X), _f(Y), 1 if bf16 else 0))
        return Y

    @property
    def last_ms(self) -> float:
        return float(self._L.pnamp_mlp_last_ms(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.pnamp_mlp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- features and targets
def mesh_features(w, pairs) -> np.ndarray:
    """Inputs of every mesh point of the pairs: [npairs, NE, nphi, nphi, 14]
    float32, Euler index iang = (ibeta * nalp + ialpha) * ngam + igamma as in
    the kernels."""
    m = w.mesh
    na, nb, ng, f = m.alp.size, m.bet.size, m.gam.size, m.nphi
    c = np.asarray(w.coords, dtype=np.float64)
    dq = np.stack([c[:, 0] * np.cos(np.radians(c[:, 1])), c[:, 0] * np.sin(np.radians(c[:, 1]))], 1)
    pairs = np.asarray(pairs).reshape(-1, 2)
    ib, ia, ig = np.meshgrid(np.arange(nb), np.arange(na), np.arange(ng), indexing="ij")
    ang = np.stack([np.cos(m.alp[ia]), np.sin(m.alp[ia]), np.cos(m.bet[ib]), np.sin(m.bet[ib]),
                    np.cos(m.gam[ig]), np.sin(m.gam[ig])], -1).reshape(-1, 6)
    phi = 2.0 * np.pi * np.arange(f) / f if f > 1 else np.zeros(1)
    pp, pn = np.meshgrid(phi, phi, indexing="ij")
    gau = np.stack([np.cos(pp), np.sin(pp), np.cos(pn), np.sin(pn)], -1)  # [f, f, 4]
    NE = ang.shape[0]
    X = np.zeros((len(pairs), NE, f, f, NIN), dtype=np.float32)
    X[..., 0:2] = dq[pairs[:, 0]][:, None, None, None, :]
    X[..., 2:4] = dq[pairs[:, 1]][:, None, None, None, :]
    X[..., 4:10] = ang[None, :, None, None, :]
    X[..., 10:14] = gau[None, None, :, :, :]
    return X


def mesh_targets(samples: np.ndarray) -> np.ndarray:
    """Targets from pnamp_mesh_samples output [..., 2] complex: log|N|,
    cos arg N, sin arg N, Re E, Im E (float64)."""
    N, E = samples[..., 0], samples[..., 1]
    a = np.abs(N)
    ph = np.angle(N)
    return np.stack([np.log(np.maximum(a, 1e-300)), np.cos(ph), np.sin(ph), E.real, E.imag], -1)


def targets_to_values(T: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Network outputs (de-standardised) -> (N, E) complex."""
    ph = np.arctan2(T[..., 2], T[..., 1])
    N = np.exp(T[..., 0]) * np.exp(1j * ph)
    return N, T[..., 3] + 1j * T[..., 4]


class Standardizer:
    def __init__(self, A: np.ndarray):
        self.mu = A.mean(0)
        self.sd = A.std(0)
        self.sd[self.sd == 0] = 1.0

    def fwd(self, A):
        return ((A - self.mu) / self.sd).astype(np.float32)

    def inv(self, A):
        return A.astype(np.float64) * self.sd + self.mu


# ---------------------------------------------------------------- the experiment
def assemble_zres(samples_exact: np.ndarray, zres_exact: np.ndarray, Npred: np.ndarray, Epred: np.ndarray,
                  threshold: float, policy: str = "replace"):
    """ZRES with the surrogate's norm and energy at some mesh points and the
    exact values at the others: slot 0 = mean of N over the gauge points,
    slot 4 = mean of N E; the other slots stay exact.
    policy "replace": surrogate where the predicted |N| >= threshold (exact
    FP64 fallback below it).  policy "screen": surrogate at every gauge point
    of the Euler points whose largest predicted |N| is below threshold, exact
    elsewhere (the Euler points the exact path could skip).
    Returns (zres, fraction of mesh points served by the surrogate)."""
    if policy == "replace":
        use = np.abs(Npred) >= threshold
    else:
        small = np.abs(Npred).reshape(Npred.shape[0], Npred.shape[1], -1).max(-1) < threshold
        use = np.broadcast_to(small[:, :, None, None], Npred.shape)
    N = np.where(use, Npred, samples_exact[..., 0])
    E = np.where(use, Epred, samples_exact[..., 1])
    z = zres_exact.copy()
    f2 = N.shape[-1] * N.shape[-2]
    z[..., 0] = N.reshape(N.shape[0], N.shape[1], -1).sum(-1) / f2
    z[..., 4] = (N * E).reshape(N.shape[0], N.shape[1], -1).sum(-1) / f2
    return z, float(use.mean())


def _levels_dE(ctx, w, kern, lv) -> float:
    """Largest |dE| (MeV) of ``kern``'s levels against ``lv`` (gcm_levels of
    the exact kernels) at the same natural-state numbers."""
    from .driver import levels_at_nos
    e = levels_at_nos(ctx, kern, w.pairs, len(w.U), w.mesh.jmax, lv, w.hw_eps)
    d = 0.0
    for J, l in lv.items():
        for x, y in zip(e[J], l["e"]):
            if x is not None and y is not None:
                d = max(d, abs(x - y))
    return d


def screen_bound(w, ctx, zres_exact, lv, taus=(1e-6, 1e-5, 1e-4, 1e-3), epss=(1e-4, 1e-3, 1e-2),
                 seed: int = 1, log=None) -> dict:
    """Ideal screen: |dE| (keV) when the Euler points with |ZRES_0| < tau x
    (the pair's largest |ZRES_0|) are set to zero, and the fraction dropped;
    surrogate accuracy: |dE| when each Euler point's six observables are
    scaled by (1 + eps g), g standard normal per point."""
    a = np.abs(zres_exact[..., 0])
    rel = a / np.maximum(a.max(1, keepdims=True), 1e-300)
    out = {"ideal_screen": [], "per_point_relative_error": []}
    for tau in taus:
        z = zres_exact.copy()
        m = rel < tau
        z[m] = 0.0
        out["ideal_screen"].append({"tau": tau, "fraction_dropped": float(m.mean()),
                                    "max_abs_dE_keV": 1e3 * _levels_dE(ctx, w, ctx.project_zres(z), lv)})
        if log:
            b = out["ideal_screen"][-1]
            log(f"surrogate: ideal screen tau {tau:g}: {100 * b['fraction_dropped']:.3f} % of Euler points "
                f"dropped, max |dE| {b['max_abs_dE_keV']:.4g} keV")
    rng = np.random.default_rng(seed)
    for eps in epss:
        z = zres_exact * (1.0 + eps * rng.normal(size=zres_exact.shape[:2]))[..., None]
        out["per_point_relative_error"].append({"eps": eps,
                                                "max_abs_dE_keV": 1e3 * _levels_dE(ctx, w, ctx.project_zres(z), lv)})
        if log:
            log(f"surrogate: per-point relative error {eps:g}: max |dE| "
                f"{out['per_point_relative_error'][-1]['max_abs_dE_keV']:.4g} keV")
    return out


def run(w, ctx, n_train: int = 200_000, n_test: int = 50_000, width: int = 128, hidden: int = 3,
        epochs: int = 60, batch: int = 4096, lr: float = 2e-3, decay: float = 0.96, threshold: float = 1e-3,
        seed: int = 3001, log=print, bound: bool = True) -> dict:
    """Train the surrogate on exact samples of the workload w (ctx: a PNAMP
    context with w's states and a factored form selected) and compare the
    Hill-Wheeler energies of surrogate+fallback kernels with the exact ones."""
    from .driver import gcm_levels

    t0 = time.perf_counter()
    zres_exact, kern_exact = ctx.kernels_batched(w.pairs, w.zval, w.nval, want_zres=True)
    t_exact_ms = ctx.timing()[0]
    samples = ctx.mesh_samples(w.pairs, w.zval, w.nval)
    X = mesh_features(w, w.pairs)
    T = mesh_targets(samples)
    nall = T[..., 0].size
    Xf, Tf = X.reshape(nall, NIN), T.reshape(nall, NOUT)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nall)
    tr, te = perm[:n_train], perm[n_train:n_train + n_test]
    sx, sy = Standardizer(Xf[tr]), Standardizer(Tf[tr])
    log(f"surrogate: {nall} mesh points, {len(tr)} train / {len(te)} test samples "
        f"({time.perf_counter() - t0:.1f} s to sample)")
    net = MLP([NIN] + [width] * hidden + [NOUT], seed=seed)
    net.set_data(sx.fwd(Xf[tr]), sy.fwd(Tf[tr]))
    t1 = time.perf_counter()
    losses = net.train(epochs, batch, lr, decay, seed)
    t_train = time.perf_counter() - t1
    log(f"surrogate: trained {epochs} epochs in {t_train:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    res = {"n_mesh_points": int(nall), "n_train": int(len(tr)), "n_test": int(len(te)),
           "layers": [NIN] + [width] * hidden + [NOUT], "epochs": epochs, "batch": batch,
           "train_seconds": t_train, "loss_first": float(losses[0]), "loss_last": float(losses[-1])}
    # test-set errors (FP32 and BF16 inference)
    for tag, bf in (("fp32", False), ("bf16", True)):
        P = sy.inv(net.predict(sx.fwd(Xf[te]), bf16=bf))
        Np, Ep = targets_to_values(P)
        Nx, Ex = samples.reshape(nall, 2)[te, 0], samples.reshape(nall, 2)[te, 1]
        big = np.abs(Nx) >= threshold
        res[f"test_{tag}"] = {
            "rmse_log_abs_N": float(np.sqrt(np.mean((P[:, 0] - Tf[te, 0]) ** 2))),
            "rmse_log_abs_N_mean_predictor": float(np.std(Tf[te, 0])),
            "median_rel_err_N_where_abs_N_ge_threshold": float(np.median(np.abs(Np[big] - Nx[big]) / np.abs(Nx[big]))),
            "median_abs_err_E_MeV_where_abs_N_ge_threshold": float(np.median(np.abs(Ep[big] - Ex[big]))),
        }
    # whole-mesh inference, surrogate + exact fallback, projection, Hill-Wheeler
    P = sy.inv(net.predict(sx.fwd(Xf), bf16=True)).reshape(T.shape)
    t_inf_ms = net.last_ms
    Np, Ep = targets_to_values(P)
    lv = gcm_levels(ctx, kern_exact, w.pairs, len(w.U), w.mesh.jmax, w.hw_eps, kmax=3)
    res.update({"threshold_abs_N": threshold, "inference_ms_all_points_bf16": t_inf_ms,
                "exact_factored_ms_all_points": t_exact_ms,
                "levels": "lowest 3 per J at PreSolve_HWG's natural-state numbers of the exact kernels"})
    for policy, thr in (("replace", threshold), ("screen", threshold), ("screen_1e-2", 1e-2)):
        zs, frac = assemble_zres(samples, zres_exact, Np, Ep, thr, policy.split("_")[0])
        d = _levels_dE(ctx, w, ctx.project_zres(zs), lv)
        res[policy] = {"threshold_abs_N": thr, "fraction_of_mesh_points_from_surrogate": frac,
                       "max_abs_dE_keV": 1e3 * d}
        log(f"surrogate: policy {policy} threshold {thr}: {100 * frac:.1f} % of points from the network, "
            f"max |dE| {1e3 * d:.4g} keV")
    if bound:
        res["bound"] = screen_bound(w, ctx, zres_exact, lv, log=log)
    net.close()
    return res


def main():
    """The learned-surrogate experiment on the pf 20-state GCM: prints one
    JSON object (errors on held-out points, fraction of mesh points served,
    inference and exact times, |dE| per Hill-Wheeler level)."""
    import argparse
    import json
    import sys

    from . import configs
    from .pnamp import FORM_FACTORED, PNAMP

    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200_000)
    ap.add_argument("--test", type=int, default=50_000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--threshold", type=float, default=1e-3)
    ap.add_argument("--no-bound", action="store_true", help="skip the bound on any screen")
    args = ap.parse_args()
    w = configs.pf_gcm()
    ctx = PNAMP(w.basis, w.inter, w.mesh)
    ctx.set_states(w.U, w.V)
    ctx.set_form(FORM_FACTORED)
    r = run(w, ctx, n_train=args.train, n_test=args.test, epochs=args.epochs, threshold=args.threshold,
            log=lambda *a: print(*a, file=sys.stderr, flush=True), bound=not args.no_bound)
    ctx.close()
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
