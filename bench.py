"""Benchmark of the PNAMP projected-kernel path on one MI355X.

Workload: the pf-space GCM (n = 40 m-states, Z,N valence 2,4, 20 HFB states,
210 pairs q <= q', Euler mesh alpha = gamma 10, beta 20, 9 x 9 gauge points,
J <= 8): 162,000 mesh points per pair, 34.02 M for the whole GCM.

A step is one pass of the chosen form over ``--pairs-per-step`` pairs
(default: all 210, i.e. the whole GCM) with inputs resident in HBM:
rotation, per-species Pfaffians and densities, the proton-neutron energy
contraction, gauge sums and the J-projection.  ``value`` is mesh points per
second (host wall clock around the timed steps; it includes the projection
and the copy of the projected kernels to the host).

Beside it, untimed by the steps: one full pass of the plain FP64 path (the
general drop-in) over all pairs, timed on its own, and the Hill-Wheeler
energies of both, whose largest difference over the lowest levels of every
J <= 8 is reported (the metric's |dE_GCM| condition).  ``roofline`` is the
dominant kernel of the chosen form (HIP events on the library stream;
algorithmic FP64 flops of the factored algorithm, see ``factored_flops``).
``cpu_baseline`` times the CPU oracle on a bounded sample of the workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix peaks are equal)


def plain_flops_per_point(n: int, nme: int) -> float:
    """Algorithmic real FP64 flops of one mesh point along the reference call
    sequence (complex multiply-add = 8 flops): 16 n^3 rotation (2 ZGEMM),
    16 n^3 Pfaf_Over inverse + ZGEMM, 19 n^3 Pfaffian of the 2n skew matrix,
    96 n^3 ROKAPPAphiomega (11 ZGEMM + inverse), 190 per interaction entry."""
    return 147.0 * n ** 3 + 190.0 * nme


def plain_terms(inter, n: int) -> int:
    """Terms of the plain kernels' regrouped energy list: entries with a != b
    and c != d, doubled when a + b != c + d."""
    import numpy as np
    idx = np.stack([inter.ka, inter.kb, inter.kc, inter.kd], 1) - 1
    keep = (idx[:, 0] != idx[:, 1]) & (idx[:, 2] != idx[:, 3])
    cross = idx[:, 0] + idx[:, 1] != idx[:, 2] + idx[:, 3]
    return int(keep.sum() + (keep & cross).sum())


def plain_own_flops(n: int, nterms: int, nphi: int) -> float:
    """Real flops of one mesh point along the plain kernels' own algorithm
    (complex multiply = 6, complex multiply-add = 8): the full-pivot Pfaffian
    (22 per trailing entry per step, 3 per pivot-scanned entry), Gauss-Jordan
    column operations on the stacked 3n x n matrix (24 n^3), the three density
    products (12 n^3), T assembly and phases (42 n^2), the regrouped list
    (26 per term), <J^2> / E_kin / traces (40 n^2), and the setup's
    gauge-independent products shared by the nphi^2 gauge points (8 n^3 / nphi^2)."""
    pf = 0.0
    for k in range(n - 1):
        m = 2 * n - 2 * k
        pf += 22.0 * (m - 2) * (m - 3) / 2 + 3.0 * m * (m - 1) / 2
    return pf + 36.0 * n ** 3 + 82.0 * n ** 2 + 26.0 * nterms + 8.0 * n ** 3 / (nphi * nphi)


def factored_flops(h: int, sum_b2: int, nphi: int, l_species: int, l_cross: int, nnz_pn: int,
                   schur: bool = False) -> float:
    """Algorithmic real flops of the factored form per (pair, Euler point),
    complex multiply-add = 8 flops.  Per species: R V1, R* U1, R Z1 (block-sparse,
    12 sum_b2 h), X, Y (8 h^3), M0 (8 sum_b2 h); then per gauge angle the
    Pfaffian of the 2h x 2h skew matrix (32 h^3 / 3), Gauss-Jordan on
    [T^T | W-right-hand sides] (20 h^3), rho, kappa10, kappa01 (12 h^3), the
    <J^2> ladder sums and traces (40 h^2) and the regrouped species list
    (28 flops per entry, 56 when its (cd, ab) partner term is present:
    ``l_species`` entries, ``l_cross`` of them with the partner).  Per Euler
    point once: G = W_pn rho_n (4 flops per CSR entry and phi_n), the nphi^2
    dot products (8 h^2 each) and the gauge sums (100 per gauge pair).  With
    ``schur`` the Pfaffian is the h x h one of the Schur complement (4 h^3 / 3)."""
    setup = 12.0 * sum_b2 * h + 8.0 * h ** 3 + 8.0 * sum_b2 * h
    pfaffian = (4.0 / 3.0 if schur else 32.0 / 3.0) * h ** 3
    per_phi = pfaffian + (20.0 + 12.0) * h ** 3 + 40.0 * h * h + 28.0 * (l_species + l_cross)
    combine = 4.0 * nnz_pn * nphi + 8.0 * nphi * nphi * h * h + 100.0 * nphi * nphi
    return 2.0 * (setup + nphi * per_phi) + combine


def traffic_per_launch(form: str, pairs: int):
    """HBM bytes per launch from the committed PMC measurement
    (profiles/traffic.json), scaled to ``pairs`` pairs per launch."""
    try:
        t = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))[form]
    except (OSError, KeyError, ValueError):
        return None
    b = 1024.0 * (t["fetch_kib_per_launch"] * t["fetch_factor"] + t["write_kib_per_launch"] * t["write_factor"])
    return b / t["pairs_per_launch"] * pairs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, threads: int):
    """The CPU oracle (this project's FP64 restatement of the reference call
    sequence, OpenMP over Euler points) on one full pair of the pf GCM
    (162,000 mesh points) plus one beta node of a strided sample of pairs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from sccm_amd.mesh import Mesh

    m = w.mesh
    o = oracle.Oracle(w.basis, w.inter, m, w.zval, w.nval, nthreads=threads)
    q0, q1 = w.pairs[3]
    t = time.perf_counter()
    o.kernel_pair(w.U[q0], w.V[q0], w.U[q1], w.V[q1])
    dt_full = time.perf_counter() - t
    log(f"bench: CPU oracle, full pair ({q0},{q1}): {dt_full:.1f} s")
    pts = m.n_points
    strided = list(range(10, len(w.pairs), 50))
    dt_s, pts_s = 0.0, 0
    for k, ip in enumerate(strided):
        ib = (7 * k) % m.bet.size
        sub = Mesh(m.alp, m.walp, m.bet[ib:ib + 1], m.wbet[ib:ib + 1], m.gam, m.wgam, m.nphi, m.jmax)
        os_ = oracle.Oracle(w.basis, w.inter, sub, w.zval, w.nval, nthreads=threads)
        p0, p1 = w.pairs[ip]
        t = time.perf_counter()
        os_.kernel_pair(w.U[p0], w.V[p0], w.U[p1], w.V[p1])
        dt_s += time.perf_counter() - t
        pts_s += sub.n_points
    return dict(value=(pts + pts_s) / (dt_full + dt_s), unit="mesh-pts/s", cores=threads, kind="port",
                cpu_model=cpu_model(), full_pair_mesh_pts_per_s=pts / dt_full,
                sample=f"pair ({q0},{q1}) of the pf GCM in full ({pts} mesh points, {dt_full:.2f} s) plus one beta "
                       f"node of pairs {strided} ({pts_s} mesh points, {dt_s:.2f} s); OpenMP over Euler points on "
                       f"{threads} threads; the 34.02 M-point GCM would take "
                       f"{34.02e6 * (dt_full + dt_s) / (pts + pts_s) / 60:.1f} min")


def gcm_compare(ctx, w, kern_a, kern_b, nlevels: int):
    """Hill-Wheeler levels of two kernel sets (a: chosen form, b: plain FP64).

    Levels are those of PreSolve_HWG's natural-state selection
    (``Codes/GCM/GCM_HWG.f90:2-100``) on the plain kernels: per J and level the
    number of natural states (NOS) of least variance, where the level sits on
    a plateau in NOS; the chosen form is solved at the same NOS.  Returns the
    largest |dE| over those levels (the metric's condition), a per-J report
    (norm spectrum, NOS, plateau, energies, |dE|) and, for comparison, the
    largest |dE| of HNDIAG's levels with every natural state above the norm
    cutoff, which the unconverged Euler mesh leaves unstable."""
    import numpy as np

    from sccm_amd.driver import gcm_energies, gcm_levels, levels_at_nos
    lv = gcm_levels(ctx, kern_b, w.pairs, len(w.U), w.mesh.jmax, w.hw_eps, kmax=nlevels)
    ea = levels_at_nos(ctx, kern_a, w.pairs, len(w.U), w.mesh.jmax, lv, w.hw_eps)
    worst, report = 0.0, {}
    for J, l in lv.items():
        d = [abs(x - y) if (x is not None and y is not None) else None for x, y in zip(ea[J], l["e"])]
        worst = max([worst] + [x for x in d if x is not None])
        report[str(J)] = {"dim": l["dim"], "natural_states_above_cutoff": l["m"],
                          "most_negative_norm_eigenvalue_rel": l["en_min_rel"],
                          "nos": l["nos"], "plateau_keV": [None if x is None else round(1e3 * x, 3) for x in l["plateau"]],
                          "E_plain_MeV": [round(x, 6) for x in l["e"]],
                          "abs_dE_keV": [None if x is None else float(1e3 * x) for x in d],
                          "E_all_above_cutoff_MeV": [round(x, 6) for x in l["e_eps"]]}
    e1 = gcm_energies(ctx, kern_a, w.pairs, len(w.U), w.mesh.jmax, w.hw_eps, nlevels=nlevels)
    e2 = gcm_energies(ctx, kern_b, w.pairs, len(w.U), w.mesh.jmax, w.hw_eps, nlevels=nlevels)
    worst_eps = max([0.0] + [float(np.abs(e1[J][:k] - e2[J][:k]).max()) for J in e1
                             for k in [min(len(e1[J]), len(e2[J]))] if k])
    return worst, report, worst_eps


def mesh_check(w, nang: int, nlevels: int) -> dict:
    """Levels of the same GCM on a refined alpha = gamma mesh (factored form),
    each by PreSolve_HWG's least-variance NOS on its own kernels, next to the
    levels of the benchmark mesh: shows which levels the 10-point alpha/gamma
    integration has converged."""
    from sccm_amd import configs
    from sccm_amd.driver import gcm_levels
    from sccm_amd.pnamp import FORM_FACTORED, PNAMP

    out = {}
    for tag, na in (("benchmark", w.mesh.alp.size), ("refined", nang)):
        wr = configs.pf_gcm(nstates=len(w.U), nalp=na, nbet=w.mesh.bet.size, nphi=w.mesh.nphi, jmax=w.mesh.jmax)
        c = PNAMP(wr.basis, wr.inter, wr.mesh)
        c.set_states(wr.U, wr.V)
        c.set_form(FORM_FACTORED)
        k = c.kernels_batched(wr.pairs, wr.zval, wr.nval)
        lv = gcm_levels(c, k, wr.pairs, len(wr.U), wr.mesh.jmax, wr.hw_eps, kmax=nlevels)
        out[tag] = {"alpha_gamma_points": na,
                    "per_J": {str(J): {"nos": l["nos"], "E_MeV": [round(x, 4) for x in l["e"]],
                                       "most_negative_norm_eigenvalue_rel": l["en_min_rel"],
                                       "E_all_above_cutoff_MeV": [round(x, 4) for x in l["e_eps"]]}
                              for J, l in lv.items()}}
        c.close()
        log(f"bench: mesh check alpha = gamma {na} done")
    return out


def kernel_errors(kern_a, kern_b, jmax: int):
    """Per J: the largest |a - b| / max|b| of the norm (N^J) and energy (H^J)
    kernels over all pairs."""
    import numpy as np
    out = {}
    for J in range(jmax + 1):
        e = {}
        for name, io in (("N", 0), ("H", 4)):
            b = kern_b[:, io, J]
            nb = np.abs(b).max()
            e[name] = float(np.abs(kern_a[:, io, J] - b).max() / nb) if nb > 0 else 0.0
        out[str(J)] = e
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs-per-step", type=int, default=0, help="0: all pairs")
    ap.add_argument("--nstates", type=int, default=20)
    ap.add_argument("--form", choices=("factored", "mixed-fp32", "mixed-bf16", "plain"), default="factored")
    ap.add_argument("--levels", type=int, default=3, help="lowest HW levels per J compared")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="factored form: keep the default launch configuration")
    ap.add_argument("--exact-pfaffian", action="store_true",
                    help="factored forms: keep the 2h x 2h Pfaffian (by default the h x h Schur-complement "
                         "Pfaffian is allowed for states with cond_1(Z_s) <= 1e6)")
    ap.add_argument("--tune", default="", help="factored form: 'pf,rest,div' launch configuration instead of "
                                               "autotuning (e.g. 6,3,2)")
    ap.add_argument("--mesh-check", type=int, default=0,
                    help="also solve the GCM on alpha = gamma = N points (same beta, gauge) and report its levels")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 512 MiB device-to-device copy (the reference kernel of the HBM counters)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain reference pass and the GCM check")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes": 512 << 20}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the pairs would shard across GPUs with no communication until the "
              "Hill-Wheeler solve); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from sccm_amd import configs
    from sccm_amd.pnamp import FORM_FACTORED, FORM_MIXED_BF16, FORM_MIXED_FP32, FORM_PLAIN, PNAMP

    log("bench: building the pf GCM workload")
    w = configs.pf_gcm(nstates=args.nstates)
    ctx = PNAMP(w.basis, w.inter, w.mesh)
    ctx.set_states(w.U, w.V)
    form = {"factored": FORM_FACTORED, "mixed-fp32": FORM_MIXED_FP32, "mixed-bf16": FORM_MIXED_BF16,
            "plain": FORM_PLAIN}[args.form]
    ctx.set_form(form)
    if form != FORM_PLAIN and not args.exact_pfaffian:
        ctx.set_schur_pfaffian(True)
    schur = form != FORM_PLAIN and ctx.schur_pfaffian_active
    log(f"bench: Pfaffian route {'h x h Schur complement' if schur else '2h x 2h'}")
    npairs = len(w.pairs)
    P = args.pairs_per_step or npairs

    def batch(s):
        idx = [(s * P + i) % npairs for i in range(P)]
        return w.pairs[idx]

    tune = None
    if form == FORM_FACTORED and args.tune:
        pf_w, rest_w, div = (int(x) for x in args.tune.split(","))
        ctx.set_tune(pf_w, rest_w, div)
        tune = {"pf_waves": pf_w, "rest_waves": rest_w, "chunk_div": div, "set": "given, not timed"}
    elif form == FORM_FACTORED and not args.no_tune:
        t = time.perf_counter()
        tune = ctx.autotune(batch(0), w.zval, w.nval)
        tune["seconds"] = time.perf_counter() - t
        log(f"bench: autotuned in {tune['seconds']:.1f} s: Pfaffian {tune['pf_waves']} waves/SIMD, "
            f"densities {tune['rest_waves']}, chunk / {tune['chunk_div']}")
    for s in range(args.warmup):
        ctx.kernels_batched(batch(s), w.zval, w.nval)
        log(f"bench: warmup step {s} done")
    k_ms = 0.0
    kern_chosen = None
    # Timed steps: all device work of a step (mesh-point kernels, gauge sums,
    # J-projection), the projected kernels left in HBM; the copy of the 6 x 9 x
    # 17 x 17 complex kernels per pair (52 MB per step) to the host is timed
    # separately below (the PCIe-inclusive rate).  Every library call returns
    # after synchronising its stream, so the device work of each timed step is
    # complete when its call returns (torch is not used here: its own HIP
    # runtime cannot share the process with the library's).
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.kernels_batched(batch(args.warmup + s), w.zval, w.nval, want_kern=False)
        ms, _ = ctx.timing()
        k_ms += ms
        log(f"bench: step {s} mesh-point kernels {ms:.1f} ms")
    elapsed = time.perf_counter() - t0
    pts_step = P * w.mesh.n_points
    value = args.steps * pts_step / elapsed
    nh = min(args.steps, 5)
    t1 = time.perf_counter()
    for s in range(nh):
        k = ctx.kernels_batched(batch(args.warmup + s), w.zval, w.nval)
        if P == npairs:
            kern_chosen = k
    pcie = {"value": nh * pts_step / (time.perf_counter() - t1), "unit": "mesh-pts/s", "steps": nh,
            "copy_bytes_per_step": int(P * np.prod(ctx._kshape()) * 16),
            "note": "the same steps with the projected kernels copied to a pageable host array"}
    log(f"bench: {value / 1e6:.1f} M mesh-pts/s (kernels resident in HBM), "
        f"{pcie['value'] / 1e6:.1f} M with the copy to the host")
    k_avg = k_ms / args.steps
    h = w.n // 2
    if form != FORM_PLAIN:
        sum_b2 = sum(s * s for (_, s, _) in w.basis.blocks()) // 2
        idx = np.stack([w.inter.ka, w.inter.kb, w.inter.kc, w.inter.kd], 1) - 1
        same = ((idx < h).all(1)) | ((idx >= h).all(1))
        cross = idx[:, 0] + idx[:, 1] != idx[:, 2] + idx[:, 3]
        l_sp = int(same.sum()) // 2
        l_cr = int((same & cross).sum()) // 2
        pn = (idx[:, 0] < h) & (idx[:, 1] >= h) & (idx[:, 2] < h) & (idx[:, 3] >= h)
        nz = set()
        for a, b, c, d in idx[pn]:
            bl, dl = b - h, d - h
            nz.add((c + a * h, dl + bl * h))
            if a + b != c + d:
                nz.add((a + c * h, bl + dl * h))
        fl = factored_flops(h, sum_b2, w.mesh.nphi, l_sp, l_cr, len(nz), schur) * w.mesh.n_euler * P
        kname = "k_fac_setup + k_fac_phi + k_fac_combine"
        tform = {FORM_FACTORED: "factored_fp64", FORM_MIXED_FP32: "mixed_fp32", FORM_MIXED_BF16: "mixed_bf16"}[form]
        if form != FORM_FACTORED:
            kname += " + k_vgemm"
    else:
        fl = plain_flops_per_point(w.n, w.inter.nme) * pts_step
        kname, tform = "k_plain_setup + k_plain_gauge + k_plain_sum", "plain_fp64"
    achieved = fl / (k_avg * 1e-3) / 1e12
    # the reference call sequence's count (147 n^3 + 190 L per mesh point) at
    # the same time: what the step's work costs done the reference's way
    plain_equiv = plain_flops_per_point(w.n, w.inter.nme) * pts_step / (k_avg * 1e-3) / 1e12
    form_text = {
        FORM_FACTORED: "factored FP64: per-species Pfaffians and densities, exact algebraic refactoring of the "
                       "reference sequence, FP64 throughout, batched launches replayed from a captured hipGraph; "
                       "no MFMA contraction, no surrogate",
        FORM_MIXED_FP32: "factored, V contraction as FP32 GEMMs on MFMA, everything else FP64, hipGraph",
        FORM_MIXED_BF16: "factored, V contraction as split-BF16 GEMMs on MFMA (FP32 accumulation), everything "
                         "else FP64, hipGraph",
        FORM_PLAIN: "plain FP64 drop-in, any states: per (pair, Euler point) the gauge-independent rotation "
                    "products, then per gauge point the full-pivot Pfaffian, T^-1 by Gauss-Jordan, densities, "
                    "regrouped energy list, <J^2>; FP64 throughout",
    }[form]
    out = {
        "metric": "mesh-pts/s: mixed+batched+surrogate vs plain FP64 GPU at |dE_GCM|<1 keV",
        "value": value,
        "unit": "mesh-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {FORM_MIXED_FP32: "f64+f32(mfma)", FORM_MIXED_BF16: "f64+bf16x3(mfma)"}.get(form, "f64"),
        "data": "synthetic (seeded pf-space interaction and HFB states)",
        "form_measured": form_text,
        "pcie_inclusive": pcie,
        "config": {
            "workload": "pf_gcm: pf space n=40, Z,N=2,4, 20 HFB states / 210 pairs, "
                        "alpha=gamma 10 x beta 20 Euler x 9x9 gauge = 162000 mesh points per pair, J<=8",
            "form": tform,
            "pairs_per_step": P,
            "mesh_points_per_step": pts_step,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic_per_launch(tform, P),
            "kernel": kname,
            "algorithmic_flops_per_launch": fl,
            "flops_count": ("the measured form's own algorithm" if form != FORM_PLAIN
                            else "the reference call sequence, 147 n^3 + 190 L per mesh point"),
            "kernel_ms_per_launch": k_avg,
            "reference_sequence_equivalent_tflops": plain_equiv,
        },
    }
    if form != FORM_PLAIN:
        out["pfaffian"] = {
            "route": "h x h Schur complement" if schur else "2h x 2h (ZPfaffianF's matrix)",
            "note": ("Pf(S_s) = Pf(-Z_s) Pf(e^{2i phi} M0 - Z_s^-1), (-Z_s)^-1 and Pf(-Z_s) once per state; "
                     "taken while cond_1(Z_s) <= 1e6 for every state, rounding up to about cond(Z_s) eps: "
                     "see gcm.max_rel_kernel_diff against the plain FP64 path"
                     if schur else "the reference's Pfaffian matrix, partial pivoting"),
        }
    if tune is not None:
        out["autotune"] = tune
    if form == FORM_PLAIN:
        own = plain_own_flops(w.n, plain_terms(w.inter, w.n), w.mesh.nphi) * pts_step
        out["roofline"]["own_algorithm_flops_per_launch"] = own
        out["roofline"]["own_algorithm_tflops"] = own / (k_avg * 1e-3) / 1e12
        out["roofline"]["own_algorithm_frac"] = own / (k_avg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
    if form != FORM_PLAIN and not args.no_plain:
        log("bench: plain FP64 reference pass over all pairs")
        ctx.set_form(FORM_PLAIN)
        t = time.perf_counter()
        kern_plain = ctx.kernels_batched(w.pairs, w.zval, w.nval)
        tp = time.perf_counter() - t
        plain_value = npairs * w.mesh.n_points / tp
        out["plain_fp64"] = {"value": plain_value, "unit": "mesh-pts/s", "pairs": npairs,
                             "kernel_ms": ctx.timing()[0]}
        out["speedup_vs_plain"] = value / plain_value
        if kern_chosen is None:
            ctx.set_form(form)
            kern_chosen = ctx.kernels_batched(w.pairs, w.zval, w.nval)
        rel = float(np.abs(kern_chosen - kern_plain).max() / np.abs(kern_plain).max())
        dE, report, dE_eps = gcm_compare(ctx, w, kern_chosen, kern_plain, args.levels)
        out["gcm"] = {"max_abs_dE_keV": 1e3 * dE,
                      "levels": "lowest %d per J at PreSolve_HWG's NOS of least variance (plain kernels)" % args.levels,
                      "hw_norm_cutoff": w.hw_eps, "max_rel_kernel_diff": rel,
                      "per_J": report,
                      "max_abs_dE_keV_all_natural_states_above_cutoff": 1e3 * dE_eps,
                      "kernel_rel_err_per_J": kernel_errors(kern_chosen, kern_plain, w.mesh.jmax)}
    if args.mesh_check and "gcm" in out:
        out["mesh_check"] = mesh_check(w, args.mesh_check, args.levels)
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        log(f"bench: CPU oracle sample on {threads} threads ({cpu_model()})")
        out["cpu_baseline"] = cpu_baseline(w, threads)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
