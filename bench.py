"""Benchmark of the PNAMP projected-kernel path on one MI355X.

Workload: the pf-space GCM (n = 40 m-states, Z,N valence 2,4, 20 HFB states,
210 pairs q <= q', Euler mesh alpha = gamma 10, beta 20, 9 x 9 gauge points,
J <= 8): 162,000 mesh points per pair, 34.02 M for the whole GCM.  A step is
one batch of ``--pairs-per-step`` pairs (cycling through the 210) evaluated by
the plain FP64 HIP path: rotation, Pfaffian norm, densities, energy, gauge sum
and J-projection, inputs resident in HBM.

Prints one JSON line.  ``value`` is mesh points per second of the whole job
(host wall clock around the timed steps, which includes the projection
kernel and the copy of the projected kernels back to the host).
``roofline`` is the mesh-point kernel's algorithmic FP64 rate from HIP events
on the library's stream.  ``cpu_baseline`` times the CPU oracle (this
project's C restatement of the reference call sequence) on a bounded sample
of the same workload on the host cores.
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


def flops_per_point(n: int, nme: int) -> float:
    """Algorithmic real FP64 flops of one mesh point, the reference call
    sequence (complex multiply-add = 8 flops): 16 n^3 rotation (2 ZGEMM),
    16 n^3 Pfaf_Over inverse + ZGEMM, 19 n^3 Pfaffian of the 2n skew matrix,
    96 n^3 ROKAPPAphiomega (11 ZGEMM + inverse), 190 per interaction entry."""
    return 147.0 * n ** 3 + 190.0 * nme


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


def cpu_baseline(w, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    from sccm_amd.mesh import Mesh

    m = w.mesh
    nb = 4  # the first four beta nodes x all alpha, gamma x all gauge points
    sub = Mesh(m.alp, m.walp, m.bet[:nb], m.wbet[:nb], m.gam, m.wgam, m.nphi, m.jmax)
    o = oracle.Oracle(w.basis, w.inter, sub, w.zval, w.nval, nthreads=threads)
    q0, q1 = w.pairs[3]
    t = time.perf_counter()
    o.kernel_pair(w.U[q0], w.V[q0], w.U[q1], w.V[q1])
    dt = time.perf_counter() - t
    pts = sub.n_points
    return dict(value=pts / dt, unit="mesh-pts/s", cores=threads, kind="port",
                sample=f"pair ({q0},{q1}) of the pf GCM, {sub.n_euler} Euler points ({nb} of 20 beta nodes) "
                       f"x {sub.n_gauge} gauge = {pts} mesh points, {dt:.2f} s, OpenMP over Euler points")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs-per-step", type=int, default=16)
    ap.add_argument("--nstates", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the pairs would shard across GPUs with no communication until the "
              "Hill-Wheeler solve); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from sccm_amd import configs
    from sccm_amd.pnamp import PNAMP

    log("bench: building the pf GCM workload")
    w = configs.pf_gcm(nstates=args.nstates)
    ctx = PNAMP(w.basis, w.inter, w.mesh)
    ctx.set_states(w.U, w.V)
    P = args.pairs_per_step
    npairs = len(w.pairs)

    def batch(s):
        idx = [(s * P + i) % npairs for i in range(P)]
        return w.pairs[idx]

    for s in range(args.warmup):
        ctx.kernels_batched(batch(s), w.zval, w.nval)
        log(f"bench: warmup step {s} done")
    k_ms = 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.kernels_batched(batch(args.warmup + s), w.zval, w.nval)
        ms, _ = ctx.timing()
        k_ms += ms
        log(f"bench: step {s} k_point {ms:.1f} ms")
    elapsed = time.perf_counter() - t0
    pts_step = P * w.mesh.n_points
    value = args.steps * pts_step / elapsed
    fpp = flops_per_point(w.n, w.inter.nme)
    k_avg = k_ms / args.steps
    achieved = fpp * pts_step / (k_avg * 1e-3) / 1e12
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
        "dtype": "f64",
        "data": "synthetic (seeded pf-space interaction and HFB states)",
        "config": {
            "workload": "pf_gcm: pf space n=40, Z,N=2,4, 20 HFB states / 210 pairs, "
                        "alpha=gamma 10 x beta 20 Euler x 9x9 gauge = 162000 mesh points per pair, J<=8",
            "form": "plain_fp64",
            "pairs_per_step": P,
            "mesh_points_per_step": pts_step,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic_per_launch("plain_fp64", P),
            "kernel": "k_point",
            "flops_per_point": fpp,
            "kernel_ms_per_launch": k_avg,
        },
    }
    if not args.no_cpu:
        threads = min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 16)
        log(f"bench: CPU oracle sample on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(w, threads)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
