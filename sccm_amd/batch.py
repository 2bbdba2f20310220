"""Uneven batch of GCM instances across model spaces (config C4).

64 GCM instances drawn (seeded) from the sd (n=24), pf (n=40) and f5pg9
(n=44) spaces, valence Z, N in {2, 4, 6} (even: the synthetic HFB vacua keep
even number parity, so U is invertible as zbareover needs), 5-30 configurations
each on a (beta2, gamma) grid, all pairs q <= q', mesh alpha=gamma 10, beta 20,
gauge 9 per species, J <= 8.  Interaction per space with the sd recipe
(J-coupled TBMEs N(0, 1.5) MeV, single-particle energies U[-5, 0] MeV).

One device context per model space holds that space's basis, interaction
and mesh and the states of all its instances (one pnamp_set_states); all pairs
of all its instances run as one batched call with per-pair valence numbers
(pnamp_kernels_batched_zn): one captured launch sequence per space covers
every instance.  The plain FP64 path is timed on a seeded sample of pairs of
every instance and its whole-batch time extrapolated per space.
"""
from __future__ import annotations

import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import configs
from .basis import make_basis
from .interaction import synthetic_interaction
from .mesh import make_mesh

SPACES = ("sd", "pf", "f5pg9")
SEEDS = {"sd": (1001, 1002), "pf": (1021, 1022), "f5pg9": (1031, 1032)}


@dataclass
class Instance:
    space: str
    zval: int
    nval: int
    nconf: int
    seed: int

    @property
    def npairs(self) -> int:
        return self.nconf * (self.nconf + 1) // 2


def draw_instances(n: int = 64, seed: int = 4001) -> list[Instance]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(Instance(SPACES[int(rng.integers(3))], int(2 * rng.integers(1, 4)), int(2 * rng.integers(1, 4)),
                            int(rng.integers(5, 31)), seed + 10 * (i + 1)))
    return out


def instance_states(inst: Instance, basis):
    coords = configs.gcm_grid(inst.nconf)
    U, V = configs.make_states(basis, inst.zval, inst.nval, coords, seed=inst.seed)
    pairs = np.array([(q, q1) for q in range(inst.nconf) for q1 in range(q, inst.nconf)], dtype=np.int32)
    return U, V, pairs


def run(n: int = 64, seed: int = 4001, nalp: int = 10, nbet: int = 20, nphi: int = 9, jmax: int = 8,
        form: str = "factored", plain_pairs: int = 2, log=print, repeats: int = 2,
        schur_pfaffian: bool = False) -> dict:
    from .pnamp import FORM_FACTORED, FORM_MIXED_BF16, FORM_MIXED_FP32, FORM_PLAIN, PNAMP
    fcode = {"factored": FORM_FACTORED, "mixed-fp32": FORM_MIXED_FP32, "mixed-bf16": FORM_MIXED_BF16}[form]
    insts = draw_instances(n, seed)
    mesh = make_mesh(nalp, nbet, nalp, nphi, jmax)
    rng = np.random.default_rng(seed + 1)
    per_space = {}
    tot_pts, t_chosen, t_plain_est, worst_rel = 0, 0.0, 0.0, 0.0
    for sp in SPACES:
        mine = [i for i in insts if i.space == sp]
        if not mine:
            continue
        b = make_basis(sp)
        ctx = PNAMP(b, synthetic_interaction(b, 1.5, *SEEDS[sp]), mesh)
        Us, Vs, pairs, zv, nv, sample = [], [], [], [], [], []
        off = 0
        for inst in mine:
            U, V, pr = instance_states(inst, b)
            Us.append(U)
            Vs.append(V)
            first = len(pairs)
            pairs.extend((p + off, q + off) for p, q in pr)
            zv.extend([inst.zval] * len(pr))
            nv.extend([inst.nval] * len(pr))
            sample.extend(first + int(k) for k in rng.choice(len(pr), size=min(plain_pairs, len(pr)), replace=False))
            off += inst.nconf
        pairs = np.array(pairs, dtype=np.int32)
        zv, nv = np.array(zv, dtype=np.int32), np.array(nv, dtype=np.int32)
        ctx.set_states(np.concatenate(Us), np.concatenate(Vs))
        ctx.set_form(fcode)
        ctx.set_schur_pfaffian(schur_pfaffian)  # taken only if every state's cond_1(Z_s) <= 1e6
        schur = ctx.schur_pfaffian_active
        tune = None
        if form == "factored":  # launch configuration timed on a sample of this space's pairs
            t = time.perf_counter()
            tune = ctx.autotune(pairs[::max(1, len(pairs) // 96)], int(zv[0]), int(nv[0]))
            tune["seconds"] = time.perf_counter() - t
        kern = ctx.kernels_batched_zn(pairs, zv, nv)  # first call captures the launch graph
        times = []
        for _ in range(repeats):
            t = time.perf_counter()
            kern = ctx.kernels_batched_zn(pairs, zv, nv)
            times.append(time.perf_counter() - t)
        dt = min(times)
        dev_ms = ctx.timing()[0]
        captures, replays, nodes = ctx.graph_info()
        pts = len(pairs) * mesh.n_points
        sample = np.array(sample)
        ctx.set_form(FORM_PLAIN)
        t = time.perf_counter()
        kp = ctx.kernels_batched_zn(pairs[sample], zv[sample], nv[sample])
        tp = time.perf_counter() - t
        rel = float(np.abs(kern[sample] - kp).max() / np.abs(kp).max())
        worst_rel = max(worst_rel, rel)
        plain_est = tp / (len(sample) * mesh.n_points) * pts
        per_space[sp] = {"instances": len(mine), "states": off, "pairs": int(len(pairs)), "mesh_points": pts,
                         "chosen_s": dt, "chosen_mesh_pts_per_s": pts / dt, "plain_sample_pairs": int(len(sample)),
                         "plain_s_extrapolated": plain_est, "rel_kernel_diff_vs_plain": rel,
                         "kernel_ms_device": dev_ms, "graph_captures": captures, "graph_replays": replays,
                         "kernel_launches_per_call": nodes, "autotune": tune,
                         "pfaffian": {2: "h x h Schur complement (every state)",
                                      1: "h x h Schur complement for the bra states that pass, 2h x 2h for the rest",
                                      0: "2h x 2h"}[schur]}
        tot_pts += pts
        t_chosen += dt
        t_plain_est += plain_est
        log(f"batch: {sp}: {len(mine)} instances, {off} states, {len(pairs)} pairs in one call, "
            f"{pts / dt / 1e6:.1f} M mesh-pts/s, rel diff vs plain {rel:.1e}")
        ctx.close()
    return {
        "metric": "mesh-pts/s, uneven GCM batch (sd/pf/f5pg9)", "form": form,
        "instances": n, "mesh_points": tot_pts, "chosen_seconds": t_chosen,
        "chosen_mesh_pts_per_s": tot_pts / t_chosen,
        "plain_fp64_mesh_pts_per_s_extrapolated": tot_pts / t_plain_est,
        "speedup_vs_plain": t_plain_est / t_chosen,
        "max_rel_kernel_diff_vs_plain_on_sampled_pairs": worst_rel,
        "calls": "one set_states and one batched call per model space (per-pair valence numbers)",
        "per_space": per_space,
        "plain_sample": f"{plain_pairs} seeded pairs per instance, whole-batch plain time extrapolated per space",
    }


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--form", default="factored", choices=("factored", "mixed-fp32", "mixed-bf16"))
    ap.add_argument("--plain-pairs", type=int, default=2)
    ap.add_argument("--schur-pfaffian", action="store_true",
                    help="allow the h x h Schur-complement Pfaffian (states with cond_1(Z_s) <= 1e6)")
    a = ap.parse_args()
    r = run(a.instances, form=a.form, plain_pairs=a.plain_pairs,
            log=lambda *x: print(*x, file=sys.stderr, flush=True), schur_pfaffian=a.schur_pfaffian)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
