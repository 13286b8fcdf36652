"""Benchmark: particle-steps/s of the extended Matter MPM step on one MI355X.

Workload: paint_4blob_256x64x256 — 1,048,576 FP32 particles in four
64x8x64-cell paint blobs on a 256x64x256 lattice (dx = 3.125 mm), von Mises /
Perzyna (yield 50 Pa, viscosity 10 Pa s), slab knife on a 40 mm circle at
0.15 m/s, dt = 1e-5 s.  Synthetic seeded inputs (no dataset exists).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints ONE JSON line.  `value` = live particles x K / wall seconds of K steps
(graph-launched, inputs resident in HBM, synchronised before and after).
`roofline` is for the dominant kernel of the step (by time), its mean
duration measured with HIP events on the library's stream over a second run of
K steps; every kernel's figure is listed under roofline.kernels.
`cpu_baseline` times the CPU restatement of the reference step (OpenMP, the
reference's cost structure) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 1M particles, learned vs plain return map, ≤1e-3 rel err"


# Algorithmic bytes per launch (FP32; particle rows and active-node rows the
# kernel must move): G2P reads x v F pig eps (92 B) and writes x v C F pig diff
# eps dgamma (136 B) per particle and reads v flip pig shear (56 B) per active
# node; P2G reads x v C F-stress pig diff (128 B) per particle and writes mass,
# momentum, force, pig, shear (60 B) per active node; the grid update reads 60 B
# and writes 56 B per active node.
def kernel_bytes(np_live, n_active):
    return {
        "k_g2p": 228.0 * np_live + 56.0 * n_active,
        "k_p2g": 128.0 * np_live + 60.0 * n_active,
        "k_grid": 116.0 * n_active,
    }


def step_bytes(np_live, n_active):
    return 356.0 * np_live + 232.0 * n_active


def cpu_baseline(scene, steps=2, threads=None):
    """Reference-structured CPU step (oracle/_build/liboracle_fast.so) on the same scene."""
    from tests import oracle_py as orc
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    o = orc.from_scene(scene, fast=True, n_threads=threads)
    o.step(1)  # warm-up (allocations, first touch)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    n_act = o.stats()["n_active"]
    return {"value": n_act * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {scene['x'].shape[0]}-particle scene after 1 warm-up step, "
                      f"{threads} OpenMP threads, per-thread grid replicas + critical merge, 64-node weight cache "
                      f"(reference cost structure), -O3 -ffast-math"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU sharding is design-only); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from pigment_mixing_amd import scenes

    scene = scenes.paint_4blob()
    n = scene["x"].shape[0]
    print(f"[bench] scene {scene['name']} particles={n}", file=sys.stderr, flush=True)
    h = scenes.make_handle(scene)
    h.step(args.warmup)
    h.sync()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    h.step(args.steps)
    h.sync()
    wall = time.perf_counter() - t0
    st = h.stats()
    live = st.n_active
    value = live * args.steps / wall
    print(f"[bench] timed {args.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)

    kt = h.profile(args.steps)
    n_act = kt.active_nodes_per_step
    kb = kernel_bytes(kt.particles_per_step, n_act)
    kms = {"k_g2p": kt.ms_g2p, "k_p2g": kt.ms_p2g, "k_grid": kt.ms_grid}
    kgbs = {k: kb[k] / (kms[k] * 1e-3) / 1e9 for k in kb}
    dom = max(kms, key=kms.get)
    step_gbs = step_bytes(kt.particles_per_step, n_act) / (kt.ms_step * 1e-3) / 1e9
    print(f"[bench] kernels ms: g2p={kt.ms_g2p:.4f} sort={kt.ms_sort:.4f} p2g={kt.ms_p2g:.4f} "
          f"grid={kt.ms_grid:.4f} step={kt.ms_step:.4f} active_nodes={n_act:.0f} "
          f"GB/s g2p={kgbs['k_g2p']:.0f} p2g={kgbs['k_p2g']:.0f} grid={kgbs['k_grid']:.0f}",
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, steps=args.cpu_steps)
        print(f"[bench] cpu baseline {cpu['value']:.3e} particle-steps/s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded box blobs, Mixbox latents of palette colours)",
        "config": {
            "workload": scene["name"],
            "particles": n,
            "lattice": [256, 64, 256],
            "dx_m": scenes.DX,
            "dt_s": float(scene["cfg"].dt),
            "return_map": "exact von Mises/Perzyna",
            "knife": "slab 60x2x30 mm tilted 30 deg, circular sweep r=40 mm at 0.15 m/s",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": kgbs[dom],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kgbs[dom] / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": kb[dom],
            "kernel_ms": kms[dom],
            "kernels": {k: {"ms": kms[k], "GB/s": kgbs[k], "frac": kgbs[k] / HBM_PEAK_GBS,
                            "algorithmic_bytes": kb[k]} for k in kb},
            "sort_ms": kt.ms_sort,
            "step_ms_events": kt.ms_step,
            "step_frac": step_gbs / HBM_PEAK_GBS,
            "active_nodes": n_act,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
