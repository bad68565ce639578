"""Benchmark: particle-substeps/s of the explicit APIC MPM substep with AnisoMPM
damage on one MI355X, FP32 form (value) next to the FP64 plain form.

Workload (default): AnisoMPM phase-field damage, 1,000,000 particles in a
1 x 0.2 x 0.2 m slab, dx = 1/128, QRStableNeoHookean E=1e4 nu=0.3, per-particle
fibre +-15 deg about x, alpha=-0.9, l0=2dx, sigma_c=5e3 Pa, top face pulled by a
moving sticky box, dt=1e-4 (seed 2).  Inputs are uploaded to HBM before the
timed region; one "step" is one full substep (binning + sort, stress + P2G,
grid update, G2P + damage + F update).

Prints ONE JSON line.  `--gpus N` with N > 1 exits: the substep runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-substeps/sec FP32-batched vs FP64-plain on 1 MI355X; % HBM peak"


def algorithmic_bytes(stage, precision, n_particles, n_nodes, damage):
    """Compulsory HBM bytes of one launch of a stage: every particle field the
    stage must read or write once, plus its grid accumulators / velocities once
    per active node (index and sort arrays excluded)."""
    w = 4 if precision == "fp32" else 8
    acc = 8  # int64 fixed point (FP32 form) or double accumulator
    nacc = 9 if damage else 7
    dmg = 1 if damage else 0
    if stage == "p2g":
        # read x v C (15) + symmetric stress (6) + m (+ d) + material id; nacc accumulators per node
        per_p = (15 + 6 + 1 + dmg) * w + 4
        per_n = nacc * acc
    elif stage == "g2p":
        # read x F m V0 d g (16) + id + material; write x v C F stress m V0 d g lap (35) + id + material;
        # read new_v (+ phase field) per node
        per_p = 16 * w + 8 + 35 * w + 8
        per_n = (3 + dmg) * w
    elif stage == "grid_update":
        per_p = 0
        per_n = nacc * acc + (3 + dmg) * w
    else:
        raise ValueError(stage)
    return per_p * n_particles + per_n * n_nodes


def cpu_baseline(scene, budget_s=12.0, max_steps=10):
    """The FP64 oracle (CPU restatement of the reference step) on the host cores."""
    from ziran2020_amd import _abi as A
    lib = A.ZmpmLib(os.path.join(ROOT, "oracle", "libzmpm_oracle.so"))
    sim = scene.make(lib, A.PREC_FP64)
    t0 = time.perf_counter()
    sim.run(1, scene.dt, 0.0)
    one = time.perf_counter() - t0
    steps = int(max(1, min(max_steps, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    sim.run(steps, scene.dt, scene.dt)
    el = time.perf_counter() - t0
    sim.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": scene.n * steps / el, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} substeps of the full {scene.n}-particle workload after 1 untimed substep, "
                      f"FP64 oracle (C++/OpenMP restatement of the reference CPU step), {threads} threads"}


def timed(scene, lib, precision, steps, warmup, timing_steps=50, autotune=True):
    """Launch configuration from the autotuner (state left unchanged), then
    throughput of `steps` substeps (HIP-graph replays, no events inside), then
    per-stage kernel times from HIP events on the handle's stream over the next
    min(steps, timing_steps) substeps launched one kernel at a time."""
    sim = scene.make(lib, precision)
    timed.launch_config = sim.autotune(scene.dt, 0.0, steps=4) if autotune else sim.launch_config()
    sim.run(warmup, scene.dt, 0.0)
    sim.synchronize()
    t_start = warmup * scene.dt
    t0 = time.perf_counter()
    sim.run(steps, scene.dt, t_start)
    sim.synchronize()
    el = time.perf_counter() - t0
    k = min(steps, timing_steps)
    sim.L.zmpm_set_timing(sim.h, 1)
    sim.run(k, scene.dt, t_start + steps * scene.dt)
    sim.synchronize()
    sim.L.zmpm_set_timing(sim.h, 0)
    stages = sim.timings()
    grid = sim.grid()
    n_nodes = int(grid["m"].shape[0])
    out = sim.download()
    sim.close()
    return el, stages, n_nodes, out


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB, ten times: the reference
    kernel whose FETCH_SIZE / WRITE_SIZE counters (known bytes per dispatch)
    calibrate the per-dispatch HBM traffic figures of the substep kernels."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "dispatches": 10}), flush=True)


def accuracy_after(scene, lib, steps=100):
    """Per-field normwise error (max |a - b| / max |b|) of the FP32 form against
    the FP64 plain form after exactly `steps` substeps from the same upload (the
    FP64 form matches the FP64 CPU oracle to 1e-9 in the parity tests)."""
    from ziran2020_amd import _abi as A
    out = {}
    for prec in (A.PREC_FP32, A.PREC_FP64):
        sim = scene.make(lib, prec)
        sim.run(steps, scene.dt, 0.0)
        out[prec] = sim.download()
        sim.close()
    errs = {}
    for k in ("x", "v", "C", "F", "d", "g"):
        a, b = out[A.PREC_FP32][k], out[A.PREC_FP64][k]
        den = np.abs(b).max()
        errs[k] = float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a).max())
    return errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="damage_slab")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch configuration")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain MiB-sized device copy (traffic-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: the substep runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import ziran2020_amd
    from ziran2020_amd import _abi as A, scenes

    lib = ziran2020_amd.load_library()
    scene = scenes.WORKLOADS[args.workload]()
    damage = any(m.has_damage for m in scene.materials)

    el32, st32, nn32, out32 = timed(scene, lib, A.PREC_FP32, args.steps, args.warmup, autotune=not args.no_autotune)
    cfg32 = timed.launch_config
    value = scene.n * args.steps / el32
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": scene.n, "dx": scene.dx, "dt": scene.dt,
                   "form": "FP32 particle state, int64 fixed-point P2G", "active_nodes": nn32,
                   "launch": cfg32, "autotuned": not args.no_autotune, "options": dict(scene.options)},
    }
    # roofline of the dominant kernel stage (HIP events on the handle's stream)
    stage_ms = {k: v for k, v in st32.items() if k in ("p2g", "g2p", "grid_update")}
    dom = max(stage_ms, key=stage_ms.get)
    byt = algorithmic_bytes(dom, "fp32", scene.n, nn32, damage)
    achieved = byt / (stage_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(f"{args.workload}/fp32/{dom}")
    result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                          "algorithmic_bytes": byt, "avg_ms": stage_ms[dom]}
    result["stage_ms"] = st32
    if not args.no_fp64:
        el64, st64, nn64, out64 = timed(scene, lib, A.PREC_FP64, args.steps, args.warmup,
                                        autotune=not args.no_autotune)
        v64 = scene.n * args.steps / el64
        errs = {}
        for k in ("x", "v", "C", "F", "d"):
            den = np.abs(out64[k]).max()
            errs[k] = float(np.abs(out32[k] - out64[k]).max() / den) if den > 0 else float(np.abs(out32[k]).max())
        result["fp64_plain"] = {"value": v64, "ms_per_step": 1e3 * el64 / args.steps, "stage_ms": st64,
                                "dtype": "fp64", "launch": timed.launch_config}
        result["fp32_over_fp64"] = value / v64
        result["fp32_vs_fp64_normwise_err"] = {"substeps": args.warmup + args.steps + min(args.steps, 50), **errs}
        result["fp32_vs_fp64_after_100_substeps"] = accuracy_after(scene, lib, 100)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
