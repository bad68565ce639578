"""Benchmark: particle-substeps/s of the explicit APIC MPM substep with AnisoMPM
damage on one MI355X, FP32 batched form (value) next to the FP64 plain form.

Workload (default, `batched_scenes`): 64 independent scenes in one launch per
stage (one handle of zmpm_create_batch) -- cubes of 10k..500k particles
(log-uniform, seed 4, 10,366,485 in all), each scene with its own dx (1/96,
1/128, 1/160), dt (8e-5..1.25e-4), start time, tilted gravity, APIC/RPIC ratio
and colliders (walls, a sticky / slipping / separating floor, moving spheres,
spinning boxes, growing capped cylinders); even scenes QRStableNeoHookean with
AnisoMPM phase-field damage, odd scenes fixed-corotated or Neo-Hookean; every
substep a HIP-graph replay.  Inputs are uploaded to HBM before the timed
region; one "step" is one full substep (binning + sort, P2G, grid update, G2P +
damage + F update + next stress) of every scene.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: the substep runs on one GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak without packed FP32: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
METRIC = "particle-substeps/sec FP32-batched vs FP64-plain on 1 MI355X; % HBM peak"
ORACLE_LIB = os.path.join(ROOT, "oracle", "libzmpm_oracle.so")


def code_only(text):
    """C++ source without its comments and blank lines (string and character
    literals kept as written)."""
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c in "\"'":
            j = i + 1
            while j < n and text[j] != c:
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            while i < n and text[i] != "\n":
                i += 1
        elif text.startswith("/*", i):
            k = text.find("*/", i + 2)
            i = n if k < 0 else k + 2
        else:
            out.append(c)
            i += 1
    lines = (ln.rstrip() for ln in "".join(out).split("\n"))
    return "\n".join(ln for ln in lines if ln)


def source_digest():
    """Digest of the device sources' code (comments aside): a traffic or
    instruction count measured on other kernels is reported as stale, not used."""
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "ziran2020_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".hpp", ".hip")):
            with open(os.path.join(csrc, name), encoding="utf-8") as f:
                h.update(name.encode())
                h.update(code_only(f.read()).encode())
    return h.hexdigest()[:16]


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads the oracle uses: OMP_NUM_THREADS (16 on the GPU box), else the CPUs
    this process may run on."""
    if os.environ.get("OMP_NUM_THREADS"):
        return int(os.environ["OMP_NUM_THREADS"])
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_sample_scene(name, scene, max_particles):
    """A bounded sample of the workload for the CPU oracle: the smallest scenes of
    a batch (each exactly as in the full batch), else the workload itself, or
    the workload's accuracy sample (its geometry with fewer particles) when the
    workload is above max_particles and the sample is smaller."""
    from ziran2020_amd import scenes
    if name == "batched_scenes":
        full_counts = np.array(scene.info["counts"])
        order = np.argsort(full_counts, kind="stable")
        keep, tot = [], 0
        for s in order:
            if tot + full_counts[s] > max_particles and keep:
                break
            keep.append(int(s))
            tot += int(full_counts[s])
        sc = scenes.batched_scenes(keep=keep)
        return sc, f"{len(keep)} of the 64 scenes ({sc.n} particles, each scene as in the full batch)"
    if scene.n > max_particles and name in ORACLE_SAMPLES:
        mk, what, _ = ORACLE_SAMPLES[name]
        sc = mk(scenes)
        if sc.n < scene.n:
            return sc, what
    return scene, f"the full {scene.n}-particle workload"


def time_oracle(scene, threads, budget_s, max_steps):
    from ziran2020_amd import _abi as A
    lib = A.ZmpmLib(ORACLE_LIB)
    lib.lib.zmpm_oracle_set_threads(threads)
    sim = scene.make(lib, A.PREC_FP64)
    t0 = time.perf_counter()
    scene.run(sim, 1)
    one = time.perf_counter() - t0
    steps = int(max(1, min(max_steps, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    scene.run(sim, steps, start=1)
    el = time.perf_counter() - t0
    sim.close()
    return scene.n * steps / el, steps


def cpu_baseline(name, scene, budget_s=10.0):
    """The FP64 oracle (C++/OpenMP restatement of the reference CPU step) on the
    host cores, at the available threads and at 1 thread, on bounded samples."""
    threads = cpu_threads()
    sc, what = cpu_sample_scene(name, scene, 1_000_000)
    v, steps = time_oracle(sc, threads, budget_s, 10)
    sc1, what1 = cpu_sample_scene(name, scene, 150_000)
    v1, steps1 = time_oracle(sc1, 1, budget_s, 5)
    return {"value": v, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} substeps of {what} after 1 untimed substep, FP64 oracle "
                      f"(C++/OpenMP restatement of the reference CPU step), {threads} threads",
            "value_1thread": v1,
            "sample_1thread": f"{steps1} substeps of {what1} after 1 untimed substep, 1 thread",
            "cpu_model": cpu_model()}


def timed(scene, lib, precision, steps, warmup, timing_steps=50, autotune=True):
    """Launch configuration from the autotuner (state left unchanged), then
    throughput of `steps` substeps (HIP-graph replays, no events inside), then
    per-stage kernel times from HIP events on the handle's stream over the next
    min(steps, timing_steps) substeps launched one kernel at a time."""
    sim = scene.make(lib, precision)
    timed.launch_config = sim.autotune(scene.dt, 0.0, steps=4) if autotune else sim.launch_config()
    scene.run(sim, warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    scene.run(sim, steps, start=warmup)
    sim.synchronize()
    el = time.perf_counter() - t0
    k = min(steps, timing_steps)
    sim.L.zmpm_set_timing(sim.h, 1)
    scene.run(sim, k, start=warmup + steps)
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


def _field_errors(a, b, keys):
    errs = {}
    for k in keys:
        den = np.abs(b[k]).max()
        errs[k] = float(np.abs(a[k] - b[k]).max() / den) if den > 0 else float(np.abs(a[k]).max())
    return errs


def _grid_errors(ga, gb):
    """Normwise error of grid m and mv over the union of the active nodes
    (matched by node coordinates; a node active in one only counts as 0 in the other)."""
    def keyed(g):
        c = g["coords"].astype(np.int64)
        return (c[:, 0] << 40) | (c[:, 1] << 20) | c[:, 2]
    ka, kb = keyed(ga), keyed(gb)
    keys = np.union1d(ka, kb)

    def spread(k, g):
        idx = np.searchsorted(keys, k)
        m, mv = np.zeros(keys.size), np.zeros((keys.size, 3))
        m[idx] = g["m"]
        mv[idx] = g["m"][:, None] * g["v"]
        return m, mv
    ma, pa = spread(ka, ga)
    mb, pb = spread(kb, gb)
    return {"m": float(np.abs(ma - mb).max() / np.abs(mb).max()),
            "mv": float(np.abs(pa - pb).max() / np.abs(pb).max()),
            "nodes_only_in_one": int(keys.size - np.intersect1d(ka, kb).size)}


def accuracy_vs_fp64(scene, lib, steps=100):
    """Per-field normwise error (max |a - b| / max |b|) of the FP32 form against
    the FP64 plain GPU form after exactly `steps` substeps of the full workload."""
    from ziran2020_amd import _abi as A
    out = {}
    for prec in (A.PREC_FP32, A.PREC_FP64):
        sim = scene.make(lib, prec)
        scene.run(sim, steps)
        out[prec] = sim.download()
        sim.close()
    return _field_errors(out[A.PREC_FP32], out[A.PREC_FP64], ("x", "v", "C", "F", "d", "g"))


ORACLE_SAMPLES = {
    # workload -> (sample scene, description, grid blocks the library reserves
    # or None for its default); the oracle runs 100 substeps of it.  The sparse
    # samples keep the workload's geometry with fewer particles, under one per
    # cell: nearly every particle opens its own grid block, more than the
    # library's default reserve for that particle count.
    "elastic_cube": (lambda S: S.elastic_cube(), "the full workload", None),
    "damage_slab": (lambda S: S.damage_slab(n=125_000), "the slab at 125,000 particles", None),
    "damage_slab_active": (lambda S: S.damage_slab_active(n=125_000), "the slab at 125,000 particles", None),
    "damage_slab_implicit": (lambda S: S.damage_slab_implicit(n=125_000), "the slab at 125,000 particles", None),
    "sparse_spheres": (lambda S: S.sparse_spheres(n=250_000), "the 20 spheres at 250,000 particles", 250_000),
    "two_material": (lambda S: S.two_material(n=250_000), "the block and column at 250,000 particles", 250_000),
}


def accuracy_vs_oracle(name, scene, lib, steps=100):
    """Per-field normwise error of the FP32 form (GPU) against the FP64 oracle
    after exactly `steps` substeps from the same upload, on a bounded sample of
    the workload (the oracle is the restated reference CPU step), including the
    grid mass and momentum of the last substep."""
    from ziran2020_amd import _abi as A, scenes
    blocks = None
    if name == "batched_scenes":
        sc, what = cpu_sample_scene(name, scene, 100_000)
    else:
        mk, what, blocks = ORACLE_SAMPLES[name]
        sc = mk(scenes)
    ref = A.ZmpmLib(ORACLE_LIB)
    ref.lib.zmpm_oracle_set_threads(cpu_threads())
    outs, grids = [], []
    for L, prec in ((lib, A.PREC_FP32), (ref, A.PREC_FP64)):
        # the library reads its block reserve from the environment at upload
        old = os.environ.get("ZMPM_BLOCK_CAPACITY")
        if blocks and L is lib:
            os.environ["ZMPM_BLOCK_CAPACITY"] = str(blocks)
        try:
            sim = sc.make(L, prec)
        finally:
            if blocks and L is lib:
                if old is None:
                    del os.environ["ZMPM_BLOCK_CAPACITY"]
                else:
                    os.environ["ZMPM_BLOCK_CAPACITY"] = old
        sc.run(sim, steps)
        outs.append(sim.download())
        grids.append(sim.grid())
        sim.close()
    errs = _field_errors(outs[0], outs[1], ("x", "v", "C", "F", "d", "g"))
    errs["grid"] = _grid_errors(grids[0], grids[1])
    return {"substeps": steps, "sample": what, "particles": sc.n, **errs}


def profile_entry(path, key, digest):
    """(value, note) of a per-dispatch counter figure committed under profiles/,
    None when missing or measured on other kernel sources."""
    if not os.path.exists(path):
        return None, "no profile"
    with open(path) as f:
        table = json.load(f)
    e = table.get(key)
    if not isinstance(e, dict):
        return None, "no profile for this workload"
    if e.get("source") != digest:
        return None, f"stale: measured on kernel sources {e.get('source')}, now {digest}"
    return e["value"], f"measured at {e.get('commit', '?')}"


def progress(msg):
    """One line on stderr as each leg of the run ends (the JSON line is stdout's)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="batched_scenes")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 100-substep accuracy runs")
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
    make = scenes.WORKLOADS[args.workload]
    scene = make()
    damage = any(m.has_damage for m in scene.materials)
    digest = source_digest()

    el32, st32, nn32, out32 = timed(scene, lib, A.PREC_FP32, args.steps, args.warmup, autotune=not args.no_autotune)
    cfg32 = timed.launch_config
    value = scene.n * args.steps / el32
    progress(f"FP32 form timed ({args.workload}, {scene.n} particles)")
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
        "config": {"workload": args.workload, "particles": scene.n,
                   "form": "FP32 particle state, int64 fixed-point P2G", "active_nodes": nn32,
                   "launch": cfg32, "autotuned": not args.no_autotune},
    }
    if isinstance(scene, scenes.SceneBatch):
        result["config"].update(scenes=len(scene.scenes), dx=sorted({s.dx for s in scene.scenes}),
                                dt=[float(scene.dts.min()), float(scene.dts.max())],
                                launches="one per stage for all scenes (zmpm_run_batch)")
    else:
        result["config"].update(dx=scene.dx, dt=scene.dt, options=dict(scene.options))
    # roofline of the dominant kernel stage (HIP events on the handle's stream)
    stage_ms = {k: v for k, v in st32.items() if k in ("p2g", "g2p", "grid_update")}
    dom = max(stage_ms, key=stage_ms.get)
    byt = algorithmic_bytes(dom, "fp32", scene.n, nn32, damage)
    achieved = byt / (stage_ms[dom] * 1e-3) / 1e9
    key = f"{args.workload}/fp32/{dom}"
    traffic, traffic_note = profile_entry(os.path.join(ROOT, "profiles", "traffic.json"), key, digest)
    valu, valu_note = profile_entry(os.path.join(ROOT, "profiles", "valu.json"), key, digest)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note, "kernel": dom,
            "algorithmic_bytes": byt, "avg_ms": stage_ms[dom], "kernel_sources": digest}
    if valu is not None:
        # SQ_INSTS_VALU (wavefront instructions per dispatch) x 64 lanes over the
        # kernel's time, against the VALU issue peak
        lane_ops = valu * 64 / (stage_ms[dom] * 1e-3)
        roof["valu"] = {"lane_ops_per_s": lane_ops, "peak": VALU_PEAK_LANE_OPS,
                        "frac": lane_ops / VALU_PEAK_LANE_OPS, "insts_per_particle": valu * 64 / scene.n,
                        "note": valu_note}
        roof["limiter"] = "valu" if lane_ops / VALU_PEAK_LANE_OPS > achieved / HBM_PEAK_GBS else "hbm"
    else:
        roof["valu"] = {"note": valu_note}
    result["roofline"] = roof
    result["stage_ms"] = st32
    if not args.no_fp64:
        el64, st64, nn64, out64 = timed(scene, lib, A.PREC_FP64, args.steps, args.warmup,
                                        autotune=not args.no_autotune)
        v64 = scene.n * args.steps / el64
        progress("FP64 form timed")
        result["fp64_plain"] = {"value": v64, "ms_per_step": 1e3 * el64 / args.steps, "stage_ms": st64,
                                "dtype": "fp64", "launch": timed.launch_config}
        result["fp32_over_fp64"] = value / v64
        if not args.no_accuracy:
            result["fp32_vs_fp64_after_100_substeps"] = accuracy_vs_fp64(scene, lib, 100)
            progress("FP32 against FP64 after 100 substeps done")
    if not args.no_accuracy:
        result["fp32_vs_oracle_after_100_substeps"] = accuracy_vs_oracle(args.workload, scene, lib, 100)
        progress("FP32 against the oracle after 100 substeps done")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args.workload, scene)
        progress("CPU baseline done")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
