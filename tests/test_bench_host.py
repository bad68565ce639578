"""The host side of bench.py, run against the oracle library on the CPU: the
accuracy and CPU-baseline legs must work on every workload's sample (a batch
included), since the bench runs them after its timed GPU legs."""
import numpy as np

import bench
from ziran2020_amd import _abi as A, scenes
from helpers import run


def test_oracle_batch_grid_is_the_scenes_grids_at_their_origins(oracle):
    b = scenes.batched_scenes(n_scenes=3, n_range=(300, 800))
    sim = run(oracle, b, A.PREC_FP64, 2)
    g = sim.grid()
    parts = []
    for i, s in enumerate(b.scenes):
        gs = run(oracle, s, A.PREC_FP64, 2).grid()
        gs["coords"] = gs["coords"] + sim.origin(i)
        parts.append(gs)
    coords = np.concatenate([p["coords"] for p in parts])
    m = np.concatenate([p["m"] for p in parts])
    order = np.lexsort((coords[:, 2], coords[:, 1], coords[:, 0]))
    assert np.array_equal(g["coords"], coords[order]) and np.array_equal(g["m"], m[order])


def test_accuracy_leg_on_a_batch(oracle, monkeypatch):
    # the bench's FP32-vs-oracle leg on the batched workload's sample, with the
    # oracle standing in for the GPU library (same calls, both FP64 here)
    monkeypatch.setattr(bench, "cpu_sample_scene",
                        lambda name, scene, n: (scenes.batched_scenes(n_scenes=2, n_range=(200, 400)), "two scenes"))
    out = bench.accuracy_vs_oracle("batched_scenes", None, oracle, steps=2)
    # the "GPU" side runs the FP32 layout (fixed-point positions): small, not zero
    assert out["particles"] > 0 and out["grid"]["m"] < 1e-4 and out["x"] < 1e-4


def test_grid_errors_helper():
    rng = np.random.default_rng(0)
    c = np.unique(rng.integers(0, 50, (200, 3)), axis=0).astype(np.int32)
    ga = dict(coords=c, m=rng.random(len(c)) + 1, v=rng.random((len(c), 3)))
    gb = dict(coords=c[3:], m=ga["m"][3:], v=ga["v"][3:])
    e = bench._grid_errors(ga, gb)
    assert e["nodes_only_in_one"] == 3 and e["m"] > 0
    assert bench._grid_errors(ga, ga)["m"] == 0.0


def test_source_digest_ignores_comments_only():
    # profile figures are tied to the kernel code: comments and blank lines do
    # not change the digest, code (string literals included) does
    a = 'int f(int x) { return x + 1; } // one\nconst char* s = "a // b";\n'
    b = '/* doc */\nint f(int x) { return x + 1; }\n\n// two\nconst char* s = "a // b";   \n'
    c = 'int f(int x) { return x + 2; }\nconst char* s = "a // b";\n'
    d = 'int f(int x) { return x + 1; }\nconst char* s = "a // c";\n'
    assert bench.code_only(a) == bench.code_only(b)
    assert bench.code_only(a) != bench.code_only(c)
    assert bench.code_only(a) != bench.code_only(d)
    assert len(bench.source_digest()) == 16
