"""Shared comparison helpers for the parity tests."""
import numpy as np


def normwise(a, b):
    """max |a-b| / max |b| over a field (relative to the field's magnitude)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    return num / den if den > 0 else num


def run(lib, scene, precision, steps, t0=0.0, **kw):
    sim = scene.make(lib, precision, **kw)
    sim.run(steps, scene.dt, t0)
    return sim
