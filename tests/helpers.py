"""Shared comparison helpers for the parity tests."""
import numpy as np


def normwise(a, b):
    """max |a-b| / max |b| over a field (relative to the field's magnitude)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    return num / den if den > 0 else num


def grid_errors(ga, gb):
    """Normwise error of grid mass and momentum (m v) of grid ga against gb over
    the union of their active nodes, matched by node coordinates (a node active
    in one only counts as 0 in the other)."""
    def keyed(g):
        c = g["coords"].astype(np.int64)
        return (c[:, 0] << 40) | (c[:, 1] << 20) | c[:, 2]
    ka, kb = keyed(ga), keyed(gb)
    keys = np.union1d(ka, kb)
    def spread(k, g):
        idx = np.searchsorted(keys, k)
        m = np.zeros(keys.size)
        mv = np.zeros((keys.size, 3))
        m[idx] = g["m"]
        mv[idx] = g["m"][:, None] * g["v"]
        return m, mv
    ma, pa = spread(ka, ga)
    mb, pb = spread(kb, gb)
    return {"m": float(np.abs(ma - mb).max() / np.abs(mb).max()),
            "mv": float(np.abs(pa - pb).max() / np.abs(pb).max()),
            "nodes_only_in_one": int(keys.size - np.intersect1d(ka, kb).size)}


def run(lib, scene, precision, steps, start=0, **kw):
    """A handle of the scene (or scene batch) after substeps start .. start + steps - 1."""
    sim = scene.make(lib, precision, **kw)
    scene.run(sim, steps, start)
    return sim


def read_state(path):
    """A state record of MpmSimulationHip::writeState (zmpm_simulation.hpp): magic,
    version, n, frame, time, time_since_last_frame, then the attribute arrays."""
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:8] == b"ZMPMSTAT"
    version, n = np.frombuffer(buf, np.int32, 1, 8)[0], int(np.frombuffer(buf, np.int64, 1, 12)[0])
    assert version == 2
    frame = int(np.frombuffer(buf, np.int32, 1, 20)[0])
    time, since = np.frombuffer(buf, np.float64, 2, 24)
    off = 40
    out = dict(frame=frame, time=float(time), time_since_last_frame=float(since))
    for name, width, dt in (("x", 3, np.float64), ("v", 3, np.float64), ("C", 9, np.float64), ("F", 9, np.float64),
                            ("m", 1, np.float64), ("V0", 1, np.float64), ("material", 1, np.int32),
                            ("d", 1, np.float64), ("g", 1, np.float64), ("laplacian", 1, np.float64),
                            ("plastic_state", 1, np.float64), ("max_dtilde", 1, np.float64), ("J", 1, np.float64)):
        cnt = n * width
        a = np.frombuffer(buf, dt, cnt, off)
        off += cnt * np.dtype(dt).itemsize
        out[name] = a.reshape(n, width) if width > 1 else a
    assert off == len(buf)
    return out
