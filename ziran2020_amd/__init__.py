"""ziran2020_amd: MI355X-native explicit APIC MPM substep (+ AnisoMPM damage).

The compute path is ``libzmpm.so`` (hand-written gfx950 HIP kernels behind the
C ABI in include/zmpm.h).  :func:`load_library` loads it and raises if it is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import os

from . import _abi
from ._abi import ZmpmError, ZmpmLib, Simulation  # noqa: F401

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, "libzmpm.so")

_lib = None


def load_library() -> ZmpmLib:
    """Load the HIP library built in-tree by ``__graft_entry__.build()``."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`; "
                "there is no CPU fallback for the MPM substep")
        _lib = ZmpmLib(LIBRARY_PATH)
    return _lib
