"""The C ABI boundary: both shared libraries load and export every function that
include/zmpm.h declares (no device work is done here, so this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zmpm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zmpm_[a-z_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared()
    for must in ("zmpm_create", "zmpm_upload", "zmpm_substep", "zmpm_run", "zmpm_download",
                 "zmpm_get_binning", "zmpm_get_grid", "zmpm_destroy"):
        assert must in names


def test_hip_library_exports_every_symbol():
    import ziran2020_amd
    lib = ctypes.CDLL(ziran2020_amd.LIBRARY_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.zmpm_abi_version.restype = ctypes.c_int32
    version = int(re.search(r"#define ZMPM_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    from ziran2020_amd import _abi as A
    assert lib.zmpm_abi_version() == version == A.ZMPM_ABI_VERSION
    lib.zmpm_build_info.restype = ctypes.c_char_p
    assert b"gfx950" in lib.zmpm_build_info()


def test_oracle_exports_every_symbol(oracle):
    missing = [n for n in declared() if not hasattr(oracle.lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    from ziran2020_amd._abi import ZmpmLib
    assert set(ZmpmLib.EXPORTS) == set(declared())


def test_struct_layouts_match_header():
    """ctypes mirrors of the ABI structs have the size and every field offset a C
    compiler gives them (a swap of two same-sized fields fails this)."""
    import subprocess
    import tempfile
    from ziran2020_amd import _abi as A
    structs = [("zmpm_scene_desc", A.SceneDesc), ("zmpm_material", A.Material), ("zmpm_collider", A.Collider),
               ("zmpm_particles", A.Particles), ("zmpm_launch_config", A.LaunchConfig)]
    c_name = {"lam": "lambda"}  # a Python keyword in the ctypes mirror
    lines = ['#include "zmpm.h"', "#include <stddef.h>", "#include <stdio.h>", "int main(){"]
    want = []
    for cname, py in structs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(ctypes.sizeof(py))
        for field in py._fields_:
            f = field[0]
            lines.append(f'printf("%zu\\n", offsetof({cname}, {c_name.get(f, f)}));')
            want.append(getattr(py, f).offset)
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write("\n".join(lines) + "\n")
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    assert got == want


def test_library_loads_without_gpu_and_fails_loudly():
    """create() needs a device; on a host without one it reports ZMPM_E_HIP instead of falling back."""
    import ziran2020_amd
    from ziran2020_amd import _abi as A
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = os.path.exists("/dev/kfd")
    if has_gpu:
        pytest.skip("a GPU is present")
    lib = ziran2020_amd.load_library()
    with pytest.raises(A.ZmpmError) as e:
        A.Simulation(lib, dx=0.01)
    assert e.value.status == A.E_HIP


def test_oracle_autotune_cached_keeps_its_configuration(oracle, tmp_path):
    # the oracle has no launches: the cached tuner returns its configuration and
    # never reports a cache hit (nor writes the file)
    from ziran2020_amd import scenes
    sc = scenes.elastic_cube(n_side_cells=2)
    sim = sc.make(oracle)
    cfg, hit = sim.autotune_cached(str(tmp_path / "t.txt"), sc.dt)
    assert not hit and cfg == sim.launch_config()
    assert not (tmp_path / "t.txt").exists()
