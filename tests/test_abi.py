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
    assert lib.zmpm_abi_version() == 6
    lib.zmpm_build_info.restype = ctypes.c_char_p
    assert b"gfx950" in lib.zmpm_build_info()


def test_oracle_exports_every_symbol(oracle):
    missing = [n for n in declared() if not hasattr(oracle.lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    from ziran2020_amd._abi import ZmpmLib
    assert set(ZmpmLib.EXPORTS) == set(declared())


def test_struct_layouts_match_header():
    """ctypes mirrors of the ABI structs have the sizes a C compiler gives them."""
    import subprocess
    import tempfile
    from ziran2020_amd import _abi as A
    src = ('#include "zmpm.h"\n#include <stdio.h>\nint main(){printf("%zu %zu %zu %zu %zu\\n", sizeof(zmpm_scene_desc),'
           ' sizeof(zmpm_material), sizeof(zmpm_collider), sizeof(zmpm_particles), sizeof(zmpm_launch_config));'
           ' return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    assert sizes == [ctypes.sizeof(A.SceneDesc), ctypes.sizeof(A.Material), ctypes.sizeof(A.Collider),
                     ctypes.sizeof(A.Particles), ctypes.sizeof(A.LaunchConfig)]


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
