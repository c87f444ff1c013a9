"""CPU checks of the C ABI and the host-side logic (no kernel launches):
the library loads and exports every symbol include/ubs_capi.h declares,
argument validation fails with an error code and message before any HIP
call, and the Python layer raises the reference's errors."""
import ctypes
import os
import re
import subprocess
import sys

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ubs_capi.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ubs_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from gsplat.cuda._backend import lib
    return lib()


def test_library_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 20, names
    for n in names:
        assert hasattr(L._dll, n), f"{n} not exported by {L.path}"
    assert L.ubs_version() >= 100


def test_every_declared_symbol_has_a_ctypes_signature():
    from gsplat.cuda._backend import _SIGNATURES
    missing = [n for n in declared_symbols() if n not in _SIGNATURES]
    assert not missing, missing


def test_ctypes_signatures_match_header_arity():
    """Every ctypes binding passes exactly the parameters the C declaration has."""
    from gsplat.cuda._backend import _SIGNATURES
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, params in re.findall(r"\b(ubs_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert len(_SIGNATURES[name][1]) == n, (name, len(_SIGNATURES[name][1]), n)


def test_argument_validation_without_gpu(L):
    sel = ctypes.c_int(0)
    rc = L.ubs_radix_sort_pairs(10, None, None, None, None, 10, 5, None, 0, ctypes.byref(sel), None)
    assert rc == 1 and "bit range" in L.last_error()
    rc = L.ubs_radix_sort_pairs(1 << 31, None, None, None, None, 0, 32, None, 0, ctypes.byref(sel), None)
    assert rc == 1 and "out of range" in L.last_error()
    rc = L.ubs_radix_sort_pairs(100, None, None, None, None, 0, 32, None, 0, ctypes.byref(sel), None)
    assert rc == 3 and "workspace" in L.last_error()
    assert L.ubs_radix_sort_workspace_bytes(10 ** 6, 0, 46) > L.ubs_radix_sort_workspace_bytes(10 ** 5, 0, 46) > 6 * 256 * 4
    rc = L.ubs_rot_scale_l_triangle_to_covar_fwd(10, 13, None, None, None, None, None, 0, 0, None, None)
    assert rc == 1 and "D must be in" in L.last_error()
    rc = L.ubs_cond_mean_convariance_opacity_fwd(10, 3, *([None] * 8), None)
    assert rc == 1
    rc = L.ubs_rasterize_to_pixels_fwd(1, 10, 0, 3, *([None] * 7), 64, 64, 12, 4, 4, *([None] * 6), None, 0, None)
    assert rc == 1 and "tile_size" in L.last_error()
    rc = L.ubs_rasterize_to_pixels_fwd(1, 10, 0, 3, *([None] * 7), 64, 64, 16, 2, 4, *([None] * 6), None, 0, None)
    assert rc == 1 and "cover" in L.last_error()
    assert L.ubs_rasterize_workspace_bytes(2, 1000, 1) == 2 * L.ubs_rasterize_workspace_bytes(2, 1000, 0) >= 64000
    # room for the tile order after the records: one uint32 per tile of a 120 x 68 grid, 2 cameras (single-tile
    # blocks by default; a build with larger blocks needs less)
    for bwd in (0, 1):
        base = L.ubs_rasterize_workspace_bytes(2, 1000, bwd)
        extra = L.ubs_rasterize_workspace_bytes_tiles(2, 1000, 120, 68, bwd) - base
        assert extra % 256 == 0 and 2 * 60 * 34 * 4 <= extra <= ((2 * 120 * 68 * 4 + 255) & ~255)
    rc = L.ubs_fully_fused_projection_fwd(1, 10, None, None, 7, *([None] * 4), 64, 64, 0.3, 0.01, 1e10, 0.0, 0,
                                          *([None] * 5), None)
    assert rc == 1


def test_noalloc_sort_rejects_before_any_hip_call(L):
    """The allocation-free sort checks its bit range and whether the sort's
    scratch fits its fixed device array before it touches the device."""
    sel = ctypes.c_int(0)
    rc = L.ubs_radix_sort_pairs_noalloc(10, None, None, None, None, 10, 5, ctypes.byref(sel), None, None)
    assert rc == 1 and "bit range" in L.last_error()
    rc = L.ubs_radix_sort_pairs_noalloc(1 << 25, None, None, None, None, 0, 64, ctypes.byref(sel), None, None)
    assert rc == 1 and "allocation-free" in L.last_error(), L.last_error()
    assert L.ubs_radix_sort_workspace_bytes(1 << 24, 0, 64) <= 24 << 20 < L.ubs_radix_sort_workspace_bytes(
        1 << 25, 0, 64)


def test_prefix_bits_match_reference_formula(L):
    import math
    for C, tw, th in ((1, 120, 68), (1, 16, 16), (8, 80, 45), (3, 1, 1), (1, 64, 64)):
        want = (math.floor(math.log2(tw * th)) + 1) + (math.floor(math.log2(C)) + 1)
        assert L.ubs_isect_prefix_bits(C, tw, th) == want


def test_python_api_errors_match_reference():
    from gsplat.cuda import _wrapper as W
    C, N = 1, 5
    args = [torch.zeros(C, N, 2), torch.zeros(C, N, 3), torch.zeros(C, N, 514), torch.zeros(C, N), torch.zeros(C, N),
            16, 16, 16, torch.zeros(C, 1, 1, dtype=torch.int32), torch.zeros(0, dtype=torch.int32)]
    with pytest.raises(ValueError):
        W.rasterize_to_pixels(*args)
    with pytest.raises(AssertionError):
        W.rasterize_to_pixels(torch.zeros(C, N, 3), *args[1:])
    with pytest.raises(RuntimeError, match="HIP/CUDA device"):
        W.rasterize_to_pixels(*args[:2], torch.zeros(C, N, 3), *args[3:])
    with pytest.raises(AssertionError):
        W.fully_fused_projection(torch.zeros(N, 3), None, None, None, torch.zeros(1, 4, 4), torch.zeros(1, 3, 3), 8, 8)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_c_shim_loads_under_another_package_name():
    """The `_C` stand-in loads the way a maintainer binds it into the
    reference's _backend.py (a file-location spec under a foreign package
    name) and provides every extension function the rasteriser path calls."""
    import importlib.util

    cuda_dir = os.path.join(ROOT, "universal-beta-splatting_amd", "gsplat", "cuda")
    spec = importlib.util.spec_from_file_location("ubs_shim_probe_pkg", os.path.join(cuda_dir, "__init__.py"),
                                                  submodule_search_locations=[cuda_dir])
    pkg = importlib.util.module_from_spec(spec)
    sys.modules["ubs_shim_probe_pkg"] = pkg
    try:
        spec.loader.exec_module(pkg)
        shim = importlib.import_module("ubs_shim_probe_pkg._c_shim")
        for name in ("cond_mean_convariance_opacity_fwd", "cond_mean_convariance_opacity_bwd",
                     "rot_scale_l_triangle_to_covar_fwd", "rot_scale_l_triangle_to_covar_bwd",
                     "l_triangle_to_rotmat_fwd", "l_triangle_to_rotmat_bwd", "fully_fused_projection_fwd",
                     "fully_fused_projection_bwd", "isect_tiles", "isect_offset_encode",
                     "rasterize_to_pixels_fwd", "rasterize_to_pixels_bwd"):
            assert callable(getattr(shim, name)), name
        with pytest.raises(RuntimeError, match="HIP/CUDA device"):
            shim.isect_offset_encode(torch.zeros(3, dtype=torch.int64), 1, 2, 2)
    finally:
        for k in [k for k in sys.modules if k.startswith("ubs_shim_probe_pkg")]:
            del sys.modules[k]


def test_constant_depth_passes_host_logic():
    """Which depth-byte sort passes isect_tiles skips, from the packed
    (~min, max) pair: byte p is constant iff min and max agree on bits >= 8p."""
    import struct

    from gsplat.cuda._wrapper import constant_depth_passes

    def pack(lo_f, hi_f):
        lo = struct.unpack("<I", struct.pack("<f", lo_f))[0]
        hi = struct.unpack("<I", struct.pack("<f", hi_f))[0]
        v = ((hi << 32) | (~lo & 0xFFFFFFFF))
        return v - (1 << 64) if v >= (1 << 63) else v  # as the int64 the device writes

    assert constant_depth_passes(pack(2.5, 5.5)) == 0b1000  # 0x40200000 .. 0x40B00000
    assert constant_depth_passes(pack(3.0, 3.0)) == 0b1111
    assert constant_depth_passes(pack(0.5, 5.5)) == 0b0000  # 0x3F000000 .. 0x40B00000
    assert constant_depth_passes(pack(3.0, 3.0000002)) == 0b1110  # differ in the lowest byte only
    assert constant_depth_passes(0) == 0  # nothing visible: ~min = 0, max = 0 -> min > max
