"""The C-ABI library loads without a GPU, exports every symbol of
include/gorender_hip.h, and its struct layouts match the ctypes binding."""
import ctypes
import os
import re
import subprocess

import pytest

from gorender_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gorender_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    fns = header_functions()
    assert "gr_draw" in fns and "gr_create" in fns and len(fns) >= 25


def test_library_exports_every_header_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (gr_[a-z0-9_]+)$", out, flags=re.M))
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == header_functions()
    lib = _lib.load()
    for name in header_functions():
        assert getattr(lib, name) is not None


def test_struct_layout_matches_c(tmp_path):
    """Compile the header as C99 and compare sizeof/offsetof with ctypes."""
    src = tmp_path / "sizes.c"
    structs = ["gr_texture_desc", "gr_mesh_desc", "gr_object", "gr_frame", "gr_stats",
               "gr_batch_frame", "gr_config", "gr_server_config", "gr_serve_result",
               "gr_tune_result"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gorender_hip.h"',
             "int main(void) {"]
    for s in structs:
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(_lib, s)._fields_:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT}/include", str(src), "-o",
                    str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for s in structs:
        cls = getattr(_lib, s)
        assert int(got[s]) == ctypes.sizeof(cls), s
        for f, _ in cls._fields_:
            assert int(got[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def test_abi_version_and_defaults():
    lib = _lib.load()
    assert lib.gr_abi_version() == 5
    f = _lib.gr_frame()
    lib.gr_frame_defaults(ctypes.byref(f), 800, 600)
    # NewRenderer / Draw defaults (renderer.go:12-17,118-136,448-449)
    assert f.width == 800 and f.height == 600 and f.ref_tiles == 16
    assert f.flags == _lib.GR_DEFAULT_FLAGS
    assert (f.ambient, f.diffuse, f.znear, f.zfar, f.dot_step) == (0.5, 0.5, 0.0, 50.0, 10)
    assert list(f.face_color) == [200, 200, 200, 255]
    assert list(f.clear_color) == [50, 50, 50, 255]
    assert list(f.dot_color) == [100, 100, 100, 255]
    assert list(f.screen) == [400, 0, 0, 400, 0, 300, 0, 300, 0, 0, 0.5, 0.5, 0, 0, 0, 1]


def test_texture_kind_for_size():
    lib = _lib.load()
    # isPowerOfTwo (utils.go:7-9) picks the fast path (texture.go:40-42)
    assert lib.gr_texture_kind_for_size(16, 16) == _lib.GR_TEXTURE_IMAGE_FAST
    assert lib.gr_texture_kind_for_size(1, 64) == _lib.GR_TEXTURE_IMAGE_FAST
    assert lib.gr_texture_kind_for_size(16, 12) == _lib.GR_TEXTURE_IMAGE
    assert lib.gr_texture_kind_for_size(0, 16) == _lib.GR_TEXTURE_IMAGE


def test_create_without_gpu_fails_cleanly():
    """Without a device gr_create reports an error instead of crashing."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.gr_create(0, ctypes.byref(h))
    if st == _lib.GR_OK:
        lib.gr_destroy(h)
        pytest.skip("a GPU is present")
    assert not h.value and lib.gr_last_error()
