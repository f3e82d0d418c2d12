"""The C-ABI library loads and exports every entry point include/swr.h declares.

No compute calls: these run without a GPU.
"""
import ctypes
import os
import re

import pytest

from swiftshader_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swr.h")
LIB = os.path.join(ROOT, "swiftshader_amd", "libswr.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(swr_\w+)\s*\(", text, flags=re.M)


def test_header_declares_the_mirrored_entry_points():
    names = declared_functions()
    assert names, "no swr_* functions found in include/swr.h"
    assert sorted(names) == sorted(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libswr.so not built")
    lib = ctypes.CDLL(LIB)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes_match_the_c_layout():
    # Offsets the C compiler gives the descriptor (x86-64 SysV, natural alignment).
    assert ctypes.sizeof(abi.Viewport) == 24
    assert ctypes.sizeof(abi.StencilOpState) == 28
    assert ctypes.sizeof(abi.ImageLevel) == 24
    assert ctypes.sizeof(abi.Texture) == 8 + 15 * 24
    assert ctypes.sizeof(abi.Attachment) == 24
    assert (abi.DrawDesc.alpha_to_coverage_enable.offset + 4 + 7) // 8 * 8 == ctypes.sizeof(abi.DrawDesc)  # tail padding to 8


def test_struct_layout_against_the_header(tmp_path):
    """Compile a probe that prints offsetof() for the descriptor and compare."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    structs = [("swr_draw_desc", abi.DrawDesc), ("swr_grad_desc", abi.GradDesc), ("swr_texture", abi.Texture),
               ("swr_resolve_desc", abi.ResolveDesc), ("swr_profile", abi.Profile), ("swr_clear_desc", abi.ClearDesc),
               ("swr_bc_desc", abi.BCDesc), ("swr_etc_desc", abi.ETCDesc)]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    want = []
    for cname, py in structs:
        for f in py._fields_:
            src.append(f'printf("%zu\\n", offsetof({cname}, {f[0]}));')
            want.append(getattr(py, f[0]).offset)
        src.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(ctypes.sizeof(py))
    src.append("return 0; }")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run([cc, str(c), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out == want


def test_version_and_error_paths_without_gpu():
    if not os.path.exists(LIB):
        pytest.skip("libswr.so not built")
    lib = abi.bind(ctypes.CDLL(LIB))
    assert lib.swr_abi_version() == abi.ABI_VERSION
    # A NULL descriptor is reported, never crashes (device or argument error).
    assert lib.swr_draw(None, None) < 0
    assert isinstance(lib.swr_last_error(), bytes)
    assert lib.swr_set_tuning(None) < 0
    assert lib.swr_resolve(None, None) < 0
    assert lib.swr_clear(None, None) < 0
    assert lib.swr_decode_bc(None, None) < 0
    assert lib.swr_batch_plan_create(None, 0, 0, None, None) < 0
    assert lib.swr_batch_plan_run(None, None) < 0
    assert lib.swr_batch_plan_destroy(None) == 0
    assert lib.swr_generate_mips(None, None) < 0
    assert lib.swr_mip_backward(None, None) < 0
