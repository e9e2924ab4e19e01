"""The C-ABI library (CPU only): it loads, exports every entry point include/swr.h declares,
and its host-only paths behave without a GPU."""
import ctypes as C
import os
import re

import pytest

import swr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "swr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swr_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported():
    lib = C.CDLL(swr.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them with a signature
    assert sorted(swr.exported_symbols()) == names


def test_struct_layouts_match_the_header(tmp_path):
    """Every field offset and record size of the ctypes mirror equals what a C compiler gives
    the header's structs."""
    records = {"swr_camera_desc": swr.CameraDesc, "swr_render_desc": swr.RenderDesc,
               "swr_shader_desc": swr.ShaderDesc, "swr_quad_light_desc": swr.QuadLightDesc,
               "swr_stats": swr.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "swr.h"', "int main(void) {"]
    for cname, cls in records.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n"):
        if line:
            rec, field, val = line.split()
            got[(rec, field)] = int(val)
    for cname, cls in records.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])


def test_null_arguments_are_rejected_without_a_device():
    lib = swr.load_library()
    assert lib.swr_create(0, None) == swr.SWR_E_INVALID
    assert lib.swr_set_budget(None, 0) == swr.SWR_E_INVALID
    assert lib.swr_cancel(None) == swr.SWR_E_INVALID
    assert lib.swr_render(None, None, None, None, None, None) == swr.SWR_E_INVALID
    assert lib.swr_scene_load_yaml(None, b"x", None, None) == swr.SWR_E_INVALID
    assert lib.swr_last_error(None) == b"null context"


def test_create_reports_missing_device():
    import torch  # noqa: F401  (plumbing only: tells whether this host has a HIP device)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert swr.load_library().swr_create(0, C.byref(h)) == swr.SWR_E_HIP
