"""The drop-in boundary: the C-ABI library loads and exports every entry point the headers
declare, the Python mirror agrees with it on struct layout, and the C++ API of
include/exposurerender.h (Source/exposurerender.h:32-42) is exported and converts statuses
into ExposureRender::Exception. The GPU half renders one scene through the C++ API and the
C-ABI and requires identical images."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(ROOT, "exposure-render_amd", "lib")
CHECK = os.path.join(LIB_DIR, "er_facade_check")

CPP_API = ["BindTracer", "BindVolume", "BindLight", "BindObject", "BindClippingObject", "BindTexture",
           "BindBitmap", "RenderEstimate", "GetEstimate", "GetAutoFocusDistance", "GetNoIterations"]


def declared_c_functions():
    names = []
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        with open(os.path.join(INCLUDE, fn)) as f:
            text = f.read()
        names += re.findall(r"^ER_API\s+[\w\s\*]+?\b(er_\w+)\s*\(", text, flags=re.M)
    return names


def test_headers_declare_entry_points():
    names = declared_c_functions()
    assert len(names) >= 35
    for must in ("er_bind_tracer", "er_render_estimate", "er_get_estimate", "er_get_no_iterations",
                 "er_get_autofocus_distance", "er_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_c_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_covers_every_entry_point():
    from exposure_render import capi
    missing = [n for n in declared_c_functions() if n not in capi.SIGNATURES]
    assert not missing, missing


def test_struct_layout_matches(lib):
    from exposure_render import capi
    capi.check_abi(lib)
    assert lib.er_abi_version() >= 1


def test_cpp_api_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(LIB_DIR, "libexposurerender_amd.so")],
                         check=True, capture_output=True, text=True).stdout
    for name in CPP_API:
        assert re.search(r"\bExposureRender::%s\(" % name, out), name


def test_cpp_header_compiles_standalone(tmp_path):
    """An application includes only exposurerender.h and needs no HIP headers."""
    src = tmp_path / "app.cpp"
    src.write_text('#include "exposurerender.h"\nint main() { ExposureRender::ErTracer t; return t.ID == -1 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", INCLUDE, str(src)], check=True)
    csrc = tmp_path / "app.c"
    csrc.write_text('#include "er_capi.h"\nint main(void) { er_tracer_desc d; (void)d; return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", INCLUDE, str(csrc)], check=True)


def test_cpp_api_raises_without_device():
    """Without a GPU the first bind fails: the C++ API throws Exception(Error) with the message."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([CHECK, "nodevice"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "throws level 2" in r.stdout


def test_status_and_last_error(lib):
    """Unknown IDs report ER_WARNING with a message (list.cuh:150-164), without a device."""
    st = lib.er_get_no_iterations(12345, C.byref(C.c_int()))
    assert st in (2, 3)
    buf = C.create_string_buffer(256)
    lib.er_last_error(buf, 256)
    assert buf.value


@pytest.mark.gpu
def test_cpp_api_matches_c_abi():
    r = subprocess.run([CHECK, "render"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
