"""The C-ABI library loads without a GPU and exports every symbol include/octmps.h declares."""
import ctypes
import os
import re

import octmps as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "octmps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(octmps_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 40
    lib = ctypes.CDLL(O.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {n for n, _, _ in O.SIGNATURES}
    assert set(declared_functions()) == bound


def test_status_strings_and_version():
    assert O.lib.octmps_status_string(0) == b"ok"
    assert O.lib.octmps_status_string(4).startswith(b"probe")
    assert b"gfx950" in O.lib.octmps_version()


def test_no_device_is_an_error_not_a_crash():
    # Without a GPU the context refuses loudly; there is no CPU fallback.
    if os.path.exists("/dev/kfd"):
        return
    try:
        O.Context(0)
    except O.OctmpsError as e:
        assert e.status == 8
    else:
        raise AssertionError("context created without a device")


# ctypes mirror -> C struct name of include/octmps.h
STRUCTS = {
    "Region": "octmps_region", "Params": "octmps_params", "GridSpec": "octmps_grid_spec",
    "AScan": "octmps_ascan", "RunStats": "octmps_run_stats", "BatchOpts": "octmps_batch_opts",
    "ServeOpts": "octmps_serve_opts", "Request": "octmps_request", "RequestInfo": "octmps_request_info",
}


def test_struct_layouts_match_the_header(tmp_path):
    # Compile a probe against the header with the host C compiler and compare the size and
    # every field offset with the ctypes mirror used by the binding.
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "octmps.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in getattr(O, py)._fields_:
            lines.append(f'printf("{py} {f} %zu\\n", offsetof({c}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        py, field, value = line.split()
        cls = getattr(O, py)
        want = ctypes.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert int(value) == want, (py, field, int(value), want)


def test_integration_example_compiles(tmp_path):
    # The C++ binding example of the integration guide, wrapped in a main that declares the names it
    # takes from the reference's host code, must compile against include/octmps.h.
    import re
    import subprocess

    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"```cpp\n(.*?)```", text, re.S).group(1)
    body = "\n".join(l for l in block.splitlines() if not l.startswith("#include"))
    src = tmp_path / "example.cpp"
    src.write_text('#include <cstdio>\n#include <vector>\n#include "octmps.h"\n'
                   "int main(int argc, char **argv) {\n"
                   "    const char *filename_opt = argv[1], *filename_bias = argv[2], *filename_mesh = argv[3];\n"
                   "    uint64_t seed = 1;\n    double x;\n" + body + "\n    return 0;\n}\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)
