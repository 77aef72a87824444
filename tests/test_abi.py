"""The C ABI library on the CPU: it loads, exports every symbol the header declares, and
its ctypes mirror has the header's struct sizes.  No kernel is launched here."""
import ctypes as C
import os
import subprocess

import pytest

from syfrack_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "syfrack_hip.h")


def test_library_exports_every_header_symbol():
    lib = abi.load()
    names = abi.header_symbols(HEADER)
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)
    assert lib.syf_abi_version() == 6


STRUCTS = [
    ("syf_fractal", abi.SyfFractal), ("syf_camera", abi.SyfCamera), ("syf_light", abi.SyfLight),
    ("syf_material", abi.SyfMaterial), ("syf_job_desc", abi.SyfJobDesc), ("syf_stats", abi.SyfStats),
    ("syf_planar_desc", abi.SyfPlanarDesc), ("syf_server_config", abi.SyfServerConfig),
    ("syf_request_info", abi.SyfRequestInfo), ("syf_server_counters", abi.SyfServerCounters),
    ("syf_finish", abi.SyfFinish), ("syf_coloring", abi.SyfColoring), ("syf_ground", abi.SyfGround),
    ("syf_fog", abi.SyfFog),
]


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and offsetof of every field, from gcc on the header, against the ctypes mirror."""
    lines, expect = [], []
    for cname, cls in STRUCTS:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(cls, fname).offset)
    src = tmp_path / "sz.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n" f'#include "{HEADER}"\n'
                   "int main(void) {\n" + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect


def test_error_text_is_thread_local_and_null_checked():
    lib = abi.load()
    rc = lib.syf_job_create(None, None, None)
    assert rc == abi.SYF_E_INVALID
    assert b"null" in lib.syf_last_error()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="no fallback"):
        abi._LIB_saved = abi._LIB
        try:
            abi._LIB = None
            abi.load(str(tmp_path / "nope.so"))
        finally:
            abi._LIB = abi._LIB_saved
