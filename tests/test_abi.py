"""The C ABI library on the CPU: it loads, exports every symbol the header declares, and
its ctypes mirror has the header's struct sizes.  No kernel is launched here."""
import ctypes as C
import os
import subprocess
import textwrap

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
    assert lib.syf_abi_version() == 1


def test_struct_sizes_match_the_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include "{HEADER}"
        int main(void) {{
          printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(syf_fractal), sizeof(syf_camera), sizeof(syf_light),
                 sizeof(syf_material), sizeof(syf_job_desc), sizeof(syf_stats), sizeof(syf_planar_desc));
          return 0;
        }}"""))
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    mine = [C.sizeof(t) for t in (abi.SyfFractal, abi.SyfCamera, abi.SyfLight, abi.SyfMaterial, abi.SyfJobDesc,
                                  abi.SyfStats, abi.SyfPlanarDesc)]
    assert sizes == mine


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
