"""The C-ABI library loads and exports every function its headers declare.

No compute calls here (no GPU in this environment).  CkFftInit without a GPU must
fail cleanly with NULL rather than crash, exactly like an allocation failure in the
reference (src/ckfft/context.cpp:61-66).
"""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ckfft_amd", "libckfft_amd.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "ckfft", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(CkFft\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_api():
    names = declared_functions()
    for f in ("CkFftInit", "CkFftRealForward", "CkFftRealInverse", "CkFftComplexForward",
              "CkFftComplexInverse", "CkFftShutdown"):
        assert f in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared_functions() - exported
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for name in declared_functions():
        assert getattr(lib, name)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_declared_symbols():
    import ckfft_amd
    assert set(ckfft_amd.EXPORTED_SYMBOLS) == declared_functions()
    assert b"gfx950" in ckfft_amd.CkFftGpuVersion()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_validation_paths_that_need_no_device():
    """Argument checks that reject before any device work (src/ckfft/ckfft.cpp:14-114)."""
    import ckfft_amd as C
    size = ctypes.c_size_t(0)
    assert not C.CkFftInit(0, 3, None, None)
    assert not C.CkFftInit(12, 3, None, None)
    assert not C.CkFftInit(16, 0, None, None)
    assert not C.CkFftInit(16, 4, None, None)
    buf = ctypes.create_string_buffer(64)
    assert not C.CkFftInit(16, 3, buf, None)                     # buf without size
    assert not C.CkFftInit(16, 3, None, ctypes.byref(size))     # query mode
    assert size.value > 0 and size.value % 8 == 0
    small = ctypes.c_size_t(size.value - 8)
    assert not C.CkFftInit(16, 3, buf, ctypes.byref(small))     # too small -> size written
    assert small.value == size.value
    dummy = ctypes.create_string_buffer(64)
    assert C.CkFftComplexForward(None, 4, dummy, buf) == 0
    assert C.CkFftRealInverse(None, 4, dummy, buf, None) == 0
    assert C.CkFftComplexForwardBatch(None, 4, 1, dummy, buf, None) == 0
    C.CkFftShutdown(None)


def test_product_does_not_reference_the_oracle():
    """The shipped package never imports or loads anything under oracle/."""
    for path in glob.glob(os.path.join(ROOT, "ckfft_amd", "**", "*"), recursive=True):
        if os.path.isfile(path) and path.endswith((".py", ".cpp", ".hip", ".h")):
            text = open(path).read()
            assert "import oracle" not in text and "libckfft_oracle" not in text and "_ref/" not in text, path


def test_headers_compile_as_c_and_link(tmp_path):
    """The headers compile as strict C (the reference's own test includes its header from
    C, src/test/test.c:1-3), and a C program taking the address of every declared function
    links against the library."""
    names = sorted(declared_functions())
    src = tmp_path / "use_api.c"
    src.write_text(
        '#include "ckfft/ckfft.h"\n#include "ckfft/ckfft_gpu.h"\n#include <stdio.h>\n'
        "int main(void) {\n"
        + "".join(f'    printf("%p\\n", (void*)&{n});\n' for n in names)
        + "    return 0;\n}\n")
    exe = tmp_path / "use_api"
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-Wno-pedantic",
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", os.path.dirname(LIB), "-lckfft_amd", "-Wl,--unresolved-symbols=ignore-in-shared-libs"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
