"""The C ABI: librecview_hip.so exports every function include/recview_hip.h
declares, and the ctypes mirrors of its structs have the layout a C compiler
gives the header.  CPU only (the library loads without a GPU; nothing is called)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from recview_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "recview_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
        head = m.group(0)
        if head.startswith(("typedef", "#")):
            continue
        names.add(m.group(1))
    for m in re.finditer(r"^extern\s+int\s+(\w+)\s*;", text, flags=re.M):
        names.add(m.group(1))
    return names


def test_header_declarations_are_all_bound():
    names = declared_functions()
    assert len(names) >= 45
    assert names == set(_lib.EXPORTS), (sorted(names - set(_lib.EXPORTS)), sorted(set(_lib.EXPORTS) - names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def _c_offsets(struct, fields, d):
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"recview_hip.h\"\nint main(void){\n"
    src += f'printf("%zu\\n", sizeof({struct}));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof({struct}, {f}));\n'
    src += "return 0;}\n"
    cfile, exe = os.path.join(d, f"{struct}.c"), os.path.join(d, f"{struct}.bin")
    open(cfile, "w").write(src)
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), cfile, "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.parametrize("struct,cls", [("RECONST_INFO", _lib.ReconstInfo), ("RvSliceSpec", _lib.RvSliceSpec),
                                        ("RvOptions", _lib.RvOptions)])
def test_struct_layout_matches_c(struct, cls, tmp_path):
    fields = [f[0] for f in cls._fields_]
    size, offs = _c_offsets(struct, fields, str(tmp_path))
    assert size == C.sizeof(cls)
    assert offs == [getattr(cls, f).offset for f in fields]
