"""The C-ABI library: it loads, exports every function include/tval3_amd.h
declares, and its structs have the layout the ctypes mirrors assume (checked by
compiling the header with gcc).  No GPU calls."""
import os
import re
import subprocess
import ctypes as C

import pytest

import tval3_amd
from tval3_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "tval3_amd.h")


def _declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tv3_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = C.CDLL(tval3_amd.lib_path())
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_loads():
    L = tval3_amd.lib()
    o = abi.Options()
    L.tv3_options_default(C.byref(o))
    assert o.mu == 256.0 and o.beta == 32.0 and o.ls_exact == 0


@pytest.mark.parametrize("name,pycls", [("tv3_options", abi.Options), ("tv3_info", abi.Info),
                                        ("tv3_trace_rec", abi.TraceRec),
                                        ("tv3_geometry", abi.Geometry),
                                        ("tv3_profile_slot", abi.ProfileSlot)])
def test_struct_layout(tmp_path, name, pycls):
    fields = [f for f, _ in pycls._fields_]
    src = tmp_path / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"',
             'int main(void) {', f'  printf("%zu\\n", sizeof({name}));']
    lines += [f'  printf("%zu\\n", offsetof({name}, {f}));' for f in fields]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert out[0] == C.sizeof(pycls)
    assert out[1:] == [getattr(pycls, f).offset for f in fields]
