"""Struct layouts of the C ABI (include/vckss_mi355x.h).

  * tests/abi/layout_asserts.c holds, as _Static_assert, every size and
    offset the reference header asserts for a struct both headers declare
    (rust/stata_backend/include/vckss_rust.h:2146-2330,
    rust/stata_backend/tests/abi_header_compat_test.c); gcc must accept it.
  * The ctypes classes of fevc_amd.abi_types (built from the header text)
    have the C compiler's size and every field offset of every struct.
  * Where the reference tree is present, a struct both headers declare has
    the same fields, size and offsets in both (compiled side by side).
"""
import ctypes as C
import os
import shutil
import subprocess
import sys

import pytest

from fevc_amd import abi_types

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "abi"))
import layout  # noqa: E402

REFERENCE_HEADER = "/root/reference/rust/stata_backend/include/vckss_rust.h"
gcc = shutil.which("gcc")
pytestmark = pytest.mark.skipif(gcc is None, reason="gcc is required to compile the layout checks")


def _offsets(tmp_path, include, structs, tag):
    src = tmp_path / f"{tag}.c"
    exe = tmp_path / tag
    src.write_text(layout.program(include, structs))
    subprocess.run([gcc, "-std=c11", "-O0", "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.splitlines())


def test_static_asserts_compile():
    here = os.path.join(os.path.dirname(__file__), "abi", "layout_asserts.c")
    r = subprocess.run([gcc, "-std=c11", "-fsyntax-only", "-Wall", "-Werror", here], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def test_ctypes_binding_matches_c(tmp_path):
    structs = layout.structs()
    assert len(structs) == len(abi_types.STRUCTS) >= 59
    got = _offsets(tmp_path, layout.HEADER, structs, "ours")
    for name, fields in structs.items():
        cls = abi_types.STRUCTS[name]
        assert C.sizeof(cls) == int(got[name]), name
        for f in fields:
            assert getattr(cls, f).offset == int(got[f"{name}.{f}"]), (name, f)


@pytest.mark.skipif(not os.path.exists(REFERENCE_HEADER), reason="reference tree not present")
def test_shared_structs_match_reference_header(tmp_path):
    ours = layout.structs()
    ref = layout.structs(REFERENCE_HEADER)
    shared = {k: v for k, v in ours.items() if k in ref}
    assert len(shared) == len(ours), sorted(set(ours) - set(ref))
    for k, fields in shared.items():
        assert fields == ref[k], k
    a = _offsets(tmp_path, layout.HEADER, shared, "ours")
    b = _offsets(tmp_path, REFERENCE_HEADER, shared, "ref")
    assert a == b
