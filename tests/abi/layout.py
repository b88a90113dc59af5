"""Struct layouts declared in include/vckss_mi355x.h, as C programs that
print every size and field offset (used by tests/test_abi_layout.py)."""
import os

from fevc_amd import abi_types

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HEADER = os.path.join(ROOT, "include", "vckss_mi355x.h")


def structs(path=HEADER):
    """{struct name: [field names]} in declaration order (every declarator of
    a multi-name field line, as the ctypes binding reads it)."""
    return {name: [f[0] for f in fields] for name, fields in abi_types._parse(path).items()}


def program(include, layout):
    """C source printing `name size` and `name.field offset` lines."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{include}"', "int main(void) {"]
    for name, fields in layout.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"
