"""ctypes structures of include/vckss_mi355x.h, built from the header itself.

Every `typedef struct Name { ... } Name;` of the header becomes a
ctypes.Structure with the same fields in the same order (fixed-width
integers, doubles, byte arrays, pointers and nested structs), so the Python
binding and the C declarations cannot drift apart.  `STRUCTS[name]` holds the
class; `new(name)` returns a zeroed instance; `CONSTANTS[name]` holds the
value of every integer `#define` (plain, `u`-suffixed or `UINT64_C(1) << k`).
"""
from __future__ import annotations

import ctypes as C
import os
import re

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "vckss_mi355x.h")

_SCALARS = {"uint8_t": C.c_uint8, "uint32_t": C.c_uint32, "int32_t": C.c_int32, "uint64_t": C.c_uint64,
            "int64_t": C.c_int64, "double": C.c_double, "VckssInterruptPollV1": C.c_void_p}
_STRUCT = re.compile(r"typedef struct(?: \w+)? \{(.*?)\} (\w+);", re.S)
_DECL = re.compile(r"^(?:const\s+)?(\w+)\s+(.*)$", re.S)
_DECLARATOR = re.compile(r"^((?:\*\s*(?:const\s*)?)*)(\w+)((?:\[\d+\])*)$")


def _parse(path):
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    out = {}
    for body, name in _STRUCT.findall(text):
        fields = []
        for stmt in body.split(";"):
            stmt = " ".join(stmt.split())
            if not stmt:
                continue
            m = _DECL.match(stmt)
            if not m:
                raise ValueError(f"{name}: cannot parse field declaration {stmt!r}")
            typ, rest = m.groups()
            for raw in rest.split(","):
                d = _DECLARATOR.match(raw.strip())
                if not d:
                    raise ValueError(f"{name}: cannot parse declarator {raw!r}")
                star, nm, dims = d.groups()
                count = 1
                for n in re.findall(r"\[(\d+)\]", dims):
                    count *= int(n)
                fields.append((nm, typ, bool(star), count if dims else None))
        out[name] = fields
    return out


def _build(parsed):
    classes = {}
    for name, fields in parsed.items():
        spec = []
        for nm, typ, ptr, count in fields:
            if ptr:
                t = C.c_void_p
            elif typ in _SCALARS:
                t = _SCALARS[typ]
            else:
                t = classes[typ]
            if count:
                t = t * count
            spec.append((nm, t))
        classes[name] = type(name, (C.Structure,), {"_fields_": spec})
    return classes


STRUCTS = _build(_parse(HEADER))


def _constants(path):
    out = {}
    for name, value in re.findall(r"^#define (\w+) (.+?)\s*$", open(path).read(), re.M):
        m = re.fullmatch(r"(\d+)u?", value)
        if m:
            out[name] = int(m.group(1))
            continue
        m = re.fullmatch(r"\(UINT64_C\((\d+)\) << (\d+)\)", value)
        if m:
            out[name] = int(m.group(1)) << int(m.group(2))
    return out


CONSTANTS = _constants(HEADER)


def new(name):
    """A zeroed instance of the named header struct."""
    return STRUCTS[name]()


def as_dict(obj):
    """Nested dict of a structure's fields (arrays as lists)."""
    out = {}
    for name, _ in obj._fields_:
        v = getattr(obj, name)
        if isinstance(v, C.Structure):
            out[name] = as_dict(v)
        elif isinstance(v, C.Array):
            out[name] = list(v)
        else:
            out[name] = v
    return out
