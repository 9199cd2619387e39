"""ctypes mirror of ``mps_params`` (include/mps_params.h).

The Structure is generated from the header itself, so the Python view and
the C layout cannot drift apart; ``check_layout`` compares the size with the
library's own ``mps_params_size()``.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

HEADER = Path(__file__).resolve().parent.parent / "include" / "mps_params.h"

_CTYPE = {"int": ctypes.c_int, "double": ctypes.c_double}


def _parse_fields(text: str):
    start = text.index("typedef struct mps_params {") + len("typedef struct mps_params {")
    body = text[start: text.index("} mps_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for stmt in body.split(";"):
        stmt = stmt.strip()
        m = re.match(r"(int|double)\s+(.+)$", stmt, flags=re.S)
        if not m:
            continue
        base = _CTYPE[m.group(1)]
        for decl in m.group(2).split(","):
            decl = decl.strip()
            a = re.match(r"(\w+)\s*\[(\d+)\]$", decl)
            if a:
                fields.append((a.group(1), base * int(a.group(2))))
            else:
                fields.append((decl, base))
    return fields


def _enums(text: str) -> dict:
    out = {}
    for m in re.finditer(r"\b(MPS_[A-Z0-9_]+)\s*=\s*(\d+)", text):
        out[m.group(1)] = int(m.group(2))
    return out


_TEXT = HEADER.read_text()


class MpsParams(ctypes.Structure):
    _fields_ = _parse_fields(_TEXT)

    def copy(self) -> "MpsParams":
        c = MpsParams()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(MpsParams))
        return c

    def as_dict(self) -> dict:
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


ENUMS = _enums(_TEXT)
globals().update(ENUMS)


def check_layout(lib) -> None:
    n = lib.mps_params_size()
    if n != ctypes.sizeof(MpsParams):
        raise RuntimeError(f"mps_params layout mismatch: C {n} bytes, ctypes {ctypes.sizeof(MpsParams)}")
