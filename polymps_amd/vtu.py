"""Reader for the VTU files that mps_write_vtu writes (ASCII, or binary with
appended raw data and UInt64 block headers), returning every DataArray as a
numpy array keyed by its Name.  Used to post-process output and by the tests.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

_TYPES = {"Float32": np.float32, "Float64": np.float64, "Int32": np.int32, "Int64": np.int64,
          "UInt8": np.uint8, "UInt64": np.uint64}
_ATTR = re.compile(r"(\w+)='([^']*)'")


def read_vtu(path) -> dict[str, np.ndarray]:
    raw = Path(path).read_bytes()
    start = raw.find(b"<AppendedData")
    head = raw if start < 0 else raw[:start]
    appended = b""
    if start >= 0:
        us = raw.index(b"_", raw.index(b">", start)) + 1
        appended = raw[us:]
    text = head.decode("ascii")
    out = {}
    for m in re.finditer(r"<DataArray([^>]*?)(/?)>", text):
        attrs = dict(_ATTR.findall(m.group(1)))
        dt = _TYPES[attrs["type"]]
        comps = int(attrs.get("NumberOfComponents", "1"))
        if attrs["format"] == "appended":
            off = int(attrs["offset"])
            nbytes = int(np.frombuffer(appended, np.uint64, 1, off)[0])
            arr = np.frombuffer(appended, dt, nbytes // np.dtype(dt).itemsize, off + 8).copy()
        elif attrs["format"] == "ascii":
            body = text[m.end():text.index("</DataArray>", m.end())]
            arr = np.array(body.split(), dtype=np.float64).astype(dt)
        else:
            raise ValueError(f"unsupported DataArray format {attrs['format']!r}")
        out[attrs["Name"]] = arr.reshape(-1, comps) if comps > 1 else arr
    return out
