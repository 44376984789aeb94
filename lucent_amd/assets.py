"""Asset ingestion through the C ABI: Radiance RGBE environment maps
(parseHDR, reference src/loaders/hdr.ts:7-93)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Union

import numpy as np

from . import _native


def decode_hdr(data: Union[bytes, bytearray, memoryview]) -> tuple[int, int, np.ndarray]:
    """A .hdr file's bytes -> (width, height, rgb float32 (height, width, 3),
    rows top to bottom), the Environment.map layout.  Raises LucentError
    with parseHDR's message on malformed input."""
    buf = np.frombuffer(bytes(data), np.uint8)
    w, h = C.c_uint32(), C.c_uint32()
    lib = _native.lib()
    _native.check(lib.lucent_decode_hdr(buf.ctypes.data, buf.size, C.byref(w), C.byref(h), None, 0))
    rgb = np.empty((h.value, w.value, 3), np.float32)
    _native.check(lib.lucent_decode_hdr(buf.ctypes.data, buf.size, C.byref(w), C.byref(h), rgb.ctypes.data, rgb.size))
    return w.value, h.value, rgb


def load_hdr(path: Union[str, Path]) -> tuple[int, int, np.ndarray]:
    return decode_hdr(Path(path).read_bytes())
