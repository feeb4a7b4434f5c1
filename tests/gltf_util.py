"""Writers for the glTF tests: a GLB container around a glTF JSON document and numpy
buffers, and PNG images with every row filter (test inputs; the loader must read back
exactly what these write)."""
from __future__ import annotations

import json
import struct
import zlib

import numpy as np

# ---------------------------------------------------------------- PNG


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def png(img, color_type=6, depth=8, palette=None, trns=None, filters=(0, 1, 2, 3, 4)):
    """PNG bytes of ``img`` ((h, w, channels) uint8 / uint16, or (h, w) palette indices);
    row y uses filter filters[y % len(filters)]."""
    img = np.asarray(img)
    h, w = img.shape[:2]
    if color_type == 3:
        idx = img.astype(np.uint8)
        rows = []
        for y in range(h):
            bits = "".join(format(int(v), f"0{depth}b") for v in idx[y])
            bits += "0" * (-len(bits) % 8)
            rows.append(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
        bpp = 1
    else:
        chans = {0: 1, 2: 3, 4: 2, 6: 4}[color_type]
        a = img.reshape(h, w, chans)
        if depth == 16:
            rows = [a[y].astype(">u2").tobytes() for y in range(h)]
            bpp = 2 * chans
        else:
            rows = [a[y].astype(np.uint8).tobytes() for y in range(h)]
            bpp = chans
    out = bytearray()
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        f = filters[y % len(filters)]
        enc = bytearray(len(row))
        for x in range(len(row)):
            a = row[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][f]
            enc[x] = (row[x] - pred) & 0xFF
        out += bytes([f]) + enc
        prev = row

    def chunk(t, data):
        return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0)
    body = chunk(b"IHDR", ihdr)
    if palette is not None:
        body += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        body += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8).reshape(-1)))
    body += chunk(b"IDAT", zlib.compress(bytes(out), 6)) + chunk(b"IEND", b"")
    return b"\x89PNG\r\n\x1a\n" + body


# ---------------------------------------------------------------- GLB


class GltfBuilder:
    """Accumulates accessors / buffer views over one binary buffer."""

    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": []}

    def _view(self, data: bytes, stride=None):
        while len(self.bin) % 4:
            self.bin += b"\0"
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, ctype=5126, type_=None, stride=None):
        arr = np.asarray(arr)
        comps = 1 if arr.ndim == 1 else arr.shape[1]
        dt = {5126: "<f4", 5125: "<u4", 5123: "<u2", 5121: "u1"}[ctype]
        data = arr.astype(dt)
        if stride:  # interleave with padding to exercise byteStride
            el = data.reshape(len(data), -1)
            pad = stride - el.itemsize * el.shape[1]
            raw = b"".join(r.tobytes() + b"\xab" * pad for r in el)
        else:
            raw = data.tobytes()
        view = self._view(raw, stride)
        t = type_ or {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}[comps]
        self.doc["accessors"].append({"bufferView": view, "componentType": ctype, "count": int(len(arr)), "type": t})
        return len(self.doc["accessors"]) - 1

    def image(self, png_bytes):
        view = self._view(png_bytes)
        self.doc.setdefault("images", []).append({"bufferView": view, "mimeType": "image/png"})
        self.doc.setdefault("textures", []).append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def glb(self):
        self.doc["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(self.doc).encode()
        js += b" " * (-len(js) % 4)
        b = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
        total = 12 + 8 + len(js) + 8 + len(b)
        return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(b), 0x004E4942) + b)
