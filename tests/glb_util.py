"""Writers for small test inputs of the asset front end: glTF binaries (.glb),
PNG images, Radiance .hdr and OpenEXR skies, from numpy arrays (test helpers only)."""
import json
import struct
import zlib

import numpy as np

COMPONENT = {np.dtype(np.float32): 5126, np.dtype(np.uint32): 5125, np.dtype(np.uint16): 5123,
             np.dtype(np.uint8): 5121}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}


class GlbBuilder:
    """Accumulates accessors / images into one BIN buffer and writes a .glb."""

    def __init__(self):
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": []}
        self.bin = bytearray()

    def _view(self, data: bytes):
        while len(self.bin) % 4:
            self.bin.append(0)
        self.doc["bufferViews"].append({"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)})
        self.bin += data
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, normalized=False):
        arr = np.ascontiguousarray(arr)
        ncomp = 1 if arr.ndim == 1 else arr.shape[1]
        acc = {"bufferView": self._view(arr.tobytes()), "componentType": COMPONENT[arr.dtype],
               "count": int(arr.shape[0]), "type": TYPE[ncomp]}
        if normalized:
            acc["normalized"] = True
        if arr.dtype == np.float32 and ncomp == 3:
            acc["min"] = arr.min(0).tolist()
            acc["max"] = arr.max(0).tolist()
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def image(self, png: bytes):
        self.doc.setdefault("images", []).append({"bufferView": self._view(png), "mimeType": "image/png"})
        self.doc.setdefault("textures", []).append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def build(self):
        doc = dict(self.doc)
        doc["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(doc).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        b = bytes(self.bin) + b"\0" * ((4 - len(self.bin) % 4) % 4)
        total = 12 + 8 + len(js) + 8 + len(b)
        return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(b), 0x004E4942) + b)


def png(arr, palette=None, trns=None):
    """PNG bytes of an HxW, HxWxC uint8/uint16 array, or HxW palette indices."""
    arr = np.asarray(arr)
    h, w = arr.shape[:2]
    c = 1 if arr.ndim == 2 else arr.shape[2]
    if palette is not None:
        ctype, depth = 3, 8
    else:
        ctype = {1: 0, 2: 4, 3: 2, 4: 6}[c]
        depth = 16 if arr.dtype == np.uint16 else 8
    raw = bytearray()
    data = arr.astype(">u2").tobytes() if depth == 16 else arr.astype(np.uint8).tobytes()
    row = len(data) // h
    for y in range(h):
        filt = y % 5  # exercise every filter type
        line = np.frombuffer(data[y * row:(y + 1) * row], np.uint8).astype(np.int32)
        prev = (np.frombuffer(data[(y - 1) * row:y * row], np.uint8).astype(np.int32) if y else np.zeros(row, np.int32))
        bpp = max(1, c * depth // 8)
        out = np.zeros(row, np.int32)
        for i in range(row):
            a = line[i - bpp] if i >= bpp else 0
            b = prev[i]
            cc = prev[i - bpp] if i >= bpp else 0
            if filt == 0:
                pred = 0
            elif filt == 1:
                pred = a
            elif filt == 2:
                pred = b
            elif filt == 3:
                pred = (a + b) >> 1
            else:
                p = a + b - cc
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - cc)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else cc)
            out[i] = (line[i] - pred) & 255
        raw.append(filt)
        raw += out.astype(np.uint8).tobytes()

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
        if trns is not None:
            out += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    out += chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")
    return out


def hdr(rgbe, rle=True):
    """Radiance .hdr bytes of an HxWx4 uint8 RGBE array (new-style RLE or flat scanlines)."""
    rgbe = np.asarray(rgbe, np.uint8)
    h, w = rgbe.shape[:2]
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if not rle:
            out += rgbe[y].tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            comp = rgbe[y, :, c]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and comp[x + run] == comp[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, comp[x]])
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and comp[x + n] == comp[x + n + 1] == comp[x + n + 2]):
                        n += 1
                    out += bytes([n]) + comp[x:x + n].tobytes()
                    x += n
    return bytes(out)


def rgbe_to_float(rgbe):
    rgbe = np.asarray(rgbe, np.uint8)
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136).astype(np.float32), np.float32(0.0))
    return rgbe[..., :3].astype(np.float32) * scale[..., None]


def exr(channels, compression=0, lines_per_chunk=None):
    """Scanline OpenEXR bytes. channels: {name: HxW array} of float16 (HALF) or
    float32 (FLOAT); compression 0 NONE, 2 ZIPS, 3 ZIP."""
    names = sorted(channels)
    h, w = channels[names[0]].shape
    lpc = lines_per_chunk or {0: 1, 2: 1, 3: 16}[compression]

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(data)) + data

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", 1 if channels[n].dtype == np.float16 else 2,
                                                       0, 0, 0, 0, 1, 1) for n in names) + b"\0"
    head = struct.pack("<II", 20000630, 2)
    head += attr("channels", "chlist", chl)
    head += attr("compression", "compression", bytes([compression]))
    head += attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    head += attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    head += attr("lineOrder", "lineOrder", b"\0")
    head += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    head += attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    chunks = []
    for y0 in range(0, h, lpc):
        raw = b"".join(channels[n][y].astype("<f2" if channels[n].dtype == np.float16 else "<f4").tobytes()
                       for y in range(y0, min(h, y0 + lpc)) for n in names)
        data = raw
        if compression in (2, 3):
            b = np.frombuffer(raw, np.uint8)
            t = np.concatenate([b[0::2], b[1::2]]).astype(np.int32)
            d = t.copy()
            d[1:] = (t[1:] - t[:-1] + 128) & 255
            z = zlib.compress(d.astype(np.uint8).tobytes())
            data = z if len(z) < len(raw) else raw
        chunks.append(struct.pack("<iI", y0, len(data)) + data)
    table_at = len(head)
    offs, o = [], table_at + 8 * len(chunks)
    for c in chunks:
        offs.append(o)
        o += len(c)
    return head + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks)
