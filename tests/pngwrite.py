"""A small PNG encoder for tests: every colour type and bit depth, per-row
filter types chosen by the caller, optional Adam7 interlacing, PLTE/tRNS.
Independent of the library's decoder, so decode(encode(x)) == x pins the
decoder's unfiltering, bit unpacking and pass layout."""
import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
         (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(kind, data):
    return (struct.pack(">I", len(data)) + kind + data +
            struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))


def _pack_row(samples, depth):
    """samples: 1-D int array of one scanline's samples."""
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    if depth == 8:
        return bytes(int(v) for v in samples)
    out = bytearray((len(samples) * depth + 7) // 8)
    for i, v in enumerate(samples):
        bit = i * depth
        out[bit // 8] |= int(v) << (8 - depth - bit % 8)
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter(row, prev, bpp, ft):
    out = bytearray(len(row))
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
        out[i] = (row[i] - pred) & 0xFF
    return bytes([ft]) + bytes(out)


def encode(samples, color_type, depth, palette=None, trns=None, interlace=False, seed=0):
    """samples: (H, W, C) integer array of raw sample values. Returns bytes."""
    samples = np.asarray(samples)
    h, w, ch = samples.shape
    assert ch == CHANNELS[color_type]
    bpp = max(1, ch * depth // 8)
    rng = np.random.default_rng(seed)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for row in sub:
            packed = _pack_row(row.reshape(-1), depth)
            raw += _filter(packed, prev, bpp, int(rng.integers(0, 5)))
            prev = packed
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, int(interlace)))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    data = zlib.compress(bytes(raw), 6)
    for k in range(0, len(data), 97):  # several IDAT chunks
        out += _chunk(b"IDAT", data[k:k + 97])
    out += _chunk(b"IEND", b"")
    return out
