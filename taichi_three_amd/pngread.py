"""PNG decoding for textures when PIL is not installed.

Reads the non-interlaced PNG images of 1, 2, 4 or 8 bits per channel (grey,
grey + alpha, RGB, RGBA, palette) and 16-bit ones (kept as uint16), following
the PNG specification (ISO/IEC 15948): the IDAT stream inflated with zlib, then
each scanline un-filtered with its filter type (None, Sub, Up, Average, Paeth).
Returns the image as an array [rows, columns, channels] with row 0 at the top,
as PIL's np.asarray(Image.open(...)) gives it.
"""
import struct
import zlib

import numpy as np

__all__ = ['read_png', 'paeth']

_SIGNATURE = b'\x89PNG\r\n\x1a\n'
_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def paeth(a, b, c):
    """The Paeth predictor of the specification (left a, up b, upper-left c), elementwise."""
    a, b, c = (np.asarray(x, dtype=np.int32) for x in (a, b, c))
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def _unfilter(raw, height, stride, bpp):
    out = np.zeros((height, stride), dtype=np.uint8)
    prev = np.zeros(stride, dtype=np.int32)
    pos = 0
    for y in range(height):
        ftype = raw[pos]
        line = np.frombuffer(raw, dtype=np.uint8, count=stride, offset=pos + 1).astype(np.int32)
        pos += stride + 1
        if ftype == 0:
            cur = line
        elif ftype == 2:
            cur = (line + prev) & 0xFF
        elif ftype == 1:
            # Sub: a running sum along each of the bpp interleaved byte lanes
            lanes = np.zeros(((stride + bpp - 1) // bpp) * bpp, dtype=np.int64)
            lanes[:stride] = line
            cur = (np.cumsum(lanes.reshape(-1, bpp), axis=0).reshape(-1)[:stride] & 0xFF).astype(np.int32)
        elif ftype in (3, 4):
            # Average and Paeth read the reconstructed byte bpp to the left: sequential
            c = line.tolist()
            up = prev.tolist()
            for x in range(stride):
                left = c[x - bpp] if x >= bpp else 0
                if ftype == 3:
                    pred = (left + up[x]) >> 1
                else:
                    ul = up[x - bpp] if x >= bpp else 0
                    pp = left + up[x] - ul
                    pa, pb, pc = abs(pp - left), abs(pp - up[x]), abs(pp - ul)
                    pred = left if (pa <= pb and pa <= pc) else (up[x] if pb <= pc else ul)
                c[x] = (c[x] + pred) & 0xFF
            cur = np.asarray(c, dtype=np.int32)
        else:
            raise ValueError(f'PNG: bad filter type {ftype} on row {y}')
        out[y] = cur
        prev = cur
    return out


def read_png(data):
    """data: the bytes of a PNG file, or a path."""
    if isinstance(data, str):
        with open(data, 'rb') as fh:
            data = fh.read()
    if data[:8] != _SIGNATURE:
        raise ValueError('not a PNG file')
    pos, idat, palette, trns, hdr = 8, [], None, None, None
    while pos < len(data):
        length, kind = struct.unpack('>I4s', data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + length]
        if zlib.crc32(kind + body) & 0xFFFFFFFF != struct.unpack('>I', data[pos + 8 + length:pos + 12 + length])[0]:
            raise ValueError(f'PNG: CRC mismatch in chunk {kind!r}')
        pos += 12 + length
        if kind == b'IHDR':
            hdr = struct.unpack('>IIBBBBB', body)
        elif kind == b'PLTE':
            palette = np.frombuffer(body, dtype=np.uint8).reshape(-1, 3)
        elif kind == b'tRNS':
            trns = np.frombuffer(body, dtype=np.uint8)
        elif kind == b'IDAT':
            idat.append(body)
        elif kind == b'IEND':
            break
    if hdr is None:
        raise ValueError('PNG: no IHDR chunk')
    width, height, depth, ctype, _, _, interlace = hdr
    if interlace:
        raise NotImplementedError('PNG: interlaced images are not supported')
    if ctype not in _CHANNELS or depth not in (1, 2, 4, 8, 16):
        raise ValueError(f'PNG: colour type {ctype} with bit depth {depth}')
    ch = _CHANNELS[ctype]
    bits = ch * depth
    stride = (width * bits + 7) // 8
    raw = zlib.decompress(b''.join(idat))
    if len(raw) < height * (stride + 1):
        raise ValueError(f'PNG: {len(raw)} bytes of scanlines for {height} rows of {stride + 1}')
    rows = _unfilter(raw, height, stride, max(1, bits // 8))
    if depth == 16:
        img = rows.view('>u2').reshape(height, width, ch).astype(np.uint16)
    elif depth == 8:
        img = rows.reshape(height, width, ch)
    else:
        per = 8 // depth
        shifts = (8 - depth) - depth * np.arange(per)
        vals = (rows[:, :, None] >> shifts[None, None, :]) & ((1 << depth) - 1)
        img = vals.reshape(height, -1)[:, :width * ch].reshape(height, width, ch).astype(np.uint8)
        if ctype == 0:
            img = (img.astype(np.uint16) * 255 // ((1 << depth) - 1)).astype(np.uint8)
    if ctype == 3:
        if palette is None:
            raise ValueError('PNG: palette image without PLTE')
        idx = img[..., 0]
        img = palette[idx]
        if trns is not None:
            alpha = np.full(len(palette), 255, dtype=np.uint8)
            alpha[:len(trns)] = trns
            img = np.concatenate([img, alpha[idx][..., None]], axis=2)
    return img[..., 0] if img.shape[2] == 1 else img
