"""Displacement maps on disk and the tile set-up of a displacement asset.

The reference reads height maps with stb_image / OpenEXR
(load_height_texture, texture_grid.cpp:10-54): 8-bit and 16-bit images
become values in [0, 1] (the mean of up to three channels / 255 or / 65535),
EXR keeps linear floats. A DisplacementAsset (displacement_asset.cpp:16-35)
box-downsamples the image to a power-of-two `resolution`, closes the tile and
builds its pyramid with strength and offset -strength * midlevel.

Here:

* PNG (grey, grey+alpha, RGB, RGBA; 8 or 16 bits; not interlaced) is
  decoded with zlib and the five PNG row filters, with the reference's
  channel mean and scaling. EXR needs a library this build does not have and
  raises.
* `.npy` (float arrays) and the tile file `.dmt` hold float texels as they
  are: a 32-byte header (magic "DMT1", n, repeat, strength, offset) and n^2
  little-endian fp32 texels, row-major [j][i], readable with numpy.memmap
  for large resident tiles.
* The reference's synthetic maps (checkerboard, gradient_map,
  gaussian_spot, texture_grid.cpp:57-97) and downsample_box (:99-116).
"""
import struct
import zlib
from pathlib import Path

import numpy as np

DMT_MAGIC = b"DMT1"
DMT_HEADER = struct.Struct("<4sIIdd4x")  # magic, n, repeat, strength, offset (32 bytes)


# ------------------------------------------------------------------ PNG

def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def _unfilter(raw, h, stride, bpp):
    """Undo the PNG row filters (None, Sub, Up, Average, Paeth)."""
    out = np.zeros((h, stride), dtype=np.int64)
    prev = np.zeros(stride, dtype=np.int64)
    pos = 0
    for y in range(h):
        ftype = raw[pos]
        line = np.frombuffer(raw, dtype=np.uint8, count=stride, offset=pos + 1).astype(np.int64)
        pos += 1 + stride
        cur = line.copy()
        if ftype == 0:
            pass
        elif ftype == 2:
            cur = (line + prev) & 0xFF
        elif ftype in (1, 3, 4):
            # left-dependent filters run byte by byte within the row
            for x in range(stride):
                left = cur[x - bpp] if x >= bpp else 0
                up = prev[x]
                if ftype == 1:
                    pred = left
                elif ftype == 3:
                    pred = (left + up) >> 1
                else:
                    ul = prev[x - bpp] if x >= bpp else 0
                    pred = int(_paeth(np.int64(left), np.int64(up), np.int64(ul)))
                cur[x] = (line[x] + pred) & 0xFF
        else:
            raise ValueError(f"PNG: unknown row filter {ftype}")
        out[y] = cur
        prev = cur
    return out


def read_png(path):
    """(values (H, W) float64 in [0, 1], raw integer image) of a PNG."""
    data = Path(path).read_bytes()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        length, ctype = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if ctype == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif ctype == b"IDAT":
            idat.append(body)
        elif ctype == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, color, _, _, interlace = hdr
    channels = {0: 1, 4: 2, 2: 3, 6: 4}.get(color)
    if channels is None or depth not in (8, 16) or interlace != 0:
        raise ValueError(f"{path}: PNG colour type {color}, depth {depth}, interlace {interlace} not supported "
                         "(grey/RGB(A), 8 or 16 bits, not interlaced)")
    bps = depth // 8
    stride = w * channels * bps
    rows = _unfilter(zlib.decompress(b"".join(idat)), h, stride, channels * bps)
    if bps == 2:
        img = (rows[:, 0::2] << 8) | rows[:, 1::2]
    else:
        img = rows
    img = img.reshape(h, w, channels)
    nc = min(channels, 3)
    # the mean of up to three channels, over the type's full scale (texture_grid.cpp:25-51)
    vals = img[:, :, :nc].sum(axis=2) / nc / (65535.0 if bps == 2 else 255.0)
    return vals, img


def write_png(path, img, depth=8, filters=None):
    """A PNG of an integer image (H, W, C) with C in 1..4 (for tests and for
    exporting tiles). `filters` picks the row filter per row (default None)."""
    img = np.asarray(img, dtype=np.int64)
    if img.ndim == 2:
        img = img[:, :, None]
    h, w, c = img.shape
    color = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    bps = depth // 8
    if bps == 2:
        b = np.stack([(img >> 8) & 0xFF, img & 0xFF], axis=-1).reshape(h, w * c * 2)
    else:
        b = (img & 0xFF).reshape(h, w * c)
    stride, bpp = b.shape[1], c * bps
    out = bytearray()
    prev = np.zeros(stride, dtype=np.int64)
    for y in range(h):
        f = 0 if filters is None else int(filters[y % len(filters)])
        line = b[y]
        enc = np.zeros(stride, dtype=np.int64)
        for x in range(stride):
            left = line[x - bpp] if x >= bpp else 0
            up = prev[x]
            ul = prev[x - bpp] if x >= bpp else 0
            pred = [0, left, up, (left + up) >> 1, int(_paeth(np.int64(left), np.int64(up), np.int64(ul)))][f]
            enc[x] = (line[x] - pred) & 0xFF
        out.append(f)
        out.extend(enc.astype(np.uint8).tobytes())
        prev = line

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b"")
    Path(path).write_bytes(png)


# ------------------------------------------------------------ tile files

def save_tile(path, texels, strength=1.0, offset=0.0, repeat=True):
    """A .dmt tile file: header + n^2 fp32 texels [j][i]."""
    t = np.ascontiguousarray(texels, dtype="<f4")
    n = t.shape[0]
    if t.shape != (n, n):
        raise ValueError("a tile is square")
    with open(path, "wb") as f:
        f.write(DMT_HEADER.pack(DMT_MAGIC, n, 1 if repeat else 0, float(strength), float(offset)))
        f.write(t.tobytes())


def load_tile(path, mmap=True):
    """(texels, strength, offset, repeat) of a .dmt tile file; the texels
    are memory-mapped unless mmap is False."""
    with open(path, "rb") as f:
        magic, n, repeat, strength, offset = DMT_HEADER.unpack(f.read(DMT_HEADER.size))
    if magic != DMT_MAGIC:
        raise ValueError(f"{path}: not a tile file")
    if mmap:
        tex = np.memmap(path, dtype="<f4", mode="r", offset=DMT_HEADER.size, shape=(n, n))
    else:
        tex = np.fromfile(path, dtype="<f4", count=n * n, offset=DMT_HEADER.size).reshape(n, n)
    return tex, strength, offset, bool(repeat)


def load_height_texture(path):
    """Texel values [j, i] of a height map file (load_height_texture,
    texture_grid.cpp:10-54): PNG as the reference scales it, .npy and .dmt
    as stored."""
    p = Path(path)
    ext = p.suffix.lower()
    if ext == ".png":
        return read_png(p)[0]
    if ext == ".npy":
        return np.load(p, allow_pickle=False).astype(np.float64)
    if ext == ".dmt":
        return np.asarray(load_tile(p, mmap=False)[0], dtype=np.float64)
    if ext == ".exr":
        raise ValueError(f"{path}: EXR needs OpenEXR, which this build does not have; convert it to .npy or .dmt")
    raise ValueError(f"{path}: unknown height map format {ext}")


# ------------------------------------------------------ synthetic maps

def checkerboard(n, cells, lo, hi):
    """texture_grid.cpp:57-69."""
    i = np.arange(n)
    ci, cj = i * cells // n, i * cells // n
    return np.where((cj[:, None] + ci[None, :]) % 2 == 0, lo, hi).astype(np.float64)


def gradient_map(n, lo, hi):
    """texture_grid.cpp:71-83."""
    x = (np.arange(n) + 0.5) / n
    return lo + (hi - lo) * 0.5 * (x[None, :] + x[:, None])


def gaussian_spot(n, cx, cy, sigma, peak, floor_value):
    """texture_grid.cpp:85-97."""
    x = (np.arange(n) + 0.5) / n
    du, dv = x[None, :] - cx, x[:, None] - cy
    return floor_value + peak * np.exp(-(du * du + dv * dv) / (2.0 * sigma * sigma))


def downsample_box(tex, n):
    """Box mean over (H/n) x (W/n) blocks (texture_grid.cpp:99-116); rows and
    columns beyond n whole blocks are dropped, as there."""
    H, W = tex.shape
    bh, bw = H // n, W // n
    if bh < 1 or bw < 1:
        raise ValueError(f"texture {W}x{H} too small for {n} nodes")
    return tex[:n * bh, :n * bw].reshape(n, bh, n, bw).mean(axis=(1, 3))


def load_asset(path, strength=1.0, midlevel=0.5, repeat=True, resolution=0, build="fold"):
    """A resident tile from a height map, as DisplacementAsset does
    (displacement_asset.cpp:16-35, create_displacement_asset :44-56):
    square map, power-of-two resolution (default the map's), box
    downsampling, h = strength * (texel - midlevel). A .dmt file carries its
    own strength, offset and wrap."""
    import dmap_amd
    p = Path(path)
    if p.suffix.lower() == ".dmt":
        tex, s, off, rep = load_tile(p)
        return dmap_amd.DisplacementTile(np.asarray(tex, dtype=np.float32), s, off, rep, build)
    img = load_height_texture(p)
    if img.shape[0] != img.shape[1]:
        raise ValueError(f"{path}: a displacement map must be square, got {img.shape[1]} x {img.shape[0]}")
    res = resolution or img.shape[0]
    if res & (res - 1) or res < 1:
        raise ValueError(f"{path}: the tile needs 2^L texels per side, got {res}")
    if res > img.shape[0]:
        raise ValueError(f"{path}: resolution {res} exceeds the map's {img.shape[0]} texels")
    tex = img if res == img.shape[0] else downsample_box(img, res)
    return dmap_amd.DisplacementTile(tex.astype(np.float32), strength, -strength * midlevel, repeat, build)
