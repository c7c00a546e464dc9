"""Film output and image comparison for rendered films.

Films are (H, W, 3) arrays whose row 0 is the bottom of the image (raster y = 0, the
layout of HDRBitmapFilm, src/liblightmetrica/hdrfilm.cpp:161-166).

* write_hdr / read_hdr -- Radiance RGBE (.hdr), the reference's default output
  (HDRBitmapFilm::RescaleAndSave, hdrfilm.cpp:199-304, through FreeImage, whose scanline 0
  is the bottom row: the saved picture is upright).  Written as flat RGBE scanlines, top row
  first ("-Y H +X W"); the reader also decodes run-length-encoded scanlines.
* write_pfm / read_pfm -- Portable Float Map, little-endian, rows bottom to top: the film
  verbatim in FP32.
* rmse -- BitmapImage::EvaluateRMSE (src/liblightmetrica/bitmap.cpp:31-52).

Command line:  python -m lightmetrica_amd.film_io rmse A B [--weight W]
"""
from __future__ import annotations

import argparse
import sys
import warnings
from pathlib import Path

import numpy as np

__all__ = ["write_hdr", "read_hdr", "write_pfm", "read_pfm", "rmse", "rgbe_encode", "rgbe_decode"]


def rgbe_encode(rgb: np.ndarray) -> np.ndarray:
    """(..., 3) float -> (..., 4) uint8 shared-exponent RGBE (Ward, Graphics Gems II):
    v = max channel, (m, e) = frexp(v), bytes = channel * m * 256 / v and e + 128; v < 1e-32 -> 0."""
    rgb = np.asarray(rgb, dtype=np.float64)
    v = rgb.max(axis=-1)
    m, e = np.frexp(v)
    out = np.zeros(rgb.shape[:-1] + (4,), dtype=np.uint8)
    ok = v >= 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out[..., :3] = np.where(ok[..., None], np.clip(np.floor(rgb * scale[..., None]), 0, 255), 0).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def rgbe_decode(rgbe: np.ndarray) -> np.ndarray:
    """(..., 4) uint8 -> (..., 3) float32: (byte + 0.5) * 2^(e - 136); e = 0 -> 0."""
    rgbe = np.asarray(rgbe, dtype=np.uint8)
    e = rgbe[..., 3].astype(np.int32)
    f = np.ldexp(1.0, e - 136)
    out = (rgbe[..., :3].astype(np.float64) + 0.5) * f[..., None]
    return np.where((e == 0)[..., None], 0.0, out).astype(np.float32)


def write_hdr(path, film: np.ndarray) -> Path:
    """Write a bottom-up film as a Radiance .hdr file (the extension is appended if missing)."""
    path = Path(path)
    if path.suffix == "":
        path = path.with_suffix(".hdr")
    film = np.asarray(film)
    h, w, _ = film.shape
    top_first = film[::-1]
    data = rgbe_encode(top_first).tobytes()
    header = f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode("ascii")
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(header + data)
    return path


def _read_rle_scanline(buf: bytes, pos: int, w: int):
    """One new-style RLE scanline (starts 2 2 hi lo): four channel planes, runs and dumps."""
    line = np.empty((4, w), dtype=np.uint8)
    pos += 4
    for ch in range(4):
        x = 0
        while x < w:
            count = buf[pos]
            pos += 1
            if count > 128:
                count -= 128
                line[ch, x:x + count] = buf[pos]
                pos += 1
            else:
                line[ch, x:x + count] = np.frombuffer(buf, dtype=np.uint8, count=count, offset=pos)
                pos += count
            x += count
    return line.T, pos


def read_hdr(path) -> np.ndarray:
    """Read a Radiance .hdr (flat or RLE, "-Y H +X W") into a bottom-up (H, W, 3) film."""
    buf = Path(path).read_bytes()
    pos = 0
    lines = []
    while True:
        end = buf.index(b"\n", pos)
        line = buf[pos:end].decode("ascii", "replace")
        pos = end + 1
        if line.startswith("-Y") or line.startswith("+Y"):
            break
        lines.append(line)
    if not lines or not lines[0].startswith("#?"):
        raise ValueError("not a Radiance HDR file")
    parts = line.split()
    if parts[0] != "-Y" or parts[2] != "+X":
        raise ValueError(f"unsupported orientation '{line}'")
    h, w = int(parts[1]), int(parts[3])
    img = np.empty((h, w, 4), dtype=np.uint8)
    for y in range(h):
        if 8 <= w < 0x8000 and buf[pos] == 2 and buf[pos + 1] == 2 and ((buf[pos + 2] << 8) | buf[pos + 3]) == w:
            img[y], pos = _read_rle_scanline(buf, pos, w)
        else:
            img[y] = np.frombuffer(buf, dtype=np.uint8, count=4 * w, offset=pos).reshape(w, 4)
            pos += 4 * w
    return rgbe_decode(img)[::-1].copy()


def write_pfm(path, film: np.ndarray) -> Path:
    """Write a bottom-up film as a little-endian colour PFM (rows bottom to top, FP32)."""
    path = Path(path)
    film = np.ascontiguousarray(film, dtype="<f4")
    h, w, _ = film.shape
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(f"PF\n{w} {h}\n-1.0\n".encode("ascii") + film.tobytes())
    return path


def read_pfm(path) -> np.ndarray:
    buf = Path(path).read_bytes()
    fields, pos = [], 0
    while len(fields) < 4:
        while buf[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not buf[end:end + 1].isspace():
            end += 1
        fields.append(buf[pos:end].decode("ascii"))
        pos = end
    pos += 1  # the single whitespace after the scale
    kind, w, h, scale = fields[0], int(fields[1]), int(fields[2]), float(fields[3])
    if kind != "PF":
        raise ValueError("only colour PFM (PF) is supported")
    dt = "<f4" if scale < 0 else ">f4"
    return np.frombuffer(buf, dtype=dt, count=w * h * 3, offset=pos).reshape(h, w, 3).astype(np.float32)


def rmse(a, b, weight: float = 1.0) -> float:
    """BitmapImage::EvaluateRMSE: sqrt(mean over all W*H*3 values of (a - weight * b)^2);
    like the reference, a size mismatch warns and gives 0."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if a.size != b.size:
        warnings.warn(f"Invalid image size : {a.size} != {b.size}")
        return 0.0
    d = a - b * weight
    return float(np.sqrt(np.mean(d * d)))


def _load(path) -> np.ndarray:
    p = Path(path)
    if p.suffix == ".hdr":
        return read_hdr(p)
    if p.suffix == ".pfm":
        return read_pfm(p)
    if p.suffix == ".npy":
        return np.load(p, allow_pickle=False)
    raise ValueError(f"unsupported image type '{p.suffix}'")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m lightmetrica_amd.film_io")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("rmse", help="RMSE between two images (.hdr, .pfm or .npy)")
    r.add_argument("a")
    r.add_argument("b")
    r.add_argument("--weight", type=float, default=1.0)
    args = ap.parse_args(argv)
    if args.cmd == "rmse":
        print(f"{rmse(_load(args.a), _load(args.b), args.weight):.9g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
