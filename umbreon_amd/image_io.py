"""Image output and comparison of the reference's bench harness
(src/bench/image/image_io.{hpp,cpp}): writeImage (binary PPM for ".ppm",
PNG otherwise; RGB sRGB-encoded, alpha stored linear), readPpm (sRGB-decoded
to linear) and compareImages (PSNR over sRGB-encoded RGB, SSIM over sRGB luma
with an 8x8 window at step 4)."""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass

import numpy as np


def linear_to_srgb(c) -> np.ndarray:
    """image_io.cpp:16-20 (clamped to [0, 1] first)."""
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)


def srgb_to_linear(c) -> np.ndarray:
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, np.power((c + 0.055) / 1.055, 2.4))


def _quantize(unit: np.ndarray) -> np.ndarray:
    """int(v * 255 + 0.5) clamped to [0, 255]."""
    return np.clip((np.asarray(unit, np.float64) * 255.0 + 0.5).astype(np.int64), 0, 255).astype(np.uint8)


def encode_8bit(pixels: np.ndarray) -> np.ndarray:
    """(H, W, 3|4) linear -> uint8: RGB sRGB-encoded, alpha linear."""
    px = np.asarray(pixels, np.float32)
    out = np.empty(px.shape, np.uint8)
    out[..., :3] = _quantize(linear_to_srgb(px[..., :3]))
    if px.shape[-1] == 4:
        out[..., 3] = _quantize(np.clip(px[..., 3], 0.0, 1.0))
    return out


def _png_bytes(img: np.ndarray) -> bytes:
    h, w, c = img.shape
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))  # filter type 0 per row

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6 if c == 4 else 2, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")


def write_image(path: str, pixels: np.ndarray) -> None:
    """writeImage: (H, W, 3|4) linear-light pixels; ".ppm" -> binary P6
    (RGB only), else PNG."""
    img = encode_8bit(pixels)
    if img.shape[-1] not in (3, 4):
        raise ValueError("channels must be 3 or 4")
    if str(path).lower().endswith(".ppm"):
        h, w = img.shape[:2]
        with open(path, "wb") as f:
            f.write(f"P6\n{w} {h}\n255\n".encode())
            f.write(np.ascontiguousarray(img[..., :3]).tobytes())
    else:
        with open(path, "wb") as f:
            f.write(_png_bytes(np.ascontiguousarray(img)))


def read_ppm(path: str) -> np.ndarray:
    """readPpm: binary P6 with maxval 255 (comments allowed in the header) ->
    (H, W, 3) linear float32."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def token() -> bytes:
        nonlocal pos
        while True:
            while pos < len(data) and data[pos:pos + 1].isspace():
                pos += 1
            if data[pos:pos + 1] == b"#":
                while pos < len(data) and data[pos:pos + 1] not in (b"\n", b"\r"):
                    pos += 1
                continue
            break
        start = pos
        while pos < len(data) and not data[pos:pos + 1].isspace():
            pos += 1
        return data[start:pos]

    if token() != b"P6":
        raise ValueError(f"not a binary PPM: {path}")
    w, h, maxval = int(token()), int(token()), int(token())
    pos += 1  # the single whitespace after the header
    if w <= 0 or h <= 0 or maxval != 255:
        raise ValueError(f"unsupported PPM header: {path}")
    px = np.frombuffer(data, np.uint8, count=w * h * 3, offset=pos).reshape(h, w, 3)
    return srgb_to_linear(px / 255.0).astype(np.float32)


@dataclass
class CompareResult:
    psnr: float = 0.0
    ssim: float = 0.0


def psnr_srgb(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR (dB) over sRGB-encoded RGB; 99 for identical images."""
    d = linear_to_srgb(np.asarray(a)[..., :3]) - linear_to_srgb(np.asarray(b)[..., :3])
    mse = float(np.mean(d * d))
    return 99.0 if mse <= 1e-12 else 10.0 * np.log10(1.0 / mse)


def compare_images(a: np.ndarray, b: np.ndarray) -> CompareResult:
    """compareImages: PSNR and SSIM of two (H, W, >=3) linear images ({0, 0}
    when the sizes differ)."""
    if a.shape[:2] != b.shape[:2] or a.shape[0] <= 0:
        return CompareResult()
    r = CompareResult(psnr=psnr_srgb(a, b))
    la = linear_to_srgb(a[..., :3]) @ np.array([0.299, 0.587, 0.114])
    lb = linear_to_srgb(b[..., :3]) @ np.array([0.299, 0.587, 0.114])
    h, w = la.shape
    c1, c2, win, step = 0.0001, 0.0009, 8, 4
    total, count = 0.0, 0
    for y in range(0, h - win + 1, step):
        for x in range(0, w - win + 1, step):
            pa, pb = la[y:y + win, x:x + win], lb[y:y + win, x:x + win]
            ma, mb = pa.mean(), pb.mean()
            va, vb = ((pa - ma) ** 2).mean(), ((pb - mb) ** 2).mean()
            cov = ((pa - ma) * (pb - mb)).mean()
            total += ((2 * ma * mb + c1) * (2 * cov + c2)) / ((ma * ma + mb * mb + c1) * (va + vb + c2))
            count += 1
    r.ssim = total / count if count else 1.0
    return r
