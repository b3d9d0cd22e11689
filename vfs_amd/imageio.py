"""File formats and preprocessing of the reference's demo pipeline, without OpenCV.

The reference's test program (test_vfs/main.cpp:103-231) and DLL (dllhook/vfs.cpp:8-50)
read their inputs through OpenCV: `cv::imread(..., IMREAD_GRAYSCALE)` for 8-bit
images and masks, `cv::readOpticalFlow` for the Middlebury `.flo` trajectory and
calibration fields, `cv::equalizeHist` and `cv::resize` for preprocessing, and
`convertTo(CV_16U, 256)` for the 16-bit depth images they write.  This module restates
those steps for the files the reference ships (grayscale, non-interlaced PNGs of 8 or
16 bits; `.flo` fields).  include/stereotgv/vfs_io.h is the same for C++ consumers.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

FLO_TAG = 202021.25  # 'PIEH' read as a little-endian float


def read_flo(path) -> np.ndarray:
    """Middlebury .flo: float tag, int32 width, int32 height, interleaved float32 (u, v)."""
    with open(path, "rb") as f:
        tag = np.fromfile(f, np.float32, 1)
        if tag.size != 1 or tag[0] != np.float32(FLO_TAG):
            raise ValueError(f"{path}: not a .flo file")
        w, h = (int(v) for v in np.fromfile(f, np.int32, 2))
        data = np.fromfile(f, np.float32, w * h * 2)
    if data.size != w * h * 2:
        raise ValueError(f"{path}: truncated")
    return data.reshape(h, w, 2)


def write_flo(path, field: np.ndarray) -> None:
    h, w = field.shape[:2]
    with open(path, "wb") as f:
        np.array([FLO_TAG], np.float32).tofile(f)
        np.array([w, h], np.int32).tofile(f)
        np.ascontiguousarray(field, np.float32).tofile(f)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_png_gray(path) -> np.ndarray:
    """Grayscale PNG (colour type 0, bit depth 8 or 16, not interlaced) as uint8 / uint16."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        pos += 12 + n
    if hdr is None:
        raise ValueError(f"{path}: no IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    if ctype != 0 or depth not in (8, 16) or interlace != 0:
        raise ValueError(f"{path}: only non-interlaced 8/16-bit grayscale PNGs are supported")
    bpp = depth // 8
    raw = zlib.decompress(b"".join(idat))
    stride = w * bpp
    if len(raw) != h * (stride + 1):
        raise ValueError(f"{path}: bad image data size")
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros(stride, np.int64)
            for i in range(stride):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                elif ft == 4:
                    p = _paeth(a, b, c)
                else:
                    raise ValueError(f"{path}: bad filter type {ft}")
                cur[i] = (line[i] + p) & 255
        out[y] = cur
        prev = cur
    if bpp == 1:
        return out
    return (out[:, 0::2].astype(np.uint16) << 8) | out[:, 1::2].astype(np.uint16)


def equalize_hist(img: np.ndarray) -> np.ndarray:
    """Histogram equalisation of an 8-bit image as OpenCV's equalizeHist computes it:
    with i the lowest occupied level, lut[i] = 0 and lut[v] = round((cdf(v) - h(i)) *
    255 / (N - h(i))) above it (float product, rounding half to even); an image of one
    level is returned unchanged."""
    img = np.asarray(img, np.uint8)
    hist = np.bincount(img.ravel(), minlength=256)
    i = int(np.flatnonzero(hist)[0])
    total = img.size
    if hist[i] == total:
        return img.copy()
    scale = np.float32(255.0) / np.float32(total - hist[i])
    lut = np.zeros(256, np.uint8)
    s = np.cumsum(hist[i + 1:]).astype(np.float32)
    lut[i + 1:] = np.clip(np.rint(s * scale), 0, 255).astype(np.uint8)
    return lut[img]


def resize_half(img: np.ndarray) -> np.ndarray:
    """Exact 2x downscale of an 8-bit image with bilinear interpolation at the coarse
    pixel centres (cv::resize, INTER_LINEAR): each output pixel is the rounded mean of
    a 2x2 block, (a + b + c + d + 2) >> 2."""
    h, w = img.shape
    a = np.asarray(img[: h - h % 2, : w - w % 2], np.int32)
    s = a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2]
    return ((s + 2) >> 2).astype(np.uint8)


def to_u16(a: np.ndarray, scale: float) -> np.ndarray:
    """convertTo(CV_16U, scale): round half to even, saturate to [0, 65535]; NaN -> 0."""
    v = np.rint(np.nan_to_num(np.asarray(a, np.float64) * scale, nan=0.0, posinf=65535.0, neginf=0.0))
    return np.clip(v, 0, 65535).astype(np.uint16)
