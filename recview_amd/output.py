"""Output side of a reconstructed slice: GPU normalisation to 16 bits and the
16-bit TIFF files RecView writes (with its ImageDescription metadata).

output_stage  -- RecviewOutputStage (gazoDoc.cpp:3993-4021) on a device image
image_description / write_tif16 -- gazoDoc.cpp:1986-1989, general.cpp:1313-1470
read_tif16    -- reads such files back (little-endian, 16-bit, strips)
"""
import ctypes as C
import struct

import numpy as np

from . import _lib


def output_stage(img, center, iInterpolation=2, stream=None):
    """img: N x N int32 or float32 device tensor.  Returns (uint16 image as an
    int16 device tensor of the same bits, min, max)."""
    import torch
    if img.dim() != 2 or img.shape[0] != img.shape[1] or not img.is_cuda or not img.is_contiguous():
        raise ValueError("img must be a contiguous square device tensor")
    if img.dtype not in (torch.int32, torch.float32):
        raise TypeError("img must be int32 or float32")
    N = img.shape[0]
    out = torch.empty(N, N, dtype=torch.int16, device=img.device)
    lo, hi = C.c_double(), C.c_double()
    s = stream if stream is not None else torch.cuda.current_stream()
    _lib.check(_lib.lib().RecviewOutputStage(C.c_void_p(img.data_ptr()), int(img.dtype == torch.float32), N,
                                             float(center), int(iInterpolation), C.c_void_p(out.data_ptr()),
                                             C.byref(lo), C.byref(hi), C.c_void_p(s.cuda_stream)),
               "RecviewOutputStage")
    return out, lo.value, hi.value


def image_description(num_sinogr, pixel_width, pixel_div, pixel_base, center, ifilter):
    buf = C.create_string_buffer(512)
    n = _lib.lib().RecviewImageDescription(buf, 512, int(num_sinogr), float(pixel_width), float(pixel_div),
                                           float(pixel_base), float(center), int(ifilter))
    if n < 0:
        raise RuntimeError("RecviewImageDescription failed")
    return buf.value.decode()


def write_tif16(path, img, desc=None, artist=None):
    """img: host int32 (or integer) array H x W; negative values are written as 0."""
    a = np.ascontiguousarray(img, np.int32)
    rc = _lib.lib().RecviewWriteTif16(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_int)), a.shape[0],
                                      a.shape[1], None if desc is None else desc.encode(),
                                      None if artist is None else artist.encode())
    if rc != 0:
        raise OSError(f"RecviewWriteTif16 failed ({rc})")


def read_tif16(path):
    """(uint16 image, ImageDescription, Artist) of a little-endian 16-bit TIFF."""
    data = open(path, "rb").read()
    if data[:4] != b"II*\x00":
        raise ValueError("not a little-endian TIFF")
    (ifd,) = struct.unpack_from("<I", data, 4)
    (n,) = struct.unpack_from("<H", data, ifd)
    tags = {}
    for i in range(n):
        tag, typ, cnt, val = struct.unpack_from("<HHII", data, ifd + 2 + 12 * i)
        if typ == 3 and cnt == 1:
            val &= 0xFFFF
        tags[tag] = (typ, cnt, val)
    W, H = tags[0x0100][2], tags[0x0101][2]
    if tags[0x0102][2] != 16:
        raise ValueError("not a 16-bit image")

    def text(tag):
        if tag not in tags:
            return None
        _, cnt, off = tags[tag]
        return data[off:off + cnt].split(b"\x00")[0].decode()

    _, ns, so = tags[0x0111]
    _, _, sc = tags[0x0117]
    offs = struct.unpack_from(f"<{ns}I", data, so) if ns > 1 else (so,)
    cnts = struct.unpack_from(f"<{ns}I", data, sc) if ns > 1 else (sc,)
    raw = b"".join(data[o:o + c] for o, c in zip(offs, cnts))
    img = np.frombuffer(raw, "<u2").reshape(H, W).copy()
    return img, text(0x010E), text(0x013B)
