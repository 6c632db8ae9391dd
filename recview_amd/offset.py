"""Offset-CT stitching (360-degree scans with the axis off the detector centre)
through librecview_hip.so: the reference's pairing (gazoDoc.cpp:2967-2978) on the
host, the stitch itself (gazoDoc.cpp:3536-3665) as a device kernel."""
import ctypes as C

import numpy as np

from . import _lib


def pairs(fdeg, binc=None):
    """Partner frame of every stitched row (-1 for unsampled rows)."""
    fd = np.ascontiguousarray(fdeg, np.float32)
    n = len(fd)
    b = np.ones(n, np.int8) if binc is None else np.ascontiguousarray(binc, np.int8)
    out = np.empty(n, np.int32)
    rc = _lib.lib().RecviewOffsetCTPairs(fd.ctypes.data_as(C.POINTER(C.c_float)),
                                         b.ctypes.data_as(C.c_char_p), n,
                                         out.ctypes.data_as(C.POINTER(C.c_int)))
    if rc < 0:
        raise ValueError(f"offset-CT pairing failed ({-rc})")
    return out[:rc]


def stitch(sino, fdeg, center_user, binc=None, binning=1, trim=0, deltac=0.0, out=None, stream=None):
    """Stitch a device sinogram [frames][W] (int16 or float32) of a 360-degree scan.

    center_user is the user centre (the reference adds HALFPIXEL_OFFSET);
    returns (stitched device rows [nout][(2W - trim)/binning], their angles)."""
    import torch
    pr = pairs(fdeg, binc)
    nout = len(pr)
    W = sino.shape[1] * binning             # raw frame width (rows are binned)
    width = (2 * W - trim) // binning
    if out is None:
        out = torch.empty(nout, width, dtype=sino.dtype, device=sino.device)
    d_pair = torch.from_numpy(pr).to(sino.device)
    st = torch.cuda.current_stream() if stream is None else stream
    rc = _lib.lib().RecviewOffsetCTStitch(C.c_void_p(sino.data_ptr()), int(sino.dtype == torch.float32), W,
                                          C.c_void_p(d_pair.data_ptr()), nout, float(center_user + 0.5),
                                          float(deltac), int(binning), int(trim), width,
                                          C.c_void_p(out.data_ptr()), C.c_void_p(st.cuda_stream))
    _lib.check(rc, "RecviewOffsetCTStitch")
    st.synchronize()
    return out, np.ascontiguousarray(np.asarray(fdeg, np.float32)[:nout])


def stitched_center(center_user, raw_width):
    """DeconvBackProj's centre remap for offset CT (gazoDoc.cpp:3832-3837)."""
    return raw_width + center_user if center_user <= raw_width / 2.0 else center_user
