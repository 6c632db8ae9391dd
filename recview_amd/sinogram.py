"""Sinogram generation through the drop-in entry point CudaSinogramHost.

Mirrors the caller side of GenerateSinogramThread (general.cpp:2377-2532,
set up by CGazoDoc, gazoDoc.cpp:3113-3471): RECONST_INFO carries the frame
rows (maxSinogrLen frames x iMultiplex detector lines, dark frame last), the
bInc flags, exposure times, angles and drift settings; the rows are converted
in place to int16 -ln(I/I0) x 3000 on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib


def generate(frames, binc, fexp, fdeg, flags=0, drift=(0, 0, 0.0, 0.0, 0), pidrift=None, is8bit=False,
             start=0, step=1, iInterpolation=2, device=0):
    """frames: uint16 [nframes][mult][W] (dark last).  Returns (int16 frames, iStatus)."""
    f = np.ascontiguousarray(frames, np.uint16).copy()
    nframes, mult, W = f.shape
    rows = (C.POINTER(C.c_short) * (nframes * mult))()
    for i in range(nframes):
        for k in range(mult):
            rows[i * mult + k] = f[i, k].ctypes.data_as(C.POINTER(C.c_short))
    binc = np.ascontiguousarray(binc, np.int8)
    fexp = np.ascontiguousarray(fexp, np.float32)
    fdeg = np.ascontiguousarray(fdeg, np.float32)
    ri = _lib.ReconstInfo()
    ri.ixdim = W
    ri.iInterpolation = iInterpolation
    ri.maxSinogrLen = nframes
    ri.iMultiplex = mult
    ri.iSinogr = C.cast(rows, C.POINTER(C.POINTER(C.c_short)))
    ri.bInc = binc.ctypes.data_as(C.POINTER(C.c_char))
    ri.fexp = fexp.ctypes.data_as(C.POINTER(C.c_float))
    ri.fdeg = fdeg.ctypes.data_as(C.POINTER(C.c_float))
    ri.dReconFlags = flags
    ri.drStart, ri.drEnd, ri.drX, ri.drY, ri.drOmit = drift
    pd = None
    if pidrift is not None:
        pd = np.ascontiguousarray(pidrift, np.int32)
        ri.piDrift = pd.ctypes.data_as(C.POINTER(C.c_int))
    ri.iFlag = 2 if is8bit else 0
    ri.iStartSino, ri.iStepSino = start, step
    ri.iStatus = 1
    _lib.lib().CudaSinogramHost(C.byref(ri), device, False)
    return f.view(np.int16), int(ri.iStatus)


def generate_device(frames, binc, fexp, fdeg, flags=0, drift=(0, 0, 0.0, 0.0, 0), pidrift=None,
                    is8bit=False, start=0, step=1, binning=1, out_float=True, out=None, stream=None):
    """Device-resident conversion (RecviewSinogramDevice).  frames: uint16 bits in an
    int16 device tensor [nframes][mult][W].  Returns (rows, number converted): rows
    float32 (or int16) [nframes][mult][W]; rows that are not converted stay zero."""
    import torch
    if frames.dtype != torch.int16 or not frames.is_cuda or frames.dim() != 3:
        raise ValueError("frames must be an int16 device tensor [nframes][mult][W]")
    nframes, mult, W = frames.shape
    if out is None:
        out = torch.zeros(nframes, mult, W, dtype=torch.float32 if out_float else torch.int16,
                          device=frames.device)
    binc = np.ascontiguousarray(binc, np.int8)
    fexp = np.ascontiguousarray(fexp, np.float32)
    fdeg = np.ascontiguousarray(fdeg, np.float32)
    pd = None if pidrift is None else np.ascontiguousarray(pidrift, np.int32)
    drs, dre, dx, dy, om = drift
    s = stream if stream is not None else torch.cuda.current_stream()
    n = _lib.lib().RecviewSinogramDevice(
        C.c_void_p(frames.contiguous().data_ptr()), nframes, mult, W, binc.ctypes.data_as(C.c_char_p),
        fexp.ctypes.data_as(C.POINTER(C.c_float)), fdeg.ctypes.data_as(C.POINTER(C.c_float)), int(flags),
        int(drs), int(dre), float(dx), float(dy), int(om),
        None if pd is None else pd.ctypes.data_as(C.POINTER(C.c_int)), int(is8bit), int(start), int(step),
        int(binning), int(out_float), C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream))
    if n == -1:
        raise RuntimeError(f"RecviewSinogramDevice failed: {_lib.last_message()}")
    return out, n
