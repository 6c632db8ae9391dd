"""Host-side mirror of RecView's reconstruction caller over the drop-in C ABI.

``deconv_back_proj`` plays the role of CGazoDoc::DeconvBackProj
(gazoDoc.cpp:3819-4027) for one device: it fills a RECONST_INFO exactly as the
reference does (gazoDoc.cpp:3929-3951), calls CudaReconstHost (the MI355X
engine), then applies the caller-side scale, min/max normalisation and circle
mask.  ``ReconstInfo`` keeps the host arrays the struct points to alive.
"""
import ctypes as C
import math
import time

import numpy as np

from . import _lib

LOG2 = 0.69314718055994530941723212145818     # constants.h:8
LOG_SCALE = 3000                               # constants.h:39
BACKPROJ_SCALE = 100                           # constants.h:41
ZERNIKE_SCALE = 1000                           # constants.h:40
HALFPIXEL_OFFSET = 0.5                         # constants.h:55
RECONST_INFO_IDLE, RECONST_INFO_BUSY, RECONST_INFO_ERROR = 0, 1, 2
RQFLAGS_ZERNIKE = 2
RQFLAGS_USEONLYEVENFRAMES = 16
RQFLAGS_USEONLYODDFRAMES = 32
RQFLAGS_SINOGRAMKEPT = 128
CGAZODOC_BINC_SAMPLE = 1


def zoom_and_binning(iInterpolation):
    """gazoDoc.cpp:3822-3826: (zoom exponent, binning, zoom factor)."""
    izoom = iInterpolation - 2 if iInterpolation > 2 else 0
    ibin = 4 if iInterpolation == 0 else (2 if iInterpolation == 1 else 1)
    return izoom, ibin, 2 ** izoom


def ndim_for(ixlen, izoom):
    """gazoDoc.cpp:3827-3828 / cudaReconstHost.cpp:638-639."""
    return 2 ** (int(math.log(ixlen) / LOG2) + 1 + izoom)


def set_filter(ndim, ifilter=0, cutoff=0.5, order=6.0):
    """CGazoDoc::SetFilter (gazoDoc.cpp:3747-3812), computed by the library."""
    out = np.empty(ndim, np.float32)
    _lib.lib().RecviewSetFilter(int(ndim), int(ifilter), float(cutoff), float(order),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


class ReconstInfo:
    """A RECONST_INFO plus the host buffers it references."""

    def __init__(self):
        self.ri = _lib.ReconstInfo()
        self._keep = []

    def set_rows(self, rows_i16):
        rows = np.ascontiguousarray(rows_i16, np.int16)
        self._rows = rows
        ptrs = (C.POINTER(C.c_short) * rows.shape[0])()
        for i in range(rows.shape[0]):
            ptrs[i] = rows[i].ctypes.data_as(C.POINTER(C.c_short))
        self._ptrs = ptrs
        self.ri.iSinogr = C.cast(ptrs, C.POINTER(C.POINTER(C.c_short)))
        self.ri.maxSinogrLen = rows.shape[0]

    def keep(self, arr):
        self._keep.append(arr)
        return arr

    def free(self, device_reset=False):
        _lib.lib().CudaReconstResourceFree(C.byref(self.ri), bool(device_reset))


def fill_info(info, rows_i16, fdeg, center, iInterpolation=2, filt=None, tilt=0.0, binc=None,
              flags=0, ixdim=None, n_start=0, n_step=1, nsino_counter=None, image=None):
    """RECONST_INFO setup of gazoDoc.cpp:3929-3951 for one worker."""
    ri = info.ri
    P = rows_i16.shape[0]
    izoom, ibin, z = zoom_and_binning(iInterpolation)
    W = rows_i16.shape[1] if ixdim is None else ixdim
    N = W * z
    ri.iStatus = RECONST_INFO_BUSY
    ri.iStepSino = n_step
    ri.iStartSino = n_start
    ri.ixdim = W
    ri.iInterpolation = iInterpolation
    ri.center = (center + HALFPIXEL_OFFSET) / ibin
    ri.iLenSinogr = P + 1
    img = info.keep(np.zeros((N, N), np.int32) if image is None else image)
    ri.iReconst = img.ctypes.data_as(C.POINTER(C.c_int))
    cnt = info.keep(np.zeros(1, np.int32) if nsino_counter is None else nsino_counter)
    ri.nSinogr = cnt.ctypes.data_as(C.POINTER(C.c_int))
    b = info.keep(np.full(P + 1, CGAZODOC_BINC_SAMPLE, np.int8) if binc is None
                  else np.ascontiguousarray(binc, np.int8))
    ri.bInc = b.ctypes.data_as(C.POINTER(C.c_char))
    fd = info.keep(np.ascontiguousarray(fdeg, np.float32))
    ri.fdeg = fd.ctypes.data_as(C.POINTER(C.c_float))
    ri.fTiltAngle = tilt
    ri.iMultiplex = 1
    ri.iOffset = 0
    info.set_rows(rows_i16)
    f = info.keep(np.ascontiguousarray(filt, np.float32))
    ri.fFilter = f.ctypes.data_as(C.POINTER(C.c_float))
    ri.dReconFlags = flags
    return img, cnt


def output_stage(img, nsino, center, iInterpolation=2, pixel_width_um=0.5, zernike=False):
    """Caller-side scale, min/max normalisation and circle mask (gazoDoc.cpp:3993-4021).

    Returns (uint16-range int32 image, pixelBase, pixelDiv)."""
    izoom, ibin, ipintp = zoom_and_binning(iInterpolation)
    N = img.shape[0]
    imax = max(int(img.max()), 0)
    imin = int(img.min())
    if imin == imax:
        imax += 1
    scale_factor = 6.2831853 * 10000. / (pixel_width_um * ibin)
    denom = (ZERNIKE_SCALE if zernike else LOG_SCALE) * BACKPROJ_SCALE * nsino
    sc = scale_factor / denom
    pixel_base = np.float32(sc * ipintp * imin)
    pixel_div = np.float32((65535 - 1) / sc / (imax - imin) / ipintp)
    r = ipintp * center / ibin
    irad = int(r) if r < (N - r) else int(N - r)
    irad2 = irad * irad
    h = N // 2
    ii = (np.arange(N) - h)[:, None] ** 2
    jj = (np.arange(N) - h)[None, :] ** 2
    inside = (ii + jj) <= irad2
    norm = ((img.astype(np.int64) - imin).astype(np.float32) * np.float32(65534)) / np.float32(imax - imin)
    out = np.where(inside, norm.astype(np.int32), 0).astype(np.int32)
    return out, float(pixel_base), float(pixel_div)


def deconv_back_proj(rows_i16, fdeg, center, iInterpolation=2, ifilter=0, cutoff=0.5, order=6.0,
                     tilt=0.0, binc=None, flags=0, device=0, report=False, stream=True,
                     pixel_width_um=0.5, info=None, normalise=True, image=None):
    """CGazoDoc::DeconvBackProj for one device (nCPU = 1).

    rows_i16 : int16 sinogram rows [P][W] (x LOG_SCALE), already binned.
    center   : user centre (the reference adds HALFPIXEL_OFFSET and divides by binning).
    image    : optional int32 [N][N] output buffer (RecView's iReconst; a pinned one,
               from CUDA_MALLOC_HOST_INT, is written by the GPU directly).
    Returns a dict with the raw int32 image, nSinogr, seconds, and (if
    normalise) the 16-bit-range image with pixelBase / pixelDiv."""
    rows = np.ascontiguousarray(rows_i16, np.int16)
    izoom, ibin, z = zoom_and_binning(iInterpolation)
    ndim = ndim_for(rows.shape[1], izoom)
    filt = set_filter(ndim, ifilter, cutoff, order)
    own = info is None
    info = ReconstInfo() if own else info
    img, cnt = fill_info(info, rows, fdeg, center, iInterpolation, filt, tilt, binc, flags, image=image)
    t0 = time.perf_counter()
    _lib.lib().CudaReconstHost(C.byref(info.ri), int(device), bool(report), bool(stream))
    t1 = time.perf_counter()
    if info.ri.iStatus == RECONST_INFO_ERROR:
        errs = _lib.errors()
        if own:
            info.free()
        raise RuntimeError(f"CudaReconstHost failed: codes {errs[-3:]} {_lib.last_message()}")
    info.ri.iStatus = RECONST_INFO_IDLE
    res = {"image": img.copy(), "nSinogr": int(cnt[0]), "seconds": t1 - t0, "info": info}
    if normalise:
        out, base, div = output_stage(img, int(cnt[0]), center, iInterpolation, pixel_width_um,
                                      bool(flags & RQFLAGS_ZERNIKE))
        res.update(normalised=out, pixelBase=base, pixelDiv=div)
    if own:
        info.free()
        res["info"] = None
    return res
