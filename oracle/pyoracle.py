"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as the
thing measured or shipped.  The product path (``recview_amd``) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Geom(C.Structure):
    _fields_ = [("ixdim", C.c_int), ("izoom", C.c_int), ("z", C.c_int), ("ndimp", C.c_int),
                ("ndim", C.c_int), ("N", C.c_int), ("h", C.c_int), ("ic", C.c_int),
                ("fcenter16", C.c_float), ("ihoffset", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def _p(t):
    return C.POINTER(t)


def _declare(L):
    fp, ip, dp, sp = _p(C.c_float), _p(C.c_int32), _p(C.c_double), _p(C.c_int16)
    gp = _p(Geom)
    L.orc_geometry.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, gp]
    L.orc_set_filter.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, fp]
    L.orc_cfft_twiddles.argtypes = [C.c_int, fp, fp]
    L.orc_fft1rev.argtypes = [C.c_int, fp, fp, fp, fp, fp]
    L.orc_fft1.argtypes = [C.c_int, fp, fp, fp, fp, fp]
    L.orc_row_prep_i16.argtypes = [sp, gp, fp]
    L.orc_row_prep_f32.argtypes = [fp, gp, fp]
    L.orc_row_prep_cudafft.argtypes = [sp, gp, fp]
    L.orc_filter_rows.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, gp, fp, fp]
    L.orc_lut16_i32.argtypes = [fp, C.c_int, C.c_int, ip]
    L.orc_lut16_f32.argtypes = [fp, C.c_int, fp]
    L.orc_bp_params.argtypes = [C.c_float, C.c_float, gp, C.c_int, fp, fp, fp]
    L.orc_bp_lut_i32.argtypes = [ip, ip, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
                                 C.c_int, C.c_int]
    L.orc_bp_lut_f64.argtypes = [dp, fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int]
    L.orc_recon_rint.argtypes = [sp, C.c_int, fp, C.c_float, gp, fp, C.c_int, C.c_int, ip]
    L.orc_recon_rf32.argtypes = [fp, C.c_int, fp, C.c_float, gp, fp, C.c_int, dp]
    L.orc_fft_f64.argtypes = [C.c_int, dp, dp, C.c_int]
    L.orc_filter_rows_f64.argtypes = [fp, C.c_int, C.c_int, gp, fp, dp]
    L.orc_recon_of64.argtypes = [fp, C.c_int, fp, C.c_float, gp, fp, dp]
    L.orc_bp_rows_linear_f64.argtypes = [fp, C.c_int, fp, C.c_float, gp, C.c_int, dp]
    L.orc_sinogram.argtypes = [_p(C.c_uint16), C.c_int, C.c_int, C.c_int, C.c_char_p, fp, fp, C.c_uint,
                               C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, ip, C.c_int, C.c_int,
                               C.c_int, C.c_int]
    L.orc_bp_points_linear_f64.argtypes = [dp, C.c_int, fp, C.c_float, gp, C.c_int, dp, dp, C.c_int, dp]
    L.orc_output_stage.argtypes = [ip, C.c_int, C.c_int, C.c_double, C.c_int, ip, ip]
    L.orc_offset_stitch.argtypes = [sp, C.c_int, C.c_int, fp, C.c_char_p, C.c_double, C.c_double,
                                    C.c_int, C.c_int, C.c_int, sp]
    L.orc_cpu_isa.restype = C.c_int
    L.orc_cpu_baseline.argtypes = [sp, C.c_int, fp, C.c_float, gp, fp, C.c_int, C.c_int, ip]
    L.orc_cpu_baseline.restype = C.c_double


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def geometry(ixdim, iInterpolation=2, center=None, float_center=True):
    if center is None:
        center = ixdim / 2.0
    g = Geom()
    rc = lib().orc_geometry(int(ixdim), int(iInterpolation), float(center), int(float_center), C.byref(g))
    if rc != 0:
        raise ValueError("detector width too small for CFft (ndimp < 3)")
    return g


def set_filter(ndim, ifilter=0, cutoff=0.5, order=6.0):
    out = np.empty(ndim, np.float32)
    lib().orc_set_filter(int(ndim), int(ifilter), float(cutoff), float(order), _ptr(out, C.c_float))
    return out


def cfft_twiddles(np_):
    n = 1 << np_
    wr = np.empty(n // 2, np.float32)
    wi = np.empty(n // 2, np.float32)
    lib().orc_cfft_twiddles(np_, _ptr(wr, C.c_float), _ptr(wi, C.c_float))
    return wr, wi


def fft1rev(x):
    """CFft::FFT1Rev of a complex vector (float32 re/im), returns complex128 view."""
    n = len(x)
    np_ = n.bit_length() - 1
    wr, wi = cfft_twiddles(np_)
    re = _f32(np.real(x)).copy()
    im = _f32(np.imag(x)).copy()
    sc = np.empty(4 * n, np.float32)
    lib().orc_fft1rev(np_, _ptr(wr, C.c_float), _ptr(wi, C.c_float), _ptr(re, C.c_float),
                      _ptr(im, C.c_float), _ptr(sc, C.c_float))
    return re.astype(np.float64) + 1j * im.astype(np.float64), (re, im)


def fft1(x):
    n = len(x)
    np_ = n.bit_length() - 1
    wr, wi = cfft_twiddles(np_)
    re = _f32(np.real(x)).copy()
    im = _f32(np.imag(x)).copy()
    sc = np.empty(4 * n, np.float32)
    lib().orc_fft1(np_, _ptr(wr, C.c_float), _ptr(wi, C.c_float), _ptr(re, C.c_float),
                   _ptr(im, C.c_float), _ptr(sc, C.c_float))
    return re.astype(np.float64) + 1j * im.astype(np.float64), (re, im)


def row_prep(strip, g, kind=0):
    out = np.empty(g.ndim, np.float32)
    if kind == 1:
        s = _f32(strip)
        lib().orc_row_prep_f32(_ptr(s, C.c_float), C.byref(g), _ptr(out, C.c_float))
    else:
        s = np.ascontiguousarray(strip, dtype=np.int16)
        fn = lib().orc_row_prep_cudafft if kind == 2 else lib().orc_row_prep_i16
        fn(_ptr(s, C.c_int16), C.byref(g), _ptr(out, C.c_float))
    return out


def filter_rows(sino, g, filt, kind=0):
    """kind 0: int16 host path; 1: float host path; 2: int16 CUDAFFT placement."""
    sino = np.ascontiguousarray(sino, dtype=np.float32 if kind == 1 else np.int16)
    P = sino.shape[0]
    px = np.empty((P, g.N), np.float32)
    filt = _f32(filt)
    lib().orc_filter_rows(sino.ctypes.data_as(C.c_void_p), kind, P, sino.shape[1], C.byref(g),
                          _ptr(filt, C.c_float), _ptr(px, C.c_float))
    return px


def filter_rows_f64(sino, g, filt):
    sino = _f32(sino)
    P = sino.shape[0]
    px = np.empty((P, g.N), np.float64)
    filt = _f32(filt)
    lib().orc_filter_rows_f64(_ptr(sino, C.c_float), P, sino.shape[1], C.byref(g),
                              _ptr(filt, C.c_float), _ptr(px, C.c_double))
    return px


def lut16_i32(px, formula=0):
    px = _f32(px)
    N = len(px)
    out = np.empty(16 * N, np.int32)
    lib().orc_lut16_i32(_ptr(px, C.c_float), N, formula, _ptr(out, C.c_int32))
    return out


def lut16_f32(px):
    px = _f32(px)
    N = len(px)
    out = np.empty(16 * N, np.float32)
    lib().orc_lut16_f32(_ptr(px, C.c_float), N, _ptr(out, C.c_float))
    return out


def bp_params(fdeg, tilt, g, icoff=0):
    c, s, o = C.c_float(), C.c_float(), C.c_float()
    lib().orc_bp_params(float(fdeg), float(tilt), C.byref(g), int(icoff), C.byref(c), C.byref(s),
                        C.byref(o))
    return c.value, s.value, o.value


def bp_lut_i32(img, lut, N, fcos, fsin, foffset, rule, iy0=0, iy1=None):
    iy1 = N if iy1 is None else iy1
    assert img.dtype == np.int32 and img.flags.c_contiguous
    lut = np.ascontiguousarray(lut, np.int32)
    lib().orc_bp_lut_i32(_ptr(img, C.c_int32), _ptr(lut, C.c_int32), N, fcos, fsin, foffset,
                         rule, iy0, iy1)


def recon_rint(sino_i16, fdeg, g, filt, formula=0, rule=0, tilt=0.0):
    sino = np.ascontiguousarray(sino_i16, np.int16)
    fdeg = _f32(fdeg)
    img = np.empty((g.N, g.N), np.int32)
    lib().orc_recon_rint(_ptr(sino, C.c_int16), sino.shape[0], _ptr(fdeg, C.c_float), tilt,
                         C.byref(g), _ptr(_f32(filt), C.c_float), formula, rule, _ptr(img, C.c_int32))
    return img


def recon_rf32(sino, fdeg, g, filt, rule=0, tilt=0.0):
    sino = _f32(sino)
    fdeg = _f32(fdeg)
    img = np.empty((g.N, g.N), np.float64)
    lib().orc_recon_rf32(_ptr(sino, C.c_float), sino.shape[0], _ptr(fdeg, C.c_float), tilt,
                         C.byref(g), _ptr(_f32(filt), C.c_float), rule, _ptr(img, C.c_double))
    return img


def recon_of64(sino, fdeg, g, filt, tilt=0.0):
    sino = _f32(sino)
    fdeg = _f32(fdeg)
    img = np.empty((g.N, g.N), np.float64)
    lib().orc_recon_of64(_ptr(sino, C.c_float), sino.shape[0], _ptr(fdeg, C.c_float), tilt,
                         C.byref(g), _ptr(_f32(filt), C.c_float), _ptr(img, C.c_double))
    return img


def bp_rows_linear_f64(px, fdeg, g, icoff=0, tilt=0.0):
    px = _f32(px)
    fdeg = _f32(fdeg)
    img = np.empty((g.N, g.N), np.float64)
    lib().orc_bp_rows_linear_f64(_ptr(px, C.c_float), px.shape[0], _ptr(fdeg, C.c_float), tilt,
                                 C.byref(g), int(icoff), _ptr(img, C.c_double))
    return img


def bp_points_linear_f64(px, fdeg, g, xs, ys, icoff=0, tilt=0.0):
    """FP64 linear back-projection of FP64 rows px[P][N] at grid points (xs, ys)."""
    px = np.ascontiguousarray(px, np.float64)
    fdeg = _f32(fdeg)
    xs = np.ascontiguousarray(xs, np.float64).ravel()
    ys = np.ascontiguousarray(ys, np.float64).ravel()
    out = np.empty(xs.size, np.float64)
    lib().orc_bp_points_linear_f64(_ptr(px, C.c_double), px.shape[0], _ptr(fdeg, C.c_float), tilt,
                                   C.byref(g), int(icoff), _ptr(xs, C.c_double), _ptr(ys, C.c_double),
                                   xs.size, _ptr(out, C.c_double))
    return out


def sinogram(frames, binc, fexp, fdeg, flags=0, drift=(0, 0, 0.0, 0.0, 0), pidrift=None, is8bit=False,
             start=0, step=1, binning=1):
    """Sinogram generation (general.cpp:2377-2532) on uint16 frames [nframes][mult][W]
    (dark frame last).  Returns (int16 frames after the in-place conversion, status)."""
    f = np.ascontiguousarray(frames, np.uint16).copy()
    nframes, mult, W = f.shape
    binc = np.ascontiguousarray(binc, np.int8)
    fexp, fdeg = _f32(fexp), _f32(fdeg)
    drs, dre, dx, dy, om = drift
    pd = None if pidrift is None else np.ascontiguousarray(pidrift, np.int32)
    st = lib().orc_sinogram(_ptr(f, C.c_uint16), nframes, mult, W, binc.ctypes.data_as(C.c_char_p),
                            _ptr(fexp, C.c_float), _ptr(fdeg, C.c_float), int(flags), int(drs),
                            int(dre), float(dx), float(dy), int(om),
                            None if pd is None else _ptr(pd, C.c_int32), int(is8bit), int(start), int(step),
                            int(binning))
    return f.view(np.int16), st


def fft_f64(x, inverse=False):
    re = np.ascontiguousarray(np.real(x), np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), np.float64).copy()
    lib().orc_fft_f64(len(re), _ptr(re, C.c_double), _ptr(im, C.c_double), int(inverse))
    return re + 1j * im


def output_stage(img, z, center_user, binning=1):
    img = np.ascontiguousarray(img, np.int32).copy()
    imin, imax = C.c_int(), C.c_int()
    lib().orc_output_stage(_ptr(img, C.c_int32), img.shape[0], z, float(center_user), binning,
                           C.byref(imin), C.byref(imax))
    return img, imin.value, imax.value


def cpu_isa():
    return lib().orc_cpu_isa()


def cpu_baseline(sino_i16, fdeg, g, filt, nthreads, isa=-1, tilt=0.0):
    """Times the reference CPU reconstruction of one int16 slice.  Returns (seconds, image)."""
    sino = np.ascontiguousarray(sino_i16, np.int16)
    fdeg = _f32(fdeg)
    img = np.empty((g.N, g.N), np.int32)
    t = lib().orc_cpu_baseline(_ptr(sino, C.c_int16), sino.shape[0], _ptr(fdeg, C.c_float), tilt,
                               C.byref(g), _ptr(_f32(filt), C.c_float), int(nthreads), int(isa),
                               _ptr(img, C.c_int32))
    if t < 0:
        raise MemoryError("cpu baseline allocation failed")
    return t, img


def offset_stitch(sino_i16, fdeg, binc, center_plus_half, bin_=1, trim=0, deltac=0.0, out_width=None):
    """Offset-CT stitch restated from gazoDoc.cpp:3536-3665 (int16 rows, already
    binned: the raw frame width is row width x bin_)."""
    sino = np.ascontiguousarray(sino_i16, np.int16)
    nframes, Wb = sino.shape
    W = Wb * bin_
    fd = _f32(fdeg)
    b = np.ascontiguousarray(binc, np.int8)
    if out_width is None:
        out_width = (2 * W - trim) // bin_
    out = np.zeros((nframes, out_width), np.int16)
    n = lib().orc_offset_stitch(_ptr(sino, C.c_int16), nframes, W, _ptr(fd, C.c_float),
                                b.ctypes.data_as(C.c_char_p), float(center_plus_half), float(deltac),
                                int(bin_), int(trim), int(out_width), _ptr(out, C.c_int16))
    if n < 0:
        raise ValueError(f"offset stitch error {-n}")
    return out[:n]
