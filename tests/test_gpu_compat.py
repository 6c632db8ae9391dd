"""GPU parity of the drop-in entry points (cudaReconst.h) against the CPU oracle.

CudaReconstHost / CudaReconstHostFFT run the HIP filter (CFft-replica LDS FFT)
and the tiled LUT16 integer back-projection; with the same operation order as
the oracle they must agree bit for bit (integer image).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import _lib, phantom, reconst

pytestmark = pytest.mark.gpu


def _case(W, P, seed, c_off=0.37, mu=0.02, zoom_interp=2):
    c_det = W / 2 + c_off
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(seed, count=8, mu_range=(0.0, mu))
    sino = phantom.sinogram(ell, W, fdeg, c_det, scale=W)
    rows = phantom.to_int16(sino).numpy()
    return rows, fdeg, c_det


def _oracle_rint(rows, fdeg, center_det, iInterp, formula=1, rule=1, ifilter=0, tilt=0.0, icoff=0,
                 g_filter=None, kind=0):
    """Reference integer reconstruction of the selected rows (oracle)."""
    g = O.geometry(rows.shape[1], iInterp, center_det, float_center=True)
    gf = g if g_filter is None else g_filter
    filt = O.set_filter(g.ndim, ifilter, 0.5, 6.0)
    px = O.filter_rows(rows, gf, filt, kind=kind)
    img = np.zeros((g.N, g.N), np.int32)
    for i in range(rows.shape[0]):
        lut = O.lut16_i32(px[i], formula)
        c, s, o = O.bp_params(fdeg[i], tilt, g, icoff)
        O.bp_lut_i32(img, lut, g.N, c, s, o, rule)
    return img


@pytest.mark.parametrize("W,P,interp,c_off,formula,rule", [
    (64, 45, 2, 0.37, 1, 1),
    (100, 60, 2, -3.2, 0, 0),        # non power of two, off-centre, CPU formula, float mask
    (128, 90, 2, 0.0, 1, 0),
    (64, 30, 3, 0.25, 1, 1),         # x2 zoom
    (48, 40, 4, 0.6, 0, 1),          # x4 zoom
    (256, 120, 2, 7.4, 1, 1),
])
def test_reconst_host_bit_exact(W, P, interp, c_off, formula, rule):
    print(f"case W={W} P={P} interp={interp} c_off={c_off} formula={formula} rule={rule}", flush=True)
    rows, fdeg, c_det = _case(W, P, 1000 + W + P, c_off)
    _lib.lib().RecviewSetCompatOptions(formula, rule)
    try:
        res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, iInterpolation=interp, normalise=False)
    finally:
        _lib.lib().RecviewSetCompatOptions(1, 1)
    ref = _oracle_rint(rows, fdeg, c_det, interp, formula, rule)
    assert res["nSinogr"] == P
    assert np.array_equal(res["image"], ref), f"max diff {np.abs(res['image'].astype(np.int64) - ref).max()}"


def test_reconst_host_pinned_image_written_directly():
    """A pinned image buffer (CUDA_MALLOC_HOST_INT, as RecView allocates iReconst for
    the CUDA path) is written by the back-projection itself: the same integer image
    as with a pageable buffer and as the oracle, also when the buffer held stale
    values, and on a second call with a kept sinogram."""
    L = _lib.lib()
    W, P = 200, 90
    rows, fdeg, c_det = _case(W, P, 4321, 1.3)
    p = C.POINTER(C.c_int)()
    assert L.CUDA_MALLOC_HOST_INT(C.byref(p), W * W * 4) == 0
    try:
        pinned = np.ctypeslib.as_array(p, shape=(W, W))
        pinned[:] = 12345
        a = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, normalise=False, image=pinned)
        b = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, normalise=False)
        ref = _oracle_rint(rows, fdeg, c_det, 2)
        assert np.array_equal(pinned, ref) and np.array_equal(a["image"], ref)
        assert np.array_equal(b["image"], ref)
        assert a["nSinogr"] == P
    finally:
        assert L.CUDA_FREE_HOST(p) == 0


def test_reconst_host_fft_variant_bit_exact():
    """CudaReconstHostFFT: zero padding, float zoom, placement n/2 (cudaReconst.cu:123-192)."""
    rows, fdeg, c_det = _case(96, 50, 31, 0.2)
    for interp in (2, 3):
        info = reconst.ReconstInfo()
        filt = reconst.set_filter(reconst.ndim_for(96, interp - 2), 0)
        img, cnt = reconst.fill_info(info, rows, fdeg, c_det - 0.5, interp, filt)
        _lib.lib().CudaReconstHostFFT(C.byref(info.ri), 0, False)
        assert info.ri.iStatus != reconst.RECONST_INFO_ERROR
        ref = _oracle_rint(rows, fdeg, c_det, interp, formula=1, rule=1, kind=2)
        info.free()
        assert np.array_equal(img, ref)


def test_kept_sinogram_centre_sweep():
    """RQFLAGS_SINOGRAMKEPT: a second call with an integer centre shift reuses the
    filtered rows and only re-back-projects with iCenterOffset (cudaReconstHost.cpp:658, 775)."""
    rows, fdeg, c_det = _case(80, 40, 7, 0.3)
    info = reconst.ReconstInfo()
    res1 = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, normalise=False, info=info)
    res2 = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5 + 2.0, normalise=False, info=info,
                                    flags=reconst.RQFLAGS_SINOGRAMKEPT)
    g0 = O.geometry(80, 2, c_det, float_center=True)
    ref1 = _oracle_rint(rows, fdeg, c_det, 2)
    ref2 = _oracle_rint(rows, fdeg, c_det + 2.0, 2, icoff=2, g_filter=g0)
    info.free()
    assert np.array_equal(res1["image"], ref1)
    assert np.array_equal(res2["image"], ref2)


def test_abort_callback_and_progress():
    """DBProjDlgCtrl is called every 20 projections; returning true stops before that
    projection (cudaReconstHost.cpp:715-717)."""
    rows, fdeg, c_det = _case(64, 100, 9)
    calls = []

    def cb(ri, step, isino, pcur):
        calls.append(isino)
        return isino >= 40

    fn = _lib.DLGCTRL(cb)
    _lib.lib().RecviewSetDlgCtrl(C.cast(fn, C.c_void_p))
    try:
        res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, report=True, normalise=False)
    finally:
        _lib.lib().RecviewSetDlgCtrl(None)
    assert calls == [0, 20, 40]
    assert res["nSinogr"] == 40
    ref = _oracle_rint(rows[:40], fdeg[:40], c_det, 2)
    assert np.array_equal(res["image"], ref)


def test_progress_segments_and_late_abort():
    """With progress reporting the back-projection runs in segments between the
    callback points while the callbacks are made: 400 projections give 20 callbacks,
    the image of a run without abort equals the unsegmented one bit for bit, and an
    abort at position 260 keeps exactly the 260 projections before it (skipped
    frames included in the positions, not in the count)."""
    rows, fdeg, c_det = _case(64, 400, 19)
    binc = np.ones(401, np.int8)
    binc[[5, 77, 261]] = 0
    plain = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, binc=binc, normalise=False)
    for stop in (None, 260):
        calls = []

        def cb(ri, step, isino, pcur):
            calls.append(isino)
            return stop is not None and isino >= stop

        fn = _lib.DLGCTRL(cb)
        _lib.lib().RecviewSetDlgCtrl(C.cast(fn, C.c_void_p))
        try:
            res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, binc=binc, report=True, normalise=False)
        finally:
            _lib.lib().RecviewSetDlgCtrl(None)
        if stop is None:
            assert calls == list(range(0, 400, 20))
            assert res["nSinogr"] == plain["nSinogr"] == 397
            assert np.array_equal(res["image"], plain["image"])
        else:
            assert calls == list(range(0, 280, 20))
            keep = [i for i in range(260) if binc[i]]
            assert res["nSinogr"] == len(keep) == 258
            assert np.array_equal(res["image"], _oracle_rint(rows[keep], fdeg[keep], c_det, 2))


def test_frame_selection_and_streams():
    """bInc skips and the even-only flag (cudaReconstHost.cpp:718-729); the stream
    and null-stream paths give identical images."""
    rows, fdeg, c_det = _case(72, 60, 11)
    binc = np.ones(61, np.int8)
    binc[[3, 4, 17, 50]] = 0
    a = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, binc=binc,
                                 flags=reconst.RQFLAGS_USEONLYEVENFRAMES, normalise=False, stream=True)
    b = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, binc=binc,
                                 flags=reconst.RQFLAGS_USEONLYEVENFRAMES, normalise=False, stream=False)
    keep = [i for i in range(60) if binc[i] and i % 2 == 0]
    ref = _oracle_rint(rows[keep], fdeg[keep], c_det, 2)
    assert a["nSinogr"] == len(keep)
    assert np.array_equal(a["image"], ref) and np.array_equal(b["image"], ref)


def test_output_stage_matches_oracle():
    rows, fdeg, c_det = _case(64, 60, 13)
    res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5)
    ref, imin, imax = O.output_stage(res["image"], 1, c_det - 0.5)
    assert np.array_equal(res["normalised"], ref)


def test_unsupported_size_sets_error():
    """A width too small for CFft (n < 8, cfft.cpp:375) sets RECONST_INFO_ERROR and
    logs 28808; transforms above 2^20 points are refused by rvPlanCreate, longer
    than the LDS transform (two passes) accepted."""
    from recview_amd import fbp
    _lib.lib().RecviewClearErrors()
    rows = np.zeros((4, 3), np.int16)
    with pytest.raises(RuntimeError):
        reconst.deconv_back_proj(rows, phantom.angles_deg(4), 1.0, normalise=False)
    assert 28808 in _lib.errors()
    with pytest.raises(RuntimeError, match="FFT range"):
        fbp.FbpPlan([fbp.SliceSpec(1100000, phantom.angles_deg(4), 550000.0)])
    plan = fbp.FbpPlan([fbp.SliceSpec(17000, phantom.angles_deg(4), 8500.0)])
    assert plan.info(0)["n_fft"] == 32768
    plan.close()


def test_kernel_wrappers():
    """CudaSinoPx2igpStream, CudaBackProjStream and CudaDeconv (cudaReconst.cu:158-246)."""
    L = _lib.lib()
    rows, fdeg, c_det = _case(64, 3, 17)
    g = O.geometry(64, 2, c_det, float_center=True)
    filt = O.set_filter(g.ndim, 0)
    px = O.filter_rows(rows, g, filt, kind=0)
    d_px = torch.from_numpy(px[0].copy()).cuda()
    d_igp = torch.zeros(16 * g.N, dtype=torch.int32, device="cuda")
    L.CudaSinoPx2igpStream(g.N, C.c_void_p(d_igp.data_ptr()), C.c_void_p(d_px.data_ptr()), None)
    torch.cuda.synchronize()
    assert np.array_equal(d_igp.cpu().numpy(), O.lut16_i32(px[0], 1))
    d_img = torch.zeros(g.N, g.N, dtype=torch.int32, device="cuda")
    ref = np.zeros((g.N, g.N), np.int32)
    for i in range(2):
        lut = O.lut16_i32(px[i], 1)
        d_igp.copy_(torch.from_numpy(lut))
        th = float(np.float32(fdeg[i])) * 0.0174532925199
        L.CudaBackProjStream(g.N, C.c_float(c_det), 0, 1, th, C.c_void_p(d_img.data_ptr()),
                             C.c_void_p(d_igp.data_ptr()), None)
        c, s, o = O.bp_params(fdeg[i], 0.0, g)
        O.bp_lut_i32(ref, lut, g.N, c, s, o, 1)
    torch.cuda.synchronize()
    assert np.array_equal(d_img.cpu().numpy(), ref)
    # CudaDeconv: CUDAFFT filter of one strip straight to the LUT
    d_strip = torch.from_numpy(rows[1].copy()).cuda()
    d_filt = torch.from_numpy(filt).cuda()
    d_p = torch.zeros(2 * g.ndim, dtype=torch.float32, device="cuda")
    L.CudaDeconv(64, 1, g.ndim, C.c_float(c_det), C.c_void_p(d_filt.data_ptr()),
                 C.c_void_p(d_strip.data_ptr()), C.c_void_p(d_igp.data_ptr()),
                 C.c_void_p(d_p.data_ptr()), None)
    torch.cuda.synchronize()
    pz = O.filter_rows(rows[1:2], g, filt, kind=2)
    assert np.array_equal(d_igp.cpu().numpy(), O.lut16_i32(pz[0], 1))


def test_device_queries():
    L = _lib.lib()
    assert L.GetCudaDeviceCount(300) >= 1
    buf = C.create_string_buffer(256)
    assert L.GetCudaDeviceName(0, buf, 256) == 0 and len(buf.value) > 0
    assert L.GetCudaWarpSize(0) == 64
    cores, procs = C.c_int(), C.c_int()
    assert L.GetCudaNumberOfCores(0, C.byref(cores), C.byref(procs)) == 0
    assert cores.value == 64 and procs.value >= 1
    p = C.POINTER(C.c_int)()
    assert L.CUDA_MALLOC_HOST_INT(C.byref(p), 4096) == 0
    p[0] = 5
    assert L.CUDA_FREE_HOST(p) == 0


def test_oversized_buffers_set_error():
    """Buffers beyond the reference's 32-bit sizes (cudaReconstHost.cpp:39) are
    refused like a failed allocation: RECONST_INFO_ERROR and code 28801, before
    any row is read (every projection pointer below aims at one host row)."""
    import ctypes as C
    from recview_amd import _lib, reconst
    W, P = 2048, 600_000                      # P x N floats > 4 GiB
    row = np.zeros(W, np.int16)
    ptrs = (C.POINTER(C.c_short) * P)()
    p0 = row.ctypes.data_as(C.POINTER(C.c_short))
    for i in range(P):
        ptrs[i] = p0
    info = reconst.ReconstInfo()
    ri = info.ri
    ri.iSinogr = C.cast(ptrs, C.POINTER(C.POINTER(C.c_short)))
    ri.maxSinogrLen = P
    ri.iStatus = 1
    ri.iStepSino, ri.iStartSino = 1, 0
    ri.ixdim, ri.iInterpolation = W, 2
    ri.center = W / 2 + 0.5
    ri.iLenSinogr = P + 1
    img = np.zeros((W, W), np.int32)
    cnt = np.zeros(1, np.int32)
    binc = np.ones(P + 1, np.int8)
    fdeg = np.linspace(0, 180, P, endpoint=False).astype(np.float32)
    filt = reconst.set_filter(4096, 0)
    ri.iReconst = img.ctypes.data_as(C.POINTER(C.c_int))
    ri.nSinogr = cnt.ctypes.data_as(C.POINTER(C.c_int))
    ri.bInc = binc.ctypes.data_as(C.POINTER(C.c_char))
    ri.fdeg = fdeg.ctypes.data_as(C.POINTER(C.c_float))
    ri.fFilter = filt.ctypes.data_as(C.POINTER(C.c_float))
    ri.iMultiplex = 1
    _lib.lib().RecviewClearErrors()
    _lib.lib().CudaReconstHost(C.byref(ri), 0, False, True)
    assert ri.iStatus == 2 and 28801 in _lib.errors()
    info.free()
