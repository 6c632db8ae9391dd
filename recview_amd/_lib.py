"""ctypes binding of librecview_hip.so (the C ABI declared in include/recview_hip.h).

The library is built in-tree by ``recview_amd.build`` (hipcc, gfx950).  There is
no CPU fallback: if the shared object is missing, ``lib()`` raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librecview_hip.so")
_LIB = None


class ReconstInfo(C.Structure):
    """RECONST_INFO (reconstinfo.h:21-101), field for field."""
    _fields_ = [
        ("iStatus", C.c_ubyte), ("ixdim", C.c_int), ("iInterpolation", C.c_int),
        ("center", C.c_double), ("iSinoCenter", C.c_int), ("iLenSinogr", C.c_int),
        ("iReconst", C.POINTER(C.c_int)), ("iStartSino", C.c_int), ("iStepSino", C.c_int),
        ("pDoc", C.POINTER(C.c_uint)), ("nSinogr", C.POINTER(C.c_int)),
        ("dataName", C.c_char * 61), ("bInc", C.POINTER(C.c_char)),
        ("fdeg", C.POINTER(C.c_float)), ("fexp", C.POINTER(C.c_float)),
        ("iMultiplex", C.c_int), ("iOffset", C.c_int), ("maxSinogrLen", C.c_int),
        ("iSinogr", C.POINTER(C.POINTER(C.c_short))), ("fFilter", C.POINTER(C.c_float)),
        ("dReconFlags", C.c_uint), ("fTiltAngle", C.c_float), ("drStart", C.c_int),
        ("drEnd", C.c_int), ("drX", C.c_double), ("drY", C.c_double), ("drOmit", C.c_int),
        ("threadID", C.c_uint), ("hThread", C.c_size_t),
        ("max_d_ifp", C.c_uint), ("max_d_igp", C.c_uint),
        ("d_ifp", C.c_void_p), ("d_igp", C.c_void_p), ("d_filt", C.c_void_p),
        ("max_d_filt", C.c_uint), ("d_strip", C.c_void_p), ("max_d_strip", C.c_uint),
        ("d_p", C.c_void_p), ("max_d_p", C.c_uint), ("d_px", C.c_void_p), ("max_d_px", C.c_uint),
        ("fftplan", C.c_void_p), ("ifftdim", C.c_uint),
        ("stream1", C.c_void_p), ("stream2", C.c_void_p),
        ("i64result", C.c_uint64), ("i64sum", C.c_uint64),
        ("ppRef", C.c_void_p), ("ppQry", C.c_void_p), ("uiMaxRef", C.c_uint),
        ("uiMaxQry", C.c_uint), ("pcxyz1", C.c_void_p), ("pcxyz2", C.c_void_p),
        ("iLsqBin", C.c_int), ("iLsqBxref", C.c_int), ("piDrift", C.POINTER(C.c_int)),
        ("iFlag", C.c_int),
    ]


class RvSliceSpec(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("P", C.c_int), ("iInterpolation", C.c_int), ("center", C.c_double),
        ("tilt", C.c_float), ("fdeg", C.POINTER(C.c_float)), ("filter", C.c_int),
        ("cutoff", C.c_double), ("order", C.c_double), ("sino_dtype", C.c_int),
        ("pad_mode", C.c_int), ("out_n", C.c_int), ("pix_zoom", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("tile", C.c_int), ("offset_ct", C.c_int),
        ("binc", C.c_char_p), ("chunk", C.c_int), ("rows_of", C.c_int),
    ]


class RvOptions(C.Structure):
    _fields_ = [
        ("interp", C.c_int), ("store", C.c_int), ("lut_formula", C.c_int),
        ("index_rule", C.c_int), ("persistent", C.c_int), ("tile", C.c_int),
        ("angles_per_chunk", C.c_int), ("lds_half", C.c_int), ("normalize", C.c_int),
        ("exact_filter", C.c_int), ("slices_per_item", C.c_int),
    ]


DLGCTRL = C.CFUNCTYPE(C.c_bool, C.POINTER(ReconstInfo), C.c_int, C.c_int, C.POINTER(C.c_int))

# Every symbol include/recview_hip.h declares (checked by the CPU tests).
EXPORTS = [
    "CudaReconstHost", "CudaReconstMemFree", "CudaReconstResourceFree", "CudaReconstHostFFT",
    "GetCudaDeviceCount", "GetCudaMaxThreadsPerBlock", "GetCudaWarpSize", "GetCudaDeviceName",
    "GetCudaDeviceComputingCapability", "GetCudaNumberOfCores", "GetCudaClockRate",
    "CudaSinogramHost", "CudaDeconv", "CudaBackProjStream", "CudaSinoPx2igpStream",
    "CUDA_MALLOC_HOST_INT", "CUDA_FREE_HOST", "blocksize", "RecviewSetDlgCtrl",
    "RecviewGetErrors", "RecviewClearErrors", "RecviewLastMessage", "RecviewSetCompatOptions",
    "rvPlanCreate", "rvPlanExecute", "rvPlanExecuteGraph", "rvPlanFilter", "rvPlanBackproject",
    "rvPlanPack", "rvPlanBackprojectPacked",
    "rvPlanFilteredRows", "rvPlanSetFilteredRows", "rvPlanInfo", "rvPlanRows", "rvPlanDestroy",
    "RecviewSetFilter", "RecviewGeometry", "RecviewOffsetCTPairs", "RecviewOffsetCTStitch",
    "rvUnetCreate", "rvUnetParamCount", "rvUnetGetParams", "rvUnetSetParams", "rvUnetForward",
    "rvUnetLossGrad", "rvUnetTrainStep", "rvUnetDestroy", "rvConv3x3Bf16", "rvConv3x3WgradBf16",
    "RecviewOutputStage", "RecviewImageDescription", "RecviewWriteTif16", "RecviewSinogramDevice",
    "RecviewDeviceAlloc", "RecviewDeviceFree", "RecviewCopyToHost", "RecviewCopyToDevice",
    "RecviewOutputStageNoAlloc",
]


def _declare(L):
    vp, ip, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)
    rip = C.POINTER(ReconstInfo)
    L.CudaReconstHost.argtypes = [rip, C.c_int, C.c_bool, C.c_bool]
    L.CudaReconstHost.restype = None
    L.CudaReconstHostFFT.argtypes = [rip, C.c_int, C.c_bool]
    L.CudaReconstHostFFT.restype = None
    L.CudaReconstMemFree.argtypes = [rip]
    L.CudaReconstMemFree.restype = None
    L.CudaReconstResourceFree.argtypes = [rip, C.c_bool]
    L.CudaReconstResourceFree.restype = None
    L.CudaSinogramHost.argtypes = [rip, C.c_int, C.c_bool]
    L.CudaSinogramHost.restype = None
    L.GetCudaDeviceCount.argtypes = [C.c_int]
    L.GetCudaMaxThreadsPerBlock.argtypes = [C.c_int]
    L.GetCudaWarpSize.argtypes = [C.c_int]
    L.GetCudaDeviceName.argtypes = [C.c_int, C.c_char_p, C.c_int]
    L.GetCudaDeviceComputingCapability.argtypes = [C.c_int, ip, ip]
    L.GetCudaNumberOfCores.argtypes = [C.c_int, ip, ip]
    L.GetCudaClockRate.argtypes = [C.c_int, ip, ip]
    L.CudaDeconv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, vp, vp, vp, vp, vp]
    L.CudaDeconv.restype = None
    L.CudaBackProjStream.argtypes = [C.c_int, C.c_float, C.c_int, C.c_int, C.c_double, vp, vp, vp]
    L.CudaBackProjStream.restype = None
    L.CudaSinoPx2igpStream.argtypes = [C.c_int, vp, vp, vp]
    L.CudaSinoPx2igpStream.restype = None
    L.CUDA_MALLOC_HOST_INT.argtypes = [C.POINTER(C.POINTER(C.c_int)), C.c_size_t]
    L.CUDA_FREE_HOST.argtypes = [vp]
    L.RecviewSetDlgCtrl.argtypes = [vp]
    L.RecviewSetDlgCtrl.restype = None
    L.RecviewGetErrors.argtypes = [ip, C.c_int]
    L.RecviewClearErrors.restype = None
    L.RecviewLastMessage.restype = C.c_char_p
    L.RecviewSetCompatOptions.argtypes = [C.c_int, C.c_int]
    L.RecviewSetCompatOptions.restype = None
    L.rvPlanCreate.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(RvSliceSpec), C.c_int,
                               C.POINTER(RvOptions)]
    L.rvPlanExecute.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), vp]
    L.rvPlanExecuteGraph.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), vp]
    L.rvPlanFilter.argtypes = [vp, C.POINTER(vp), vp]
    L.rvPlanBackproject.argtypes = [vp, C.POINTER(vp), vp]
    L.rvPlanPack.argtypes = [vp, vp]
    L.rvPlanBackprojectPacked.argtypes = [vp, C.POINTER(vp), vp]
    L.rvPlanFilteredRows.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), ip]
    L.rvPlanSetFilteredRows.argtypes = [vp, C.c_int, vp, vp]
    L.rvPlanInfo.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip]
    L.rvPlanRows.argtypes = [vp, C.c_int, ip, ip]
    L.rvPlanDestroy.argtypes = [vp]
    L.rvPlanDestroy.restype = None
    L.RecviewSetFilter.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, fp]
    L.RecviewSetFilter.restype = None
    L.RecviewGeometry.argtypes = [C.c_int, C.c_int, ip, ip]
    L.RecviewOffsetCTPairs.argtypes = [fp, C.c_char_p, C.c_int, ip]
    L.RecviewOffsetCTStitch.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_double, C.c_double,
                                        C.c_int, C.c_int, C.c_int, vp, vp]
    L.RecviewSinogramDevice.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_char_p, fp, fp, C.c_uint, C.c_int,
                                        C.c_int, C.c_double, C.c_double, C.c_int, ip, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, vp, vp]
    L.RecviewOutputStage.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), vp]
    L.RecviewOutputStageNoAlloc.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp]
    L.RecviewImageDescription.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_int]
    L.RecviewWriteTif16.argtypes = [C.c_char_p, ip, C.c_int, C.c_int, C.c_char_p, C.c_char_p]
    L.rvUnetCreate.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_ulonglong]
    L.rvUnetParamCount.argtypes = [vp]
    L.rvUnetParamCount.restype = C.c_longlong
    L.rvUnetGetParams.argtypes = [vp, fp, C.c_longlong]
    L.rvUnetSetParams.argtypes = [vp, fp, C.c_longlong]
    L.rvUnetForward.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, vp]
    L.rvUnetLossGrad.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, fp, fp, vp]
    L.rvUnetTrainStep.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, fp, vp]
    L.rvUnetDestroy.argtypes = [vp]
    L.rvUnetDestroy.restype = None
    L.rvConv3x3Bf16.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    L.rvConv3x3WgradBf16.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, vp,
                                     C.c_int, C.c_int, C.c_int, vp]


def lib():
    """Loads librecview_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "librecview_hip.so is not built; run __graft_entry__.build() "
                "(hipcc --offload-arch=gfx950) first -- there is no CPU fallback")
        _LIB = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _declare(_LIB)
    return _LIB


def last_message():
    m = lib().RecviewLastMessage()
    return m.decode() if m else ""


def errors():
    buf = (C.c_int * 1000)()
    n = lib().RecviewGetErrors(buf, 1000)
    return list(buf[:n])


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_message()}")
