"""Native batched FBP (Part 2 of the C ABI): plans over many slices on one GPU.

torch supplies device memory and streams only; every operation on the data
runs in librecview_hip.so's kernels.
"""
import ctypes as C

import numpy as np

from . import _lib

INTERP = {"linear": 0, "lut16_f32": 1, "lut16_i32": 2}
STORE = {"f32": 0, "f16": 1, "bf16": 2}
SINO = {"i16": 0, "f32": 1}
PAD = {"replicate": 0, "zero": 1}
FILTERS = {"hann": 0, "hamming": 1, "ramp": 2, "parzen": 3, "butterworth": 4, "shepp-logan": 5}


def _stream_handle(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _ptrs(tensors, host_ok=False):
    """Device pointers of contiguous device tensors (or, with host_ok, of pinned
    host tensors: pinned memory is mapped into the GPU's address space, so a kernel
    writes it across the host link)."""
    arr = (C.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        if not t.is_contiguous() or not (t.is_cuda or (host_ok and t.is_pinned())):
            raise ValueError("tensors must be contiguous device tensors" +
                             (" or pinned host tensors" if host_ok else ""))
        arr[i] = t.data_ptr()
    return arr


class SliceSpec:
    """One slice: W detector samples x P projections (angles fdeg, degrees)."""

    def __init__(self, W, fdeg, center, iInterpolation=2, tilt=0.0, filter="hann", cutoff=0.5,
                 order=6.0, sino_dtype="f32", pad="replicate", out_n=0, pix_zoom=1.0, x0=0.0,
                 y0=0.0, tile=-1, offset_ct=False, binc=None, chunk=0, rows_from=-1):
        """center follows RECONST_INFO.center (user centre + 0.5); offset_ct: the
        rows are a 360-degree scan (angles over 0..360) with an off-centre axis,
        stitched to double width on the GPU; tile = -1 uses the plan's tile;
        chunk = 0 the plan's angles per LDS chunk (see autotune.apply); rows_from =
        index of an earlier slice whose filtered rows this one back-projects (its
        integer centre offset applied as the reference's kept-sinogram path does)."""
        self.tile = int(tile)
        self.chunk = int(chunk)
        self.rows_from = int(rows_from)
        self.offset_ct = bool(offset_ct)
        self.binc = None if binc is None else np.ascontiguousarray(binc, np.int8)
        self.W = int(W)
        self.fdeg = np.ascontiguousarray(fdeg, np.float32)
        self.P = len(self.fdeg)
        self.center = float(center)
        self.iInterpolation = int(iInterpolation)
        self.tilt = float(tilt)
        self.filter = FILTERS[filter] if isinstance(filter, str) else int(filter)
        self.cutoff = float(cutoff)
        self.order = float(order)
        self.sino_dtype = SINO[sino_dtype]
        self.pad = PAD[pad]
        self.out_n = int(out_n)
        self.pix_zoom = float(pix_zoom)
        self.x0 = float(x0)
        self.y0 = float(y0)

    def to_c(self):
        s = _lib.RvSliceSpec()
        s.W, s.P, s.iInterpolation, s.center, s.tilt = self.W, self.P, self.iInterpolation, self.center, self.tilt
        s.fdeg = self.fdeg.ctypes.data_as(C.POINTER(C.c_float))
        s.filter, s.cutoff, s.order = self.filter, self.cutoff, self.order
        s.sino_dtype, s.pad_mode = self.sino_dtype, self.pad
        s.out_n, s.pix_zoom, s.x0, s.y0 = self.out_n, self.pix_zoom, self.x0, self.y0
        s.tile, s.offset_ct = self.tile, int(self.offset_ct)
        s.binc = None if self.binc is None else self.binc.ctypes.data_as(C.c_char_p)
        s.chunk = self.chunk
        s.rows_of = self.rows_from + 1
        return s


class FbpPlan:
    """A reconstruction plan for a batch of slices (see rvPlanCreate)."""

    def __init__(self, slices, interp="linear", store="f32", lut_formula=1, index_rule=1,
                 persistent=False, tile=0, angles_per_chunk=0, lds_half=False, normalize=False,
                 exact_filter=False, slices_per_item=0, device=0):
        """normalize: images in attenuation per detector sample (x 2 pi / projections)
        instead of raw back-projected sums.  exact_filter: LINEAR modes keep CFft's
        operation order (bit-identical filtered rows) instead of two rows per transform.
        slices_per_item: LINEAR modes back-project slices of identical geometry
        together (0: four per work item with FP32 values in LDS -- the same results
        as one per item; a class's leftovers as a pair and a single -- and with FP16
        pairs (lds_half) sixteen on the 32x64 and 32x32 tiles, eight on 64x64; or 1,
        2, 4, 8, 16, 32, where 8 and up need FP16 pairs, 16 the 32x32 or 32x64
        tile, and 32 selects the matrix-core
        kernel for every 32 slices of one geometry; with FP16 pairs and more than
        two per item, the pairs and singles a class leaves over run the FP32-pair
        kernels on the 16-bit rows)."""
        self.slices = list(slices)
        L = _lib.lib()
        specs = (_lib.RvSliceSpec * len(self.slices))(*[s.to_c() for s in self.slices])
        opt = _lib.RvOptions(INTERP[interp], STORE[store], int(lut_formula), int(index_rule),
                             int(persistent), int(tile), int(angles_per_chunk), int(lds_half),
                             int(normalize), int(exact_filter), int(slices_per_item))
        self.interp = interp
        self.store = store
        h = C.c_void_p()
        _lib.check(L.rvPlanCreate(C.byref(h), int(device), specs, len(self.slices), C.byref(opt)),
                   "rvPlanCreate")
        self._h = h

    def info(self, s):
        v = [C.c_int() for _ in range(5)]
        _lib.check(_lib.lib().rvPlanInfo(self._h, s, *[C.byref(x) for x in v]), "rvPlanInfo")
        out_n, n_fft, tiles, tw, th = (x.value for x in v)
        return {"out_n": out_n, "n_fft": n_fft, "tiles": tiles, "tile_w": tw, "tile_h": th}

    def out_n(self, s):
        return self.info(s)["out_n"]

    def rows(self, s):
        """(rows, width) slice s reconstructs, after offset-CT stitching."""
        n, w = C.c_int(), C.c_int()
        _lib.check(_lib.lib().rvPlanRows(self._h, s, C.byref(n), C.byref(w)), "rvPlanRows")
        return n.value, w.value

    def _count(self, tensors, what, host_ok=False):
        if len(tensors) != len(self.slices):
            raise ValueError(f"{what}: {len(tensors)} tensors for a plan of {len(self.slices)} slices")
        return _ptrs(tensors, host_ok)

    def execute(self, sinos, outs, stream=None, graph=False):
        """Reconstructs every slice.  Images may be device tensors or pinned host
        tensors (the back-projection then writes each finished tile across the host
        link, so the download overlaps the kernel)."""
        fn = _lib.lib().rvPlanExecuteGraph if graph else _lib.lib().rvPlanExecute
        _lib.check(fn(self._h, self._count(sinos, "sinograms"), self._count(outs, "images", True),
                      _stream_handle(stream)), "rvPlanExecute")

    def filter(self, sinos, stream=None):
        _lib.check(_lib.lib().rvPlanFilter(self._h, self._count(sinos, "sinograms"), _stream_handle(stream)),
                   "rvPlanFilter")

    def backproject(self, outs, stream=None):
        _lib.check(_lib.lib().rvPlanBackproject(self._h, self._count(outs, "images"), _stream_handle(stream)),
                   "rvPlanBackproject")

    def pack(self, stream=None):
        """The first half of backproject(): grouped slices' filtered rows into the
        form their kernels stage."""
        _lib.check(_lib.lib().rvPlanPack(self._h, _stream_handle(stream)), "rvPlanPack")

    def backproject_packed(self, outs, stream=None):
        """The second half of backproject(): the back-projection kernels alone, on
        the rows the last pack() (or backproject()) wrote."""
        _lib.check(_lib.lib().rvPlanBackprojectPacked(self._h, self._count(outs, "images"), _stream_handle(stream)),
                   "rvPlanBackprojectPacked")

    def filtered_rows(self, s):
        rows, scales, n = C.c_void_p(), C.c_void_p(), C.c_int()
        _lib.check(_lib.lib().rvPlanFilteredRows(self._h, s, C.byref(rows), C.byref(scales), C.byref(n)),
                   "rvPlanFilteredRows")
        return rows.value, scales.value, n.value

    def filtered_rows_tensor(self, s):
        """Copy of slice s's filtered rows as a float32 [P, N] device tensor."""
        import torch
        rows, scales, n = self.filtered_rows(s)
        P = self.rows(s)[0]
        if self.store == "f32":
            src = _wrap_device(rows, (P, n), torch.float32)
            return src.clone()
        dt = torch.float16 if self.store == "f16" else torch.bfloat16
        src = _wrap_device(rows, (P, n), dt)
        sc = _wrap_device(scales, (P,), torch.float32)
        return src.float() * sc[:, None]

    def set_filtered_rows(self, s, rows, stream=None):
        _lib.check(_lib.lib().rvPlanSetFilteredRows(self._h, s, C.c_void_p(rows.data_ptr()),
                                                    _stream_handle(stream)), "rvPlanSetFilteredRows")

    def close(self):
        if self._h:
            _lib.lib().rvPlanDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceArray:
    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def _wrap_device(ptr, shape, dtype):
    """A torch view of library-owned device memory (valid while the plan lives)."""
    import torch
    n = int(np.prod(shape))
    esz = torch.empty((), dtype=dtype).element_size()
    t = torch.as_tensor(_DeviceArray(ptr, n, "<f4" if esz == 4 else "<i2"), device="cuda")
    return t.view(dtype).view(shape)


def reconstruct(sino, fdeg, center, out=None, stream=None, **kw):
    """One-slice convenience wrapper: float32 or int16 device sinogram [P, W] -> image."""
    import torch
    dtype = "i16" if sino.dtype == torch.int16 else "f32"
    plan_kw = {k: kw.pop(k) for k in list(kw) if k in ("interp", "store", "lut_formula", "index_rule",
                                                       "persistent", "tile", "angles_per_chunk",
                                                       "lds_half", "normalize", "exact_filter")}
    spec = SliceSpec(sino.shape[1], fdeg, center, sino_dtype=dtype, **kw)
    plan = FbpPlan([spec], **plan_kw)
    n = plan.out_n(0)
    if out is None:
        odt = torch.int32 if plan.interp == "lut16_i32" else torch.float32
        out = torch.empty(n, n, device=sino.device, dtype=odt)
    plan.execute([sino.contiguous()], [out], stream=stream)
    torch.cuda.synchronize()
    plan.close()
    return out


def centre_sweep(sino, fdeg, centres, stream=None, **kw):
    """Axis-scan images: one sinogram reconstructed at every centre in `centres`
    (RECONST_INFO.center convention) in one persistent launch, the batched form of
    the reference's kept-sinogram centre sweep (gazoDoc.cpp:1994-2025,
    cudaReconstHost.cpp:775-777): the sinogram is filtered once, at the centre of
    median integer part, and every other centre back-projects those rows shifted by
    its integer centre offset.  Returns a [K, N, N] float32 device tensor."""
    import torch
    dtype = "i16" if sino.dtype == torch.int16 else "f32"
    plan_kw = {k: kw.pop(k) for k in list(kw) if k in ("interp", "store", "persistent", "tile",
                                                       "angles_per_chunk", "lds_half", "normalize",
                                                       "exact_filter")}
    plan_kw.setdefault("persistent", True)
    centres = [float(c) for c in centres]
    order = sorted(range(len(centres)), key=lambda k: int(centres[k]))
    src = order[len(order) // 2]                    # the rows are filtered at this centre
    seq = [src] + [k for k in range(len(centres)) if k != src]
    specs = [SliceSpec(sino.shape[1], fdeg, centres[k], sino_dtype=dtype, rows_from=(0 if j else -1), **kw)
             for j, k in enumerate(seq)]
    plan = FbpPlan(specs, **plan_kw)
    n = plan.out_n(0)
    out = torch.empty(len(specs), n, n, device=sino.device)
    rows = sino.contiguous()
    plan.execute([rows] * len(specs), [out[k] for k in seq], stream=stream)
    torch.cuda.synchronize()
    plan.close()
    return out
