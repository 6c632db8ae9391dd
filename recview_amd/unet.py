"""Learned restoration of sparse-view slices: the native residual U-Net.

Thin ctypes layer over Part 3 of the C ABI (rvUnet*, rvConv3x3*).  torch
supplies device memory and streams; every layer runs in librecview_hip.so's
kernels (BF16 MFMA convolutions with FP32 accumulation).
"""
import ctypes as C

import numpy as np

from . import _lib

DEFAULT_BASE = 32
DEFAULT_LEVELS = 4


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev(t, dtype_name):
    import torch
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("tensors must be contiguous device tensors")
    if str(t.dtype) != dtype_name:
        raise TypeError(f"expected {dtype_name}, got {t.dtype}")
    return C.c_void_p(t.data_ptr())


def param_layout(base=DEFAULT_BASE, levels=DEFAULT_LEVELS):
    """[(name, cin, cout, w_offset, b_offset)] in parameter-vector order
    (weights [cout][3][3][cin] then biases [cout] per layer)."""
    Cl = lambda l: base << l
    specs = [("enc0a_in", 1, Cl(0)), ("enc0b", Cl(0), Cl(0))]
    for l in range(1, levels):
        specs += [(f"enc{l}a", Cl(l - 1), Cl(l)), (f"enc{l}b", Cl(l), Cl(l))]
    for l in range(levels - 2, -1, -1):
        specs += [(f"dec{l}a", Cl(l + 1) + Cl(l), Cl(l)), (f"dec{l}b", Cl(l), Cl(l))]
    specs.append(("out", Cl(0), 1))
    out, off = [], 0
    for name, cin, cout in specs:
        w = off
        off += 9 * cin * cout
        out.append((name, cin, cout, w, off))
        off += cout
    return out


def flops_per_pixel(base=DEFAULT_BASE, levels=DEFAULT_LEVELS):
    """Multiply-add FLOPs of one forward pass per input pixel (2 per MAC, all
    3x3 convolutions; level l runs on 4^-l of the pixels)."""
    total = 0.0
    for name, cin, cout, _, _ in param_layout(base, levels):
        lev = 0 if name == "out" else int(name[3])
        total += 2.0 * 9 * cin * cout / 4 ** lev
    return total


class Unet:
    """Residual U-Net on one GPU: out = x + net(x * scale) / scale."""

    def __init__(self, base=DEFAULT_BASE, levels=DEFAULT_LEVELS, seed=0, device=0):
        self.base, self.levels, self.device = int(base), int(levels), int(device)
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.rvUnetCreate(C.byref(h), self.device, self.base, self.levels, C.c_ulonglong(seed)),
                   "rvUnetCreate")
        self._h = h
        self.n = L.rvUnetParamCount(h)
        lay = param_layout(self.base, self.levels)
        assert lay[-1][4] + lay[-1][2] == self.n, "parameter layout disagrees with the library"

    def get_params(self):
        p = np.empty(self.n, np.float32)
        _lib.check(_lib.lib().rvUnetGetParams(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), self.n),
                   "rvUnetGetParams")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters, got {p.size}")
        _lib.check(_lib.lib().rvUnetSetParams(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), self.n),
                   "rvUnetSetParams")

    @staticmethod
    def _bhw(x):
        if x.dim() == 2:
            return 1, x.shape[0], x.shape[1]
        if x.dim() == 3:
            return tuple(x.shape)
        raise ValueError("images must be [H, W] or [B, H, W]")

    def forward(self, x, scale, out=None, stream=None):
        import torch
        B, H, W = self._bhw(x)
        if out is None:
            out = torch.empty_like(x)
        _lib.check(_lib.lib().rvUnetForward(self._h, _dev(x, "torch.float32"), _dev(out, "torch.float32"),
                                            B, H, W, C.c_float(scale), _stream(stream)), "rvUnetForward")
        return out

    def loss_grad(self, x, target, scale, stream=None):
        B, H, W = self._bhw(x)
        g = np.empty(self.n, np.float32)
        loss = C.c_float()
        _lib.check(_lib.lib().rvUnetLossGrad(self._h, _dev(x, "torch.float32"), _dev(target, "torch.float32"),
                                             B, H, W, C.c_float(scale), C.byref(loss),
                                             g.ctypes.data_as(C.POINTER(C.c_float)), _stream(stream)),
                   "rvUnetLossGrad")
        return loss.value, g

    def train_step(self, x, target, scale, lr, want_loss=True, stream=None):
        B, H, W = self._bhw(x)
        loss = C.c_float()
        _lib.check(_lib.lib().rvUnetTrainStep(self._h, _dev(x, "torch.float32"),
                                              _dev(target, "torch.float32"), B, H, W, C.c_float(scale),
                                              C.c_float(lr), C.byref(loss) if want_loss else None,
                                              _stream(stream)), "rvUnetTrainStep")
        return loss.value if want_loss else None

    def save(self, path, scale, meta=None, bf16=True):
        """Parameters (as bf16 bit patterns, or float32) plus scale and shape."""
        p = self.get_params()
        if bf16:
            u = p.view(np.uint32).astype(np.uint64)
            stored = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
        else:
            stored = p
        extra = {k: np.asarray(v) for k, v in (meta or {}).items()}
        np.savez(path, params=stored, base=self.base, levels=self.levels, scale=np.float64(scale), **extra)

    @staticmethod
    def read(path):
        """(params float32, base, levels, scale, extra dict) from a file written by save()."""
        with np.load(path, allow_pickle=False) as z:
            p = z["params"]
            if p.dtype == np.uint16:
                p = (p.astype(np.uint32) << 16).view(np.float32)
            extra = {k: z[k] for k in z.files if k not in ("params", "base", "levels", "scale")}
            return p.astype(np.float32), int(z["base"]), int(z["levels"]), float(z["scale"]), extra

    @classmethod
    def load(cls, path, device=0):
        p, base, levels, scale, extra = cls.read(path)
        net = cls(base, levels, device=device)
        net.set_params(p)
        net.scale = scale
        net.meta = extra
        return net

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().rvUnetDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def conv3x3(x, w, b, cin, cout, relu=True, in_stride=None, in_off=0, out=None, out_stride=None,
            out_off=0, mask=None, mask_stride=None, mask_off=0, data_gradient=False, shape=None,
            stream=None):
    """One 3x3 convolution on NHWC bf16 tensors (int16 storage of bf16 bits).

    Forward: x has cin channels (pixel pitch in_stride), out cout channels.
    data_gradient: x holds dY (cout channels), out receives dX (cin channels).
    w is float32 [cout, 3, 3, cin] on the device, b float32 [cout] (or None).
    """
    import torch
    Bn, H, W = shape
    kin, kout = (cout, cin) if data_gradient else (cin, cout)
    in_stride = in_stride or kin
    if out is None:
        out_stride = kout
        out = torch.zeros(Bn, H, W, kout, dtype=torch.int16, device=x.device)
    out_stride = out_stride or kout
    _lib.check(_lib.lib().rvConv3x3Bf16(
        C.c_void_p(x.data_ptr()), in_stride, in_off, cin, C.c_void_p(w.data_ptr()),
        C.c_void_p(b.data_ptr()) if b is not None else None, C.c_void_p(out.data_ptr()), out_stride,
        out_off, cout, C.c_void_p(mask.data_ptr()) if mask is not None else None,
        mask_stride or 0, mask_off, int(relu), int(data_gradient), Bn, H, W, _stream(stream)),
        "rvConv3x3Bf16")
    return out


def conv3x3_wgrad(dy, x, cin, cout, shape, dy_stride=None, dy_off=0, x_stride=None, x_off=0,
                  stream=None):
    """(dW float32 [cout, 3, 3, cin], db float32 [cout]) on the device."""
    import torch
    Bn, H, W = shape
    dw = torch.empty(cout, 3, 3, cin, dtype=torch.float32, device=dy.device)
    db = torch.empty(cout, dtype=torch.float32, device=dy.device)
    _lib.check(_lib.lib().rvConv3x3WgradBf16(
        C.c_void_p(dy.data_ptr()), dy_stride or cout, dy_off, cout, C.c_void_p(x.data_ptr()),
        x_stride or cin, x_off, cin, C.c_void_p(dw.data_ptr()), C.c_void_p(db.data_ptr()), Bn, H, W,
        _stream(stream)), "rvConv3x3WgradBf16")
    return dw, db


def to_bf16_bits(t):
    """float tensor -> int16 tensor holding the bfloat16 bits (round to nearest even)."""
    import torch
    return t.to(torch.bfloat16).view(torch.int16)


def from_bf16_bits(t):
    import torch
    return t.view(torch.bfloat16).float()
