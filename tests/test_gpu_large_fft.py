"""Transforms longer than the LDS holds (n = 32768 and 65536): RecView's x4 / x8
zoom of 2048- and 4096-sample detectors (cudaReconstHost.cpp:636-641 accepts any
ndim).  The filter runs CFft's radix-2 network in two passes per transform
through HBM work rows; every butterfly is the one-pass network's, so:

* filtered rows are bit-identical to the oracle's CFft restatement (FP32 rows,
  int16 and float inputs, the CUDAFFT placement), and the 16-bit stores equal
  those rows scaled by the per-row power of two and rounded to nearest even;
* the drop-in CudaReconstHost at 4096 samples x4 (a 16384^2 image) equals the
  reference integer path R-int bit for bit;
* a LINEAR plan at 2048 samples x8 matches FP64 filtering + FP64 linear
  back-projection on sampled rows: in FP32 no worse than the reference's own
  FP32 filter back-projected in FP64 (its twiddles rounded from double, the
  plan's rows are more accurate), within 1e-4 of the range with the FP16 store;
* CudaDeconv's single-strip filter at n = 32768 gives the reference LUT.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import _lib, fbp, phantom, reconst

pytestmark = pytest.mark.gpu


def _rows(W, P, seed, interp, int16=False):
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(seed, count=10, mu_range=(0.0, 2.0 / W))
    s = phantom.sinogram(ell, W, fdeg, c_det, scale=W).float()
    if int16:
        s = phantom.to_int16(s.double())
    return s, fdeg, c_det


def _lp_expected(rows32, store):
    """Per-row power-of-two scale with |v| / scale in [2^13, 2^14), then FP16 or BF16
    rounded to nearest even (the 16-bit store of the filter)."""
    mx = np.abs(rows32).max(axis=1).astype(np.float32)
    m, e = np.frexp(mx / np.float32(16384.0))
    e = np.where(mx > 0, e, 0)
    sc = np.ldexp(np.float32(1.0), e).astype(np.float32)
    v = (rows32 * (np.float32(1.0) / sc)[:, None]).astype(np.float32)
    if store == "f16":
        q = v.astype(np.float16).astype(np.float32)
    else:
        u = v.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        q = u.astype(np.uint32).view(np.float32)
    return q * sc[:, None], sc


@pytest.mark.parametrize("W,interp,int16", [(4096, 4, False), (2048, 5, True), (4096, 5, False)])
def test_two_pass_filter_rows_bit_identical(W, interp, int16):
    P = 9
    s, fdeg, c_det = _rows(W, P, 4000 + W + interp, interp, int16)
    g = O.geometry(W, interp, c_det)
    assert g.ndim in (32768, 65536)
    print(f"case two-pass filter W={W} n={g.ndim} int16={int16}", flush=True)
    spec = fbp.SliceSpec(W, fdeg, c_det, iInterpolation=interp, filter="hamming",
                         sino_dtype="i16" if int16 else "f32")
    filt = O.set_filter(g.ndim, 1)
    ref = O.filter_rows(s.numpy(), g, filt, kind=0 if int16 else 1)
    for store in ("f32", "f16", "bf16"):
        plan = fbp.FbpPlan([spec], store=store, exact_filter=True)
        assert plan.info(0)["n_fft"] == g.ndim
        plan.filter([s.cuda().contiguous()])
        torch.cuda.synchronize()
        got = plan.filtered_rows_tensor(0).cpu().numpy()
        plan.close()
        if store == "f32":
            assert np.array_equal(got, ref), f"max {np.abs(got - ref).max()}"
        else:
            exp, _ = _lp_expected(ref, store)
            assert np.array_equal(got, exp), (store, np.abs(got - exp).max())
    # LINEAR plans run the same network with twiddles rounded from double: no less
    # accurate against the FP64 filter than the reference's float-angle table
    plan = fbp.FbpPlan([spec])
    plan.filter([s.cuda().contiguous()])
    torch.cuda.synchronize()
    lin = plan.filtered_rows_tensor(0).cpu().numpy()
    plan.close()
    ref64 = O.filter_rows_f64(s.numpy().astype(np.float32), g, filt)
    e_lin, e_ref = np.abs(lin - ref64).max(), np.abs(ref - ref64).max()
    print(f"  max error vs FP64: LINEAR {e_lin:.2e}, reference order {e_ref:.2e}")
    assert e_lin <= e_ref


def test_reconst_host_x4_zoom_bit_exact():
    """4096 samples at x4 zoom: n = 32768, a 16384^2 int32 image (1 GiB)."""
    W, P = 4096, 12
    s, fdeg, c_det = _rows(W, P, 4100, 4, int16=True)
    rows = s.numpy()
    print(f"case CudaReconstHost x4 zoom {P}x{W}", flush=True)
    res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, iInterpolation=4, normalise=False)
    g = O.geometry(W, 4, c_det, float_center=True)
    assert res["image"].shape == (g.N, g.N) == (16384, 16384)
    ref = O.recon_rint(rows, fdeg, g, O.set_filter(g.ndim, 0, 0.5, 6.0), formula=1, rule=1)
    assert np.array_equal(res["image"], ref), f"{int((res['image'] != ref).sum())} pixels differ"
    assert ref.max() - ref.min() > 100                  # a non-trivial image


@pytest.mark.parametrize("store,tol", [("f32", 2e-6), ("f16", 1e-4)])
def test_linear_x8_zoom_vs_fp64(store, tol):
    W, P, interp = 2048, 24, 5
    s, fdeg, c_det = _rows(W, P, 4200, interp)
    g = O.geometry(W, interp, c_det)
    N = g.N
    print(f"case LINEAR x8 zoom W={W} N={N} n={g.ndim} store={store}", flush=True)
    plan = fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c_det, iInterpolation=interp)], store=store, persistent=True)
    out = torch.empty(N, N, device="cuda")
    plan.execute([s.cuda().contiguous()], [out])
    torch.cuda.synchronize()
    plan.close()
    px64 = O.filter_rows_f64(s.numpy(), g, O.set_filter(g.ndim, 0))
    ys = np.linspace(0.2 * N, 0.8 * N, 3).astype(int)
    yy, xx = np.meshgrid(ys, np.arange(N), indexing="ij")
    inside = ((xx - N / 2) ** 2 + (yy - N / 2) ** 2) <= (0.45 * N) ** 2
    ref = O.bp_points_linear_f64(px64, fdeg, g, xx[inside], yy[inside])
    got = out.cpu().numpy()[yy[inside], xx[inside]].astype(np.float64)
    rng = ref.max() - ref.min()
    rms = float(np.sqrt(np.mean((got - ref) ** 2)) / rng)
    # the reference's FP32 filter (CFft order) back-projected in FP64: at 8x zoom of
    # 24 projections its own rounding is ~2e-5 of the range, above the 2e-6 that
    # unzoomed slices of many projections reach
    px32 = O.filter_rows(s.numpy(), g, O.set_filter(g.ndim, 0), kind=1).astype(np.float64)
    r32 = O.bp_points_linear_f64(px32, fdeg, g, xx[inside], yy[inside])
    rms_ref = float(np.sqrt(np.mean((r32 - ref) ** 2)) / rng)
    print(f"  rms {rms:.2e} of range (reference FP32 filter: {rms_ref:.2e})")
    assert rms <= max(tol, 1.1 * rms_ref)


def test_cuda_deconv_two_pass():
    L = _lib.lib()
    W, z = 4096, 8
    s, fdeg, c_det = _rows(W, 2, 4300, 5, int16=True)
    rows = s.numpy()
    g = O.geometry(W, 5, c_det, float_center=True)
    assert g.ndim == 65536
    filt = O.set_filter(g.ndim, 0)
    d_strip = torch.from_numpy(rows[1].copy()).cuda()
    d_filt = torch.from_numpy(filt).cuda()
    d_p = torch.zeros(2 * g.ndim, dtype=torch.float32, device="cuda")
    d_igp = torch.zeros(16 * g.N, dtype=torch.int32, device="cuda")
    L.CudaDeconv(W, z, g.ndim, C.c_float(c_det), C.c_void_p(d_filt.data_ptr()),
                 C.c_void_p(d_strip.data_ptr()), C.c_void_p(d_igp.data_ptr()),
                 C.c_void_p(d_p.data_ptr()), None)
    torch.cuda.synchronize()
    pz = O.filter_rows(rows[1:2], g, filt, kind=2)
    assert np.array_equal(d_igp.cpu().numpy(), O.lut16_i32(pz[0], 1))
