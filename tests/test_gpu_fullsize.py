"""Parity at the full sizes of the named configurations.

* The drop-in CudaReconstHost / CudaReconstHostFFT on one 1800 x 2048 int16
  slice (the C2 shape): the LUT16 integer image, whose back-projection splits
  every tile's projections over four work items summed atomically, equals the
  reference-semantics oracle R-int bit for bit, for both LUT formulas and both
  index rules.
* The reference-order filter (CFft network) at transform lengths 4096, 8192 and
  16384: filtered rows bit-identical to the oracle's CFft restatement.
* A batch of slices at the C3 shapes in one persistent launch: widths 512..4096,
  projections 450..3600, an offset-axis 360-degree scan stitched to 8192 samples
  (16384-point transform, 8192^2 image), FP32 and FP16 filtered sinograms;
  every slice against FP64 filtering + FP64 linear back-projection on sampled
  rows, the batch equal to each slice planned alone, hipGraph replay equal to
  the eager result.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import _lib, fbp, offset, phantom, reconst

pytestmark = pytest.mark.gpu


def _c2_rows(seed=2002):
    W, P = 2048, 1800
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(seed, count=10, mu_range=(0.0, 2.0 / W))
    clean = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(seed)
    noisy = phantom.poisson(clean, 1e4, generator=gen)
    return phantom.to_int16(noisy.double().cpu()).numpy(), fdeg, c_det


@pytest.mark.parametrize("formula,rule", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_reconst_host_c2_bit_exact(formula, rule):
    rows, fdeg, c_det = _c2_rows()
    W = rows.shape[1]
    print(f"case CudaReconstHost {rows.shape[0]}x{W} formula={formula} rule={rule}", flush=True)
    _lib.lib().RecviewSetCompatOptions(formula, rule)
    try:
        res = reconst.deconv_back_proj(rows, fdeg, c_det - 0.5, normalise=False)
    finally:
        _lib.lib().RecviewSetCompatOptions(1, 1)
    g = O.geometry(W, 2, c_det, float_center=True)
    ref = O.recon_rint(rows, fdeg, g, O.set_filter(g.ndim, 0, 0.5, 6.0), formula=formula, rule=rule)
    diff = np.abs(res["image"].astype(np.int64) - ref)
    print(f"  max |GPU - R-int| = {diff.max()}, image range {ref.min()}..{ref.max()}")
    assert np.abs(ref).max() > 10 ** 5
    assert np.array_equal(res["image"], ref)


def test_reconst_host_fft_c2_bit_exact():
    """CudaReconstHostFFT: zero padding and placement n/2 of the CUDAFFT build."""
    rows, fdeg, c_det = _c2_rows(2003)
    W = rows.shape[1]
    print(f"case CudaReconstHostFFT {rows.shape[0]}x{W}", flush=True)
    info = reconst.ReconstInfo()
    filt = reconst.set_filter(reconst.ndim_for(W, 0), 0)
    img, _ = reconst.fill_info(info, rows, fdeg, c_det - 0.5, 2, filt)
    _lib.lib().CudaReconstHostFFT(C.byref(info.ri), 0, False)
    status = info.ri.iStatus
    info.free()
    assert status != reconst.RECONST_INFO_ERROR
    g = O.geometry(W, 2, c_det, float_center=True)
    f = O.set_filter(g.ndim, 0, 0.5, 6.0)
    px = O.filter_rows(rows, g, f, kind=2)
    ref = np.zeros((g.N, g.N), np.int32)
    for i in range(rows.shape[0]):
        c, s, o = O.bp_params(fdeg[i], 0.0, g, 0)
        O.bp_lut_i32(ref, O.lut16_i32(px[i], 1), g.N, c, s, o, 1)
    assert np.array_equal(img, ref)


@pytest.mark.parametrize("W,interp,kind", [(2048, 2, 1), (4096, 2, 1), (4096, 2, 0), (4096, 3, 1)])
def test_exact_filter_rows_bit_identical_large(W, interp, kind):
    """Reference-order filter at 4096, 8192 and 16384 points (the last with the
    LDS raised above 64 KiB): rows bit-identical to the CFft restatement, float
    and int16 inputs."""
    P = 40
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(2100 + W + interp, count=10, mu_range=(0.0, 2.0 / W))
    s = phantom.sinogram(ell, W, fdeg, c_det, scale=W).float()
    if kind == 0:
        s = phantom.to_int16(s.double())
    g = O.geometry(W, interp, c_det)
    print(f"case exact filter W={W} n={g.ndim} kind={kind}", flush=True)
    spec = fbp.SliceSpec(W, fdeg, c_det, iInterpolation=interp, filter="hamming",
                         sino_dtype="i16" if kind == 0 else "f32")
    plan = fbp.FbpPlan([spec], exact_filter=True)
    plan.filter([s.cuda().contiguous()])
    torch.cuda.synchronize()
    got = plan.filtered_rows_tensor(0).cpu().numpy()
    plan.close()
    ref = O.filter_rows(s.numpy(), g, O.set_filter(g.ndim, 1), kind=kind)
    assert np.array_equal(got, ref), f"max {np.abs(got - ref).max()}"


def _c3_slices():
    """(W, P, offset) at the C3 shapes: the extremes of both axes, an offset-axis
    scan stitched to double width, and a pair of equal geometry (grouped)."""
    return [(4096, 3600, False), (512, 450, False), (4096, 900, True), (1024, 3600, False),
            (2048, 1800, False), (2048, 1800, False), (512, 1800, True), (1024, 450, False)]


def _c3_case(i, W, P, off):
    if off:
        c_det = W / 2 + (0.3 if i % 2 else -0.3) * W + 0.37
        fdeg = (np.arange(2 * P + 1) * (180.0 / P)).astype(np.float32)
        ell = phantom.random_ellipses(3300 + i, count=10, mu_range=(0.0, 1.0 / W))
        fov = 2 * min(c_det, W - c_det) + W
        s = phantom.sinogram(ell, W, fdeg, c_det, scale=0.9 * fov, device="cuda")
        # int16 rows, so that the oracle's stitch (int16) applies exactly
        return phantom.to_int16(s.double()).contiguous(), fdeg, c_det
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(3300 + (i if i != 5 else 4), count=10, mu_range=(0.0, 2.0 / W))
    s = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device="cuda").float()
    if i == 5:
        s = s.flip(1).contiguous()          # equal geometry to slice 4, different data
    return s.contiguous(), fdeg, c_det


def _c3_reference(sino, fdeg, c_det, W, off, ys):
    """FP64 filtering and FP64 linear back-projection of the rows the plan
    reconstructs, at pixel rows ys (every column)."""
    s = sino.cpu().numpy()
    if off:
        rows = O.offset_stitch(s, fdeg, np.ones(len(fdeg), np.int8), c_det)
        ang = fdeg[:rows.shape[0]]
        c = offset.stitched_center(c_det - 0.5, W) + 0.5
        g = O.geometry(2 * W, 2, c)
        s = rows.astype(np.float32)
    else:
        ang = fdeg
        g = O.geometry(W, 2, c_det)
    px64 = O.filter_rows_f64(s, g, O.set_filter(g.ndim, 0))
    yy, xx = np.meshgrid(ys, np.arange(g.N), indexing="ij")
    return O.bp_points_linear_f64(px64, ang, g, xx, yy).reshape(yy.shape), g


@pytest.mark.parametrize("store,lds_half,rms_tol", [("f32", False, 2e-6), ("f16", True, 1e-4)])
def test_c3_shapes_batch(store, lds_half, rms_tol):
    cases = _c3_slices()
    specs, sinos, info = [], [], []
    for i, (W, P, off) in enumerate(cases):
        s, fdeg, c_det = _c3_case(i, W, P, off)
        specs.append(fbp.SliceSpec(W, fdeg, c_det, offset_ct=off,
                                   sino_dtype="i16" if s.dtype == torch.int16 else "f32"))
        sinos.append(s)
        info.append((W, P, off, fdeg, c_det))
    print(f"case C3 batch of {len(cases)} slices store={store} lds_half={lds_half}", flush=True)
    plan = fbp.FbpPlan(specs, store=store, lds_half=lds_half, persistent=True)
    outs = [torch.empty(plan.out_n(i), plan.out_n(i), device="cuda") for i in range(len(cases))]
    plan.execute(sinos, outs)
    torch.cuda.synchronize()
    assert plan.info(2)["n_fft"] == 16384 and plan.out_n(2) == 8192
    eager = [o.cpu().numpy() for o in outs]
    # one hipGraph: equal to the eager execution on every replay
    for rep in range(2):
        for o in outs:
            o.fill_(float("nan"))
        plan.execute(sinos, outs, graph=True)
        torch.cuda.synchronize()
        for i, o in enumerate(outs):
            assert np.array_equal(o.cpu().numpy(), eager[i]), f"replay {rep} slice {i}"
    plan.close()
    for i, (W, P, off, fdeg, c_det) in enumerate(info):
        N = eager[i].shape[0]
        ys = np.linspace(0.15 * N, 0.85 * N, 4).astype(int)
        ref, g = _c3_reference(sinos[i], fdeg, c_det, W, off, ys)
        assert g.N == N
        yy, xx = np.meshgrid(ys, np.arange(N), indexing="ij")
        inside = ((xx - N / 2) ** 2 + (yy - N / 2) ** 2) <= (0.45 * N) ** 2
        got = eager[i][ys][inside].astype(np.float64)
        r = ref[inside]
        rng = r.max() - r.min()
        rms = float(np.sqrt(np.mean((got - r) ** 2)) / rng)
        print(f"  slice {i}: W={W} P={P} offset={off} N={N}: rms {rms:.2e} of range", flush=True)
        assert rms <= rms_tol, (i, rms)
        # the same slice planned alone: equal to its image in the batch (one slice
        # per work item vs grouped is bit-identical for FP32 pairs; FP16 groups share
        # one item scale, so a lone slice may round differently -- to the FP64 bound)
        alone = fbp.FbpPlan([specs[i]], store=store, lds_half=lds_half)
        o1 = torch.empty(N, N, device="cuda")
        alone.execute([sinos[i]], [o1])
        torch.cuda.synchronize()
        alone.close()
        a = o1.cpu().numpy()
        if store == "f32":
            assert np.array_equal(a, eager[i]), f"slice {i}: batch differs from the slice alone"
        else:
            d = (a[ys][inside].astype(np.float64) - r)
            assert float(np.sqrt(np.mean(d * d)) / rng) <= rms_tol
