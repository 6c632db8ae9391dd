"""Full-size parity of the named configurations against the CPU oracle.

C1: one slice, 900 projections x 1024 detector samples (float32 analytic
    10-ellipse phantom, centre +0.37, Shepp-Logan filter) -> 1024^2.  LUT16 in
    float against the reference CPU path R-f32 (projx64 index rule, CPU LUT
    formula): max |diff| <= 1e-5 of the dynamic range; LINEAR against the FP64
    oracle: RMS and max reported.
C2: 16 slices of one geometry at 1800 x 2048 -> 2048^2, grouped as the plan
    does (FP32 slice pairs, FP16 pairs eight per work item), against FP64
    filtering + FP64 linear back-projection on sampled rows; and the benchmark's
    own 64-slice textured stack, eight slices against FP64 and one in LUT16
    float against R-f32 over the whole image.
C4: one 4096^2 slice from 3600 projections x 4096 samples with 1e5-photon
    Poisson noise, and a 4x zoomed 1024^2 off-centre sub-region; filtered
    sinograms stored as FP32, FP16 and BF16 (per-row scale, FP32 accumulation).
    RMS and max error against FP64 filtering + FP64 linear back-projection on a
    sample of rows (the oracle cost of every pixel would dominate the suite).
    With RV_REPORT_DIR set, the numbers are also written to
    $RV_REPORT_DIR/c4_precision.json.
"""
import json
import os
import time

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import fbp, phantom

pytestmark = pytest.mark.gpu


def _circle(N, frac=0.45):
    yy, xx = np.mgrid[0:N, 0:N]
    return ((xx - N / 2) ** 2 + (yy - N / 2) ** 2) <= (frac * N) ** 2


def _run(specs, sinos, **kw):
    plan = fbp.FbpPlan(specs, **kw)
    outs = [torch.empty(plan.out_n(s), plan.out_n(s), device="cuda") for s in range(len(specs))]
    plan.execute(sinos, outs)
    torch.cuda.synchronize()
    return plan, outs


def test_c1_full_size_parity():
    W, P = 1024, 900
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(1001, count=10, axis_range=(0.05, 0.4), mu_range=(0.0, 0.02))
    sino = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device="cuda").float()
    s_np = sino.cpu().numpy()
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 5)
    spec = fbp.SliceSpec(W, fdeg, c_det, filter="shepp-logan")
    m = _circle(W)
    _, (lut,) = _run([spec], [sino.contiguous()], interp="lut16_f32", lut_formula=0, index_rule=0)
    ref = O.recon_rf32(s_np, fdeg, g, filt, rule=0)
    rng = ref[m].max() - ref[m].min()
    mx = float(np.abs(lut.cpu().numpy() - ref)[m].max() / rng)
    print(f"C1 LUT16_F32 vs R-f32: max {mx:.2e} of range")
    assert mx <= 1e-5
    _, (lin,) = _run([spec], [sino.contiguous()], interp="linear")
    ref64 = O.recon_of64(s_np, fdeg, g, filt)
    d = (lin.cpu().numpy().astype(np.float64) - ref64)[m]
    r64 = ref64[m].max() - ref64[m].min()
    rms, mxl = float(np.sqrt(np.mean(d * d)) / r64), float(np.abs(d).max() / r64)
    print(f"C1 LINEAR vs O-f64: rms {rms:.2e} max {mxl:.2e} of range")
    assert rms <= 2e-6 and mxl <= 1e-4


def _errors(got, ref):
    rng = ref.max() - ref.min()
    d = got.astype(np.float64) - ref
    return float(np.sqrt(np.mean(d * d)) / rng), float(np.abs(d).max() / rng)


def test_c2_grouped_parity():
    """C2 geometry (1800 projections x 2048 samples -> 2048^2) with 16 slices of
    one geometry and different noisy phantoms, as the plan groups them: FP32
    slice pairs (the default) and FP16 pairs sixteen per work item (lds_half).
    Every slice against FP64 filtering + FP64 linear back-projection on six rows
    across the reconstruction circle."""
    W, P, S = 2048, 1800, 16
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 0)
    N = g.N
    rows = np.linspace(0.1 * N, 0.9 * N, 6).astype(int)
    ys, xs = np.meshgrid(rows, np.arange(N), indexing="ij")
    inside = ((xs - N / 2) ** 2 + (ys - N / 2) ** 2) <= (0.45 * N) ** 2
    sinos, refs = [], []
    for i in range(S):
        ell = phantom.random_ellipses(1100 + i, count=10, mu_range=(0.0, 2.0 / W))
        clean = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(1100 + i)
        sino = phantom.poisson(clean, 1e4, generator=gen).contiguous()
        px64 = O.filter_rows_f64(sino.cpu().numpy(), g, filt)
        refs.append(O.bp_points_linear_f64(px64, fdeg, g, xs[inside], ys[inside]))
        sinos.append(sino)
    specs = [fbp.SliceSpec(W, fdeg, c_det) for _ in range(S)]
    for store, lds_half, tol in (("f32", False, 2e-6), ("f16", True, 1e-4)):
        plan, outs = _run(specs, sinos, store=store, lds_half=lds_half, persistent=True)
        worst = 0.0
        for i in range(S):
            rms, mx = _errors(outs[i].cpu().numpy()[ys[inside], xs[inside]], refs[i])
            worst = max(worst, rms)
            assert rms <= tol, (store, i, rms, mx)
        print(f"C2 grouped {store}{'+lds_half' if lds_half else ''}: worst slice rms {worst:.2e} of range",
              flush=True)
        plan.close()


def test_c2_stack_parity():
    """The headline C2 inputs themselves (bench.make_stack: a 64-slice stack of a
    40-ellipsoid phantom with the 2% white-noise texture and Poisson noise at 1e4
    photons), reconstructed the way the benchmark does (FP32 slice pairs, tuned tile
    and chunk, one hipGraph; its FP16 line: 16-bit filtered rows, sixteen FP16 pairs
    per work item; and the opt-in matrix-core kernel, two 32-slice items): eight
    slices across the stack against FP64 filtering
    + FP64 linear back-projection on sampled rows, RMS <= 2e-6 (FP32) / 1e-4 (FP16)
    of the range; and one slice in LUT16 float against the reference FP32 path R-f32
    (projx64 index rule, CPU LUT formula) over the whole image, max <= 1e-5 of the
    range."""
    import bench
    from recview_amd import autotune
    W, P, S = 2048, 1800, 64
    sinos, fdeg, c_det = bench.make_stack(S, P, W, 1002, torch.device("cuda"))
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 0)
    rows = np.linspace(0.1 * W, 0.9 * W, 5).astype(int)
    ys, xs = np.meshgrid(rows, np.arange(W), indexing="ij")
    inside = ((xs - W / 2) ** 2 + (ys - W / 2) ** 2) <= (0.45 * W) ** 2
    refs = {}
    for i in range(0, S, 9):
        px64 = O.filter_rows_f64(sinos[i].cpu().numpy(), g, filt)
        refs[i] = O.bp_points_linear_f64(px64, fdeg, g, xs[inside], ys[inside])
    for store, lds_half, spi, tol in (("f32", False, 0, 2e-6), ("f16", True, 0, 1e-4), ("f16", True, 32, 1e-4)):
        specs = [fbp.SliceSpec(W, fdeg, c_det) for _ in range(S)]
        autotune.apply(specs, [W] * S, autotune.mode_of("linear", store, lds_half, spi))
        plan = fbp.FbpPlan(specs, store=store, lds_half=lds_half, slices_per_item=spi, persistent=True)
        if spi == 32:
            assert plan.info(0)["tile_w"] == 32                  # the matrix-core items
        outs = [torch.empty(W, W, device="cuda") for _ in range(S)]
        plan.execute([sinos[i] for i in range(S)], outs, graph=True)
        torch.cuda.synchronize()
        plan.close()
        worst = 0.0
        for i, ref in refs.items():
            rms, mx = _errors(outs[i].cpu().numpy()[ys[inside], xs[inside]], ref)
            worst = max(worst, rms)
            assert rms <= tol, (store, i, rms, mx)
        print(f"C2 stack (texture) {store} slices_per_item={spi}: worst sampled-slice rms {worst:.2e} "
              f"of range vs FP64", flush=True)
        del outs
    s_np = sinos[S // 2].cpu().numpy()
    spec = fbp.SliceSpec(W, fdeg, c_det)
    _, (lut,) = _run([spec], [sinos[S // 2]], interp="lut16_f32", lut_formula=0, index_rule=0)
    ref = O.recon_rf32(s_np, fdeg, g, filt, rule=0)
    m = _circle(W)
    rng = ref[m].max() - ref[m].min()
    mxl = float(np.abs(lut.cpu().numpy() - ref)[m].max() / rng)
    print(f"C2 slice {S // 2} LUT16_F32 vs R-f32: max {mxl:.2e} of range", flush=True)
    assert mxl <= 1e-5


def test_c4_precision_report():
    W, P = 4096, 3600
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(1004, count=10, mu_range=(0.0, 2.0 / W))
    clean = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(1004)
    sino = phantom.poisson(clean, 1e5, generator=gen).contiguous()
    s_np = sino.cpu().numpy()
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 0)
    t0 = time.time()
    px64 = O.filter_rows_f64(s_np, g, filt)
    N = g.N
    # full slice: 24 rows spread over the reconstruction circle, all columns
    rows = np.linspace(0.08 * N, 0.92 * N, 24).astype(int)
    ys, xs = np.meshgrid(rows, np.arange(N), indexing="ij")
    inside = ((xs - N / 2) ** 2 + (ys - N / 2) ** 2) <= (0.45 * N) ** 2
    ref_full = O.bp_points_linear_f64(px64, fdeg, g, xs[inside], ys[inside])
    # 4x zoom of an off-centre 1024^2 region (256 x 256 reference pixels)
    z, n_roi = 4.0, 1024
    x0, y0 = 0.62 * N - 128, 0.33 * N - 128
    roi_rows = np.arange(0, n_roi, 64)
    ry, rx = np.meshgrid(roi_rows, np.arange(n_roi), indexing="ij")
    ref_roi = O.bp_points_linear_f64(px64, fdeg, g, x0 + rx / z, y0 + ry / z).reshape(ry.shape)
    t_oracle = time.time() - t0
    report = {"config": "C4: 4096^2 from 3600 proj x 4096 det, 1e5 photons; 4x zoom 1024^2 ROI at "
                        f"({x0:.0f}, {y0:.0f})", "oracle": "FP64 filter + FP64 linear BP on "
              f"{int(inside.sum())} full-slice pixels and {ry.size} ROI pixels ({t_oracle:.1f} s)",
              "results": {}}
    limits = {"f32": 2e-6, "f16": 1e-4, "bf16": 2e-3}
    for store in ("f32", "f16", "bf16"):
        for lds_half in ((False, True) if store != "f32" else (False,)):
            key = store + ("+lds_half" if lds_half else "")
            spec = fbp.SliceSpec(W, fdeg, c_det)
            plan, (img,) = _run([spec], [sino], store=store, lds_half=lds_half, persistent=True)
            out = torch.empty(N, N, device="cuda")
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(3):
                plan.execute([sino], [out])
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / 3
            full_rms, full_max = _errors(img.cpu().numpy()[ys[inside], xs[inside]], ref_full)
            roi_spec = fbp.SliceSpec(W, fdeg, c_det, out_n=n_roi, pix_zoom=z, x0=x0, y0=y0)
            _, (roi,) = _run([roi_spec], [sino], store=store, lds_half=lds_half, persistent=True)
            roi_rms, roi_max = _errors(roi.cpu().numpy()[roi_rows], ref_roi)
            report["results"][key] = {"full_rms": full_rms, "full_max": full_max, "roi_rms": roi_rms,
                                      "roi_max": roi_max, "ms_per_slice": ms,
                                      "tau_ns": ms * 1e-3 / (N * N * P) * 1e9}
            print(f"C4 {key}: full rms {full_rms:.2e} max {full_max:.2e}; roi rms {roi_rms:.2e} "
                  f"max {roi_max:.2e}; {ms:.2f} ms/slice", flush=True)
            assert full_rms <= limits[store] and roi_rms <= limits[store], key
    out_dir = os.environ.get("RV_REPORT_DIR")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "c4_precision.json"), "w") as f:
            json.dump(report, f, indent=1)
