"""Benchmark: filtered back-projection throughput on one MI355X.

Default workload (the configuration the headline metric is quoted on): a
64-slice stack, each slice 1800 projections over 180 deg x 2048 detector
samples (float32 line integrals of a seeded ellipsoid phantom with Poisson
noise at 1e4 photons/ray), reconstructed to 2048 x 2048 float32 with the Hann
filter and linear detector interpolation.  One step = filter + back-projection
of the whole stack, sinograms already resident in HBM, images written to HBM.

--workload c5: sparse-view restoration -- 16 slices reconstructed from 225 of
1800 angles and restored by the BF16-MFMA U-Net, beside full-angle FBP of the
same slices (rate ratio, PSNR / SSIM).

--workload dropin: the host-buffer entry point CudaReconstHost, one call per
slice (int16 rows in, int32 image out): the PCIe-inclusive rate.

--workload c3: 256 slices of uneven width {512..4096} and projection count
{450..3600}, a quarter of them 360-degree offset-axis scans stitched on the GPU
to double width, per-shape autotuned tiles, one hipGraph per step.

Prints one JSON line.  value = slices/s for the whole job; tau = time per
pixel*projection (README.md:96 of the reference); roofline = FP32 VALU
fraction of the back-projection kernel (6 algorithmic FLOP per pixel*projection
update), timed with HIP events on the launch stream; cpu_baseline = the
reference CPU path (projx64 AVX-512 loop restated in C, reference threading)
timed on this host (default workload only).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FBP tau (ns/pixel·projection) + slices/s at 2048²×1800; sparse+CNN PSNR"
REF_TAU_NS = 0.029          # RTX A4000, Build-220422, README.md:53 (2048^2 x 1800)
FP32_PEAK_TFLOPS = 157.3    # MI355X published FP32 vector peak
FP16_PEAK_TFLOPS = 2500.0   # MI355X published dense FP16 MFMA peak
FLOP_PER_UPDATE = 6         # t = fma, frac, lerp fma, accumulate


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_stack(nslices, P, W, seed, device, photons=1e4, texture=0.02):
    """Float32 sinograms [nslices][P][W] of a seeded 40-ellipsoid phantom with a
    white-noise texture (each pixel's attenuation times 1 + texture * N(0, 1), the
    texture projected by the ray-driven projector), Poisson noise and -ln."""
    import torch
    from recview_amd import phantom
    # attenuation per sample scaled with the width so that rays are not photon starved
    ells = phantom.random_ellipsoids(seed, count=40, mu_range=(0.0, 4.0 / W))
    fdeg = phantom.angles_deg(P)
    c_det = W / 2 + 0.37
    gen = torch.Generator(device=device).manual_seed(seed)
    out = torch.empty(nslices, P, W, dtype=torch.float32, device=device)
    for s in range(nslices):
        z = -0.3 + 0.6 * (s + 0.5) / nslices
        ell = phantom.ellipsoid_slice(ells, z)
        sino = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device=device)
        if texture:
            img = phantom.image(ell, W, W, device=device, dtype=torch.float32)
            noise = torch.randn(W, W, device=device, generator=gen)
            sino = sino + phantom.texture_sinogram(img * noise * texture, W, fdeg, c_det).double()
        out[s] = phantom.poisson(sino, photons, generator=gen)
    return out, fdeg, c_det


def make_c3(nslices, seed, device, interp, store, lds_half=False, slices_per_item=0):
    """256 uneven slices; a quarter are 360-degree offset-axis scans (stitched to 2W)."""
    import numpy as np
    import torch
    from recview_amd import autotune, fbp, phantom
    rng = phantom.SplitMix64(seed)
    widths, projs = (512, 1024, 2048, 4096), (450, 900, 1800, 3600)
    specs, sinos, shapes = [], [], []
    table = autotune.load()
    for s in range(nslices):
        W = widths[rng.next_u64() % 4]
        P = projs[rng.next_u64() % 4]
        offset = (s % 4 == 3)
        ell = phantom.random_ellipses(seed * 1000 + s, count=10, mu_range=(0.0, 2.0 / W))
        if offset:
            side = 1 if rng.next_u64() % 2 else -1
            c_det = W / 2 + side * 0.3 * W + 0.37
            fdeg = (np.arange(2 * P + 1) * (180.0 / P)).astype(np.float32)
            fov = 2 * min(c_det, W - c_det) + W       # the stitched FOV
            sino = phantom.sinogram(ell, W, fdeg, c_det, scale=0.9 * fov, device=device).float()
            N, Pe = 2 * W, P + 1
        else:
            c_det = W / 2 + 0.37
            fdeg = phantom.angles_deg(P)
            sino = phantom.sinogram(ell, W, fdeg, c_det, scale=W, device=device).float()
            N, Pe = W, P
        specs.append(fbp.SliceSpec(W, fdeg, c_det, offset_ct=offset))
        sinos.append(sino.contiguous())
        shapes.append((N, Pe))
    mode = autotune.mode_of(interp, store, lds_half, slices_per_item)
    if mode:
        autotune.apply(specs, [n for n, _ in shapes], mode, table)
    return specs, sinos, shapes


def read_traffic(name):
    """HBM bytes per launch of the back-projection kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_cpu():
    """(model, MHz, instruction-set flags, usable threads) of this host."""
    model, mhz, flags = "unknown", None, []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model == "unknown":
                    model = v.strip()
                elif k == "cpu MHz" and mhz is None:
                    mhz = float(v)
                elif k == "flags" and not flags:
                    have = set(v.split())
                    flags = [f for f in ("avx512f", "avx512dq", "avx512vl", "avx2", "fma") if f in have]
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, mhz, flags, usable


def cpu_baseline(sino_f32, fdeg, c_det, threads):
    """The reference CPU reconstruction of one slice (int16 input, LUT16, AVX-512)."""
    from oracle import pyoracle as O
    from recview_amd import phantom
    rows = phantom.to_int16(sino_f32.double().cpu()).numpy()
    W = rows.shape[1]
    g = O.geometry(W, 2, c_det, float_center=False)
    filt = O.set_filter(g.ndim, 0, 0.5, 6.0)
    # small warm-up to fault in code pages, then the timed slice
    O.cpu_baseline(rows[:8], fdeg[:8], g, filt, threads)
    secs, _ = O.cpu_baseline(rows, fdeg, g, filt, threads)
    isa = {2: "avx512", 1: "avx2", 0: "scalar"}[O.cpu_isa()]
    return secs, isa


BF16_PEAK_TFLOPS = 2500.0   # MI355X published dense BF16 MFMA peak


def run_c5(args):
    """Sparse-view (every 8th of P angles) FBP + the restoration U-Net on a batch
    of slices, against full-angle FBP of the same slices."""
    import torch
    from recview_amd import fbp, learned, phantom, unet

    dev = torch.device("cuda:0")
    S, P, W = (args.slices or 16), args.proj, args.width
    weights = os.path.join(ROOT, "recview_amd", "weights", "unet_sparse8.npz")
    trained = os.path.exists(weights)
    net = unet.Unet.load(weights) if trained else unet.Unet(seed=0)
    scale = net.scale if trained else W / 2.0
    t0 = time.time()
    seeds = list(range(3_000_001, 3_000_001 + S))
    sinos = [learned.slice_sinogram(s, W, P, dev)[0].contiguous() for s in seeds]
    sparse = [x[::learned.SPARSE].contiguous() for x in sinos]
    fdeg = phantom.angles_deg(P)
    c = W / 2.0
    sp_plan = fbp.FbpPlan([fbp.SliceSpec(W, fdeg[::learned.SPARSE], c, filter=learned.FILTER)] * S,
                          persistent=True, normalize=True)
    full_plan = fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c, filter=learned.FILTER)] * S, persistent=True,
                            normalize=True)
    N = sp_plan.out_n(0)
    sp_img = torch.empty(S, N, N, device=dev)
    restored = torch.empty_like(sp_img)
    full_img = torch.empty_like(sp_img)
    sp_views, full_views = list(sp_img.unbind(0)), list(full_img.unbind(0))
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    log(f"inputs generated in {time.time() - t0:.1f}s (trained weights: {trained})")

    def sparse_step():
        sp_plan.execute(sparse, sp_views, stream=stream, graph=True)
        net.forward(sp_img, scale, out=restored, stream=stream)

    def full_step():
        full_plan.execute(sinos, full_views, stream=stream, graph=True)

    def timed(fn, steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    for _ in range(args.warmup):
        sparse_step()
        full_step()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fbp_ms, cnn_ms = [], []
    for _ in range(max(1, min(3, args.steps))):
        ev[0].record(stream)
        sp_plan.execute(sparse, sp_views, stream=stream, graph=True)
        ev[1].record(stream)
        net.forward(sp_img, scale, out=restored, stream=stream)
        ev[2].record(stream)
        torch.cuda.synchronize()
        fbp_ms.append(ev[0].elapsed_time(ev[1]))
        cnn_ms.append(ev[1].elapsed_time(ev[2]))
    fbp_ms, cnn_ms = sum(fbp_ms) / len(fbp_ms), sum(cnn_ms) / len(cnn_ms)
    t_sparse = timed(sparse_step, args.steps)
    t_full = timed(full_step, args.steps)

    yy, xx = torch.meshgrid(torch.arange(N, device=dev), torch.arange(N, device=dev), indexing="ij")
    mask = ((xx - N / 2 + 0.5) ** 2 + (yy - N / 2 + 0.5) ** 2) <= (0.9 * N / 2) ** 2
    q = {"psnr_vs_full": {"sparse": [], "restored": []}, "ssim_vs_full": {"sparse": [], "restored": []},
         "psnr_phantom": {"sparse": [], "restored": [], "full": []}}
    for i in range(min(S, 4)):
        truth = learned.slice_image(seeds[i], W, dev)
        rf = float(full_img[i][mask].max() - full_img[i][mask].min())
        rt = float(truth[mask].max() - truth[mask].min())
        for k, img in (("sparse", sp_img[i]), ("restored", restored[i])):
            q["psnr_vs_full"][k].append(learned.psnr(img, full_img[i], mask, rf))
            q["ssim_vs_full"][k].append(learned.ssim(img, full_img[i], mask, rf))
        for k, img in (("sparse", sp_img[i]), ("restored", restored[i]), ("full", full_img[i])):
            q["psnr_phantom"][k].append(learned.psnr(img, truth, mask, rt))
    quality = {m: {k: round(sum(v) / len(v), 3) for k, v in d.items()} for m, d in q.items()}

    flops = unet.flops_per_pixel(net.base, net.levels) * S * N * N
    achieved = flops / (cnn_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import unet_ref
        crop = 128
        p = net.get_params().astype(np.float64)
        x = sp_img[0, N // 2 - crop // 2:N // 2 + crop // 2, N // 2 - crop // 2:N // 2 + crop // 2]
        x = x.double().cpu().numpy()[None]
        ref = unet_ref.UnetRef(net.base, net.levels)
        t = time.perf_counter()
        ref.forward(p, x, scale, rnd=None)
        secs = (time.perf_counter() - t) * (N * N) / (crop * crop)
        cpu = {"value": round(1.0 / secs, 6), "unit": "slices/s", "cores": _blas_threads(), "kind": "port",
               "sample": f"network only: FP64 numpy forward of one {crop}x{crop} crop, scaled to a "
                         f"{N}x{N} slice"}
    result = {
        "metric": "sparse-view FBP + restoration U-Net slices/s at 2048^2 (225 of 1800 angles)",
        "value": round(S / t_sparse, 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_sparse * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "full_fbp_slices_per_s": round(S / t_full, 3),
        "rate_vs_full_fbp": round(t_full / t_sparse, 3),
        "sparse_fbp_ms": round(fbp_ms, 3),
        "cnn_ms": round(cnn_ms, 3),
        "quality": quality,
        "trained": trained,
        "config": {"workload": f"C5: {S} slices, {P // learned.SPARSE} of {P} proj x {W} det -> {N}^2, "
                               f"sparse FBP + U-Net (4 levels, 32 base ch, bf16 MFMA)",
                   "slices": S, "projections": P // learned.SPARSE, "detector": W},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def run_dropin(args):
    """The drop-in entry point with host buffers: CudaReconstHost on int16 rows
    (LUT16 integer semantics), image copied back -- the PCIe-inclusive rate.
    value: the C++ host (recview_amd/recview_harness, RECONST_INFO filled like
    CGazoDoc::DeconvBackProj) calling CudaReconstHost repeatedly on one slice;
    python_ms_per_slice: the same calls from the ctypes caller recview_amd.reconst
    (RECONST_INFO set-up and image copies in Python included)."""
    import subprocess
    import tempfile
    import torch
    from recview_amd import build as rvbuild, phantom, reconst

    dev = torch.device("cuda:0")
    S, P, W = (args.slices or 8), args.proj, args.width
    sinos, fdeg, c_det = make_stack(S, P, W, 1002, dev)
    rows = [phantom.to_int16(sinos[s].double().cpu()).numpy() for s in range(S)]
    info = reconst.ReconstInfo()
    c_user = c_det - 0.5
    for _ in range(max(1, args.warmup)):
        reconst.deconv_back_proj(rows[0], fdeg, c_user, info=info, normalise=False)
    t = time.perf_counter()
    for k in range(args.steps):
        for s in range(S):
            reconst.deconv_back_proj(rows[s], fdeg, c_user, info=info, normalise=False)
    py_el = (time.perf_counter() - t) / args.steps / S
    info.free()
    if not os.path.exists(rvbuild.HARNESS):
        raise RuntimeError("recview_harness is not built (python -m recview_amd.build)")
    reps = max(1, args.steps) * S
    with tempfile.TemporaryDirectory() as d:
        rows[0].astype(np.int16).tofile(os.path.join(d, "sino.i16"))
        np.asarray(fdeg, np.float32).tofile(os.path.join(d, "fdeg.f32"))
        out = subprocess.run([rvbuild.HARNESS, "--repeat", str(reps), os.path.join(d, "sino.i16"), str(P), str(W),
                              repr(float(c_user)), os.path.join(d, "fdeg.f32"), os.path.join(d, "out.i32")],
                             capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"recview_harness failed ({out.returncode}): {out.stderr.strip()}")
    ms = float(re.search(r"ms_per_call\s+([0-9.]+)", out.stdout).group(1))
    result = {
        "metric": "drop-in CudaReconstHost slices/s incl. host transfers (2048^2 x 1800, LUT16 int)",
        "value": round(1e3 / ms, 3), "unit": "slices/s", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "ms_per_slice": round(ms, 3),
        "tau_ns": round(ms * 1e-3 / (W * W * P) * 1e9, 7),
        "python_ms_per_slice": round(py_el * 1e3, 3),
        "config": {"workload": f"drop-in: {P} proj x {W} det int16 host rows -> int32 host image, one "
                               "CudaReconstHost call per slice from the C++ host (pinned image buffer)",
                   "calls": reps},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def d2h_inclusive(plan, specs, sino_list, outs, args, stream, S, updates):
    """Steps whose window includes the image download, as the reference's timer
    does (gazoDoc.cpp:3873-3991, D2H at cudaReconstHost.cpp:811): two image sets
    and two plans alternate, so step k's images travel to pinned host memory on a
    copy stream (default priority) while step k+1 back-projects into the other set.
    Returns slices/s over args.steps such steps, from the first launch to the last
    byte on the host."""
    import torch
    from recview_amd import fbp
    plan_b = fbp.FbpPlan(specs, interp=args.interp, store=args.store, persistent=bool(args.persistent),
                         tile=max(args.tile, 0), angles_per_chunk=args.chunk, lds_half=bool(args.lds_half),
                         slices_per_item=args.slices_per_item)
    outs_b = [torch.empty_like(o) for o in outs]
    host = torch.empty((S,) + tuple(outs[0].shape), dtype=outs[0].dtype, pin_memory=True)
    copy = torch.cuda.Stream()
    sets = [(plan, outs), (plan_b, outs_b)]
    done = [None, None]

    def step(k):
        p, o = sets[k % 2]
        if done[k % 2] is not None:
            stream.wait_event(done[k % 2])          # the set's previous download has finished
        p.execute(sino_list, o, stream=stream, graph=bool(args.graph))
        ready = torch.cuda.Event()
        ready.record(stream)
        copy.wait_event(ready)
        with torch.cuda.stream(copy):
            for s in range(S):
                host[s].copy_(o[s], non_blocking=True)
        d = torch.cuda.Event()
        d.record(copy)
        done[k % 2] = d

    for k in range(max(2, args.warmup)):
        step(k)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    if not torch.equal(host[S - 1], sets[(args.steps - 1) % 2][1][S - 1].cpu()):
        raise RuntimeError("downloaded image differs from the device image")
    plan_b.close()
    del outs_b
    dma = (S * args.steps / el, el)

    # Direct: the images are pinned host memory, mapped into the GPU's address
    # space; the back-projection writes each finished tile across the host link.
    hosts = [torch.empty(tuple(outs[0].shape), dtype=outs[0].dtype, pin_memory=True) for _ in range(S)]
    for _ in range(max(1, args.warmup)):
        plan.execute(sino_list, hosts, stream=stream, graph=bool(args.graph))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(sino_list, hosts, stream=stream, graph=bool(args.graph))
    torch.cuda.synchronize()
    el_d = time.perf_counter() - t
    plan.execute(sino_list, outs, stream=stream, graph=bool(args.graph))
    torch.cuda.synchronize()
    for s in (0, S - 1):
        if not torch.equal(hosts[s], outs[s].cpu()):
            raise RuntimeError("image written to host memory differs from the device image")
    nbytes = sum(o.numel() * o.element_size() for o in outs)
    best = min(el, el_d)
    return {"value": round(S * args.steps / best, 3), "unit": "slices/s",
            "ms_per_step": round(1e3 * best / args.steps, 3),
            "tau_ns": round(best / args.steps / updates * 1e9, 7),
            "image_bytes_per_step": nbytes,
            "dma": {"value": round(dma[0], 3), "ms_per_step": round(1e3 * dma[1] / args.steps, 3),
                    "window": "filter + back-projection, then every image copied to pinned host memory on a "
                              "copy stream, overlapped with the next step's back-projection"},
            "direct": {"value": round(S * args.steps / el_d, 3), "ms_per_step": round(1e3 * el_d / args.steps, 3),
                       "window": "filter + back-projection writing every image straight into pinned host "
                                 "memory (mapped into the GPU address space)"}}


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        return os.cpu_count() or 1


def dtype_of(args):
    """Arithmetic of the back-projection: FP16 samples and weights with FP32 sums when
    the filtered rows are 16-bit and stay 16-bit in LDS, else FP32 (LUT16 integer: int32)."""
    if args.interp == "lut16_i32":
        return "int32"
    return "f16" if (args.store != "f32" and args.lds_half) else "f32"


def roofline_of(args, S, achieved, traffic):
    """FP32 kernels against the FP32 vector peak; the opt-in matrix-core kernel (FP16
    rows, --slices-per-item 32) against the dense FP16 MFMA peak.  achieved is
    algorithmic (6 FLOP per update) either way; the MFMA kernel executes 32 FLOP per
    update (16-sample windows), reported beside it."""
    m32 = (args.store != "f32" and args.lds_half and args.slices_per_item == 32 and S >= 32
           and args.interp == "linear")
    if m32:
        return {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP16_PEAK_TFLOPS, 4), "traffic": traffic,
                "mfma_executed": round(achieved * 32 / FLOP_PER_UPDATE, 3)}
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c5", "dropin"])
    ap.add_argument("--slices", type=int, default=0, help="default 64 (c2) / 256 (c3) / 16 (c5)")
    ap.add_argument("--proj", type=int, default=1800)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--interp", default="linear", choices=["linear", "lut16_f32", "lut16_i32"])
    ap.add_argument("--store", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--tile", type=int, default=-1, help="-1: tile and chunk from the tuned table")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--lds-half", type=int, default=0, help="FP16 pairs in LDS for 16-bit stores")
    ap.add_argument("--persistent", type=int, default=1)
    ap.add_argument("--slices-per-item", type=int, default=0,
                    help="LINEAR: slices of identical geometry per work item (0 = four with FP32 "
                         "rows, sixteen with FP16 pairs; or 1, 2, 4, 8, 16, 32)")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--texture", type=float, default=0.02,
                    help="c2: white-noise texture of the phantom (fraction of the attenuation; 0 = none)")
    ap.add_argument("--d2h", type=int, default=1,
                    help="c2: also time steps that download every image to pinned host memory")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(slices shard across GPUs with no communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c5":
        run_c5(args)
        return
    if args.workload == "dropin":
        run_dropin(args)
        return

    import torch
    from recview_amd import fbp

    dev = torch.device("cuda:0")
    t0 = time.time()
    odt = torch.int32 if args.interp == "lut16_i32" else torch.float32
    if args.workload == "c2":
        S, P, W = (args.slices or 64), args.proj, args.width
        sinos, fdeg, c_det = make_stack(S, P, W, 1002, dev, texture=args.texture)
        if args.interp == "lut16_i32":
            from recview_amd import phantom
            sinos = phantom.to_int16(sinos)
        dtype = "i16" if args.interp == "lut16_i32" else "f32"
        specs = [fbp.SliceSpec(W, fdeg, c_det, sino_dtype=dtype) for _ in range(S)]
        if args.tile < 0:
            # tile and angle chunk of the plan's mode from the tuned table
            from recview_amd import autotune
            mode = autotune.mode_of(args.interp, args.store, bool(args.lds_half), args.slices_per_item)
            if mode:
                autotune.apply(specs, [W] * S, mode)
        sino_list = [sinos[s] for s in range(S)]
        shapes = [(W, P)] * S
        desc = (f"C2: {S}-slice stack, {P} proj x {W} det -> {W}^2, {args.interp} interpolation, "
                f"{args.store} filtered sinogram")
    else:
        S = args.slices or 256
        specs, sino_list, shapes = make_c3(S, 1003, dev, args.interp, args.store, bool(args.lds_half),
                                           args.slices_per_item)
        desc = (f"C3: {S} uneven slices (W 512..4096, P 450..3600, 1/4 offset-axis 360-deg scans "
                f"stitched to 2W), per-shape tiles, {args.interp} {args.store}")
    torch.cuda.synchronize()
    log(f"inputs generated in {time.time() - t0:.1f}s")

    plan = fbp.FbpPlan(specs, interp=args.interp, store=args.store, persistent=bool(args.persistent),
                       tile=max(args.tile, 0), angles_per_chunk=args.chunk, lds_half=bool(args.lds_half),
                       slices_per_item=args.slices_per_item)
    outs = [torch.empty(plan.out_n(s), plan.out_n(s), dtype=odt, device=dev) for s in range(S)]
    stream = torch.cuda.current_stream()
    updates = float(sum(n * n * p for n, p in shapes))

    def step():
        plan.execute(sino_list, outs, stream=stream, graph=bool(args.graph))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Stage timing with HIP events on the launch stream (eager passes): filter,
    # pack, and the back-projection kernels alone (the roofline's kernel).
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    bp_ms, pack_ms = [], []
    for _ in range(max(1, min(3, args.steps))):
        ev[0].record(stream)
        plan.filter(sino_list, stream=stream)
        ev[1].record(stream)
        plan.pack(stream=stream)
        ev[2].record(stream)
        plan.backproject_packed(outs, stream=stream)
        ev[3].record(stream)
        torch.cuda.synchronize()
        pack_ms.append(ev[1].elapsed_time(ev[2]))
        bp_ms.append(ev[2].elapsed_time(ev[3]))
    filt_ms = ev[0].elapsed_time(ev[1])
    bp_ms = sum(bp_ms) / len(bp_ms)
    pack_ms = sum(pack_ms) / len(pack_ms)

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    slices_per_s = S * args.steps / elapsed
    tau_ns = elapsed / args.steps / updates * 1e9
    achieved = FLOP_PER_UPDATE * updates / (bp_ms * 1e-3) / 1e12
    for o in outs:
        if not torch.isfinite(o.float()).all():
            raise RuntimeError("non-finite image values")

    d2h = None
    if args.workload == "c2" and args.d2h:
        d2h = d2h_inclusive(plan, specs, sino_list, outs, args, stream, S, updates)

    cpu = None
    if not args.no_cpu and args.workload == "c2":
        model, mhz, flags, usable = host_cpu()
        # every usable host thread, up to the 16 that make one GPU's share of a GPU host
        threads = args.cpu_threads or min(16, usable)
        src = sino_list[0] if args.interp != "lut16_i32" else sino_list[0].float() / 3000.0
        secs, isa = cpu_baseline(src, fdeg, c_det, threads)
        N = W
        clock = f" @ {mhz:.0f} MHz" if mhz else ""
        cpu = {"value": round(1.0 / secs, 4), "unit": "slices/s", "cores": threads, "kind": "port",
               "sample": f"1 slice {P}x{W} -> {N}^2, int16 input, reference LUT16 path "
                         f"({isa} restatement of projx64, row bands + per-thread filtering, "
                         f"partial-image sum), tau {secs / (N * N * P) * 1e9:.4f} ns; host: {model}{clock}, "
                         f"{threads} threads of {usable} usable, ISA flags {' '.join(flags) or 'none'}"}

    traffic = None
    if args.workload == "c2" and args.interp == "linear" and args.store == "f32" and S == 64 \
            and args.proj == 1800 and args.width == 2048 and args.slices_per_item in (0, 2, 4) \
            and args.tile in (-1, 0) and args.chunk == 0:
        # the FP32 default runs four slices per work item; two per item has its own figure
        traffic = read_traffic("bp_linear_c2_traffic.md" if args.slices_per_item == 2
                               else "bp_linear4_c2_traffic.md")
    config = {"workload": desc, "slices": S, "interp": args.interp, "store": args.store,
              "persistent": bool(args.persistent), "graph": bool(args.graph),
              "lds_half": bool(args.lds_half), "slices_per_item": args.slices_per_item, "updates": updates}
    if args.workload == "c2":
        config.update({"projections": P, "detector": W, "image": W,
                       "tile": specs[0].tile if args.tile < 0 else args.tile,
                       "chunk": specs[0].chunk if args.tile < 0 else args.chunk})
    result = {
        "metric": METRIC,
        "value": round(slices_per_s, 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(REF_TAU_NS / tau_ns, 3) if args.workload == "c2" else None,
        "dtype": dtype_of(args),
        "data": "synthetic",
        "tau_ns": round(tau_ns, 7),
        "tau_bp_ns": round(bp_ms * 1e-3 / updates * 1e9, 7),
        "filter_ms": round(filt_ms, 3),
        "pack_ms": round(pack_ms, 3),
        "bp_ms": round(bp_ms, 3),
        "config": config,
        "d2h_inclusive": d2h,
        "roofline": roofline_of(args, S, achieved, traffic),
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
