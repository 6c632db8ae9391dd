"""GPU parity of the native plan API (rvPlan*) against the CPU oracle.

Tolerances: integer modes are bit-exact; LUT16_F32 is within 1e-5 of the image
dynamic range of R-f32 (reference FP32 path); the production LINEAR mode is
compared with the FP64 oracle O-f64 (RMS and max stated relative to the
dynamic range inside the reconstruction circle).
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import fbp, phantom

pytestmark = pytest.mark.gpu


def _sino(W, P, seed, c_off=0.37, count=8, photons=None):
    c_det = W / 2 + c_off
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(seed, count=count)
    s = phantom.sinogram(ell, W, fdeg, c_det, scale=W)
    if photons:
        g = torch.Generator().manual_seed(seed)
        s = phantom.poisson(s, photons, generator=g).double()
    return s.float().numpy(), fdeg, c_det


def _circle(N, frac=0.45):
    yy, xx = np.mgrid[0:N, 0:N]
    return ((xx - N / 2) ** 2 + (yy - N / 2) ** 2) <= (frac * N) ** 2


def _errors(got, ref, mask):
    rng = ref[mask].max() - ref[mask].min()
    d = (got.astype(np.float64) - ref)[mask]
    return float(np.sqrt(np.mean(d * d)) / rng), float(np.abs(d).max() / rng)


def _run(specs, sinos, **kw):
    plan = fbp.FbpPlan(specs, **kw)
    outs = []
    for s in range(len(specs)):
        n = plan.out_n(s)
        odt = torch.int32 if kw.get("interp") == "lut16_i32" else torch.float32
        outs.append(torch.empty(n, n, dtype=odt, device="cuda"))
    plan.execute([x.cuda() for x in sinos], outs)
    torch.cuda.synchronize()
    return plan, [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("W,P,interp,formula,rule", [
    (64, 45, 2, 1, 1), (96, 60, 2, 0, 0), (64, 40, 3, 1, 0), (200, 90, 2, 0, 1)])
def test_plan_lut16_i32_bit_exact(W, P, interp, formula, rule):
    print(f"case W={W} P={P} interp={interp}", flush=True)
    s, fdeg, c = _sino(W, P, 300 + W)
    rows = phantom.to_int16(torch.from_numpy(s)).numpy()
    spec = fbp.SliceSpec(W, fdeg, c, iInterpolation=interp, sino_dtype="i16")
    _, (img,) = _run([spec], [torch.from_numpy(rows)], interp="lut16_i32", lut_formula=formula,
                     index_rule=rule)
    g = O.geometry(W, interp, c)
    ref = O.recon_rint(rows, fdeg, g, O.set_filter(g.ndim, 0), formula=formula, rule=rule)
    assert np.array_equal(img, ref)


@pytest.mark.parametrize("W,P,rule", [(128, 90, 0), (256, 180, 1)])
def test_plan_lut16_f32_vs_reference_fp32(W, P, rule):
    """LUT16 in float against R-f32: max |diff| <= 1e-5 of the dynamic range."""
    s, fdeg, c = _sino(W, P, 500 + W)
    spec = fbp.SliceSpec(W, fdeg, c)
    _, (img,) = _run([spec], [torch.from_numpy(s)], interp="lut16_f32", index_rule=rule)
    g = O.geometry(W, 2, c)
    ref = O.recon_rf32(s, fdeg, g, O.set_filter(g.ndim, 0), rule=rule)
    rms, mx = _errors(img, ref, _circle(W))
    print(f"LUT16_F32 vs R-f32 W={W} P={P}: rms {rms:.2e} max {mx:.2e} of range")
    assert mx <= 1e-5


def test_plan_filtered_rows_bit_exact():
    """The LDS FFT filter reproduces CFft's rounding: float rows identical to the oracle."""
    for W, P, interp in ((128, 20, 2), (300, 10, 2), (64, 8, 4)):
        s, fdeg, c = _sino(W, P, 700 + W)
        spec = fbp.SliceSpec(W, fdeg, c, iInterpolation=interp, filter="hamming")
        plan = fbp.FbpPlan([spec], exact_filter=True)
        plan.filter([torch.from_numpy(s).cuda()])
        torch.cuda.synchronize()
        got = plan.filtered_rows_tensor(0).cpu().numpy()
        g = O.geometry(W, interp, c)
        ref = O.filter_rows(s, g, O.set_filter(g.ndim, 1), kind=1)
        plan.close()
        assert np.array_equal(got, ref), f"W={W}: max {np.abs(got - ref).max()}"
        # default LINEAR plans filter two rows per transform with the symmetrised
        # filter: the same operator, so equal to the FP64 filter to FP32 rounding
        for store in ("f32", "f16"):
            plan = fbp.FbpPlan([spec], store=store)
            plan.filter([torch.from_numpy(s).cuda()])
            torch.cuda.synchronize()
            pair = plan.filtered_rows_tensor(0).cpu().numpy()
            plan.close()
            ref64 = O.filter_rows_f64(s, g, O.set_filter(g.ndim, 1))
            scale = np.abs(ref64).max()
            tol = 5e-6 if store == "f32" else 2e-3   # two rows share the rounding of one transform
            assert np.abs(pair - ref64).max() <= tol * scale, (W, store, np.abs(pair - ref64).max() / scale)
            assert np.abs(got - ref64).max() <= 5e-6 * scale      # CFft order itself: FP32 rounding


def test_plan_linear_bp_vs_fp64():
    """LINEAR back-projection of identical filtered rows vs FP64 linear interpolation."""
    W, P = 256, 200
    s, fdeg, c = _sino(W, P, 901)
    g = O.geometry(W, 2, c)
    px = O.filter_rows(s, g, O.set_filter(g.ndim, 0), kind=1)
    plan = fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c)])
    plan.set_filtered_rows(0, torch.from_numpy(px).cuda())
    out = torch.empty(W, W, device="cuda")
    plan.backproject([out])
    torch.cuda.synchronize()
    ref = O.bp_rows_linear_f64(px, fdeg, g)
    rms, mx = _errors(out.cpu().numpy(), ref, _circle(W))
    print(f"LINEAR BP vs FP64: rms {rms:.2e} max {mx:.2e}")
    assert rms <= 1e-6 and mx <= 1e-5


@pytest.mark.parametrize("store,lds_half,rms_tol", [("f32", False, 2e-6), ("f16", False, 1e-4),
                                                    ("bf16", False, 2e-3), ("f16", True, 1e-4),
                                                    ("bf16", True, 2e-3)])
def test_plan_linear_full_vs_fp64(store, lds_half, rms_tol):
    """Whole pipeline (filter + LINEAR BP) vs O-f64 on a noisy phantom; FP16/BF16
    filtered sinograms with per-row power-of-two scales, expanded to FP32 pairs in
    LDS or kept as FP16 pairs (per-chunk scale)."""
    W, P = 256, 180
    s, fdeg, c = _sino(W, P, 1234, photons=1e4)
    _, (img,) = _run([fbp.SliceSpec(W, fdeg, c)], [torch.from_numpy(s)], store=store, lds_half=lds_half)
    g = O.geometry(W, 2, c)
    ref = O.recon_of64(s, fdeg, g, O.set_filter(g.ndim, 0))
    rms, mx = _errors(img, ref, _circle(W))
    print(f"LINEAR store={store} vs O-f64: rms {rms:.2e} max {mx:.2e} of range")
    assert rms <= rms_tol


def test_plan_batch_uneven_equals_single():
    """Slices of different width / projection count in one plan give the same
    images as one plan per slice (the work list only reorders tiles)."""
    cases = [(64, 30), (96, 45), (128, 60), (80, 90)]
    specs, sinos = [], []
    for i, (W, P) in enumerate(cases):
        s, fdeg, c = _sino(W, P, 40 + i)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
    for persistent in (False, True):
        _, batch = _run(specs, sinos, persistent=persistent)
        for i in range(len(cases)):
            _, (one,) = _run([specs[i]], [sinos[i]])
            assert np.array_equal(batch[i], one)


@pytest.mark.parametrize("store,persistent,normalize", [("f32", True, False), ("f32", False, True),
                                                         ("f16", True, False), ("bf16", False, False)])
def test_plan_slice_groups_identical(store, persistent, normalize):
    """Slices of identical geometry back-projected two or four per work item give
    exactly the images of one slice per item: five equal-geometry slices (two pairs
    and one left over, or four and one) with different data, plus two slices of
    other geometries; four per item on every tile configuration."""
    print(f"case store={store} persistent={persistent} normalize={normalize}", flush=True)
    specs, sinos = [], []
    for i in range(5):
        s, fdeg, c = _sino(128, 90, 500 + i)
        specs.append(fbp.SliceSpec(128, fdeg, c))
        sinos.append(torch.from_numpy(s))
    for i, (W, P) in enumerate([(128, 91), (96, 90)]):
        s, fdeg, c = _sino(W, P, 510 + i)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
    kw = dict(store=store, persistent=persistent, normalize=normalize)
    # each tile configuration against one slice per item on the same tiles (the
    # FP32 coordinate is formed relative to the tile origin)
    singles = {}
    for k, tile in ((0, 0), (2, 0), (4, 0), (4, 1), (4, 2), (4, 3), (4, 4)):
        if tile not in singles:
            singles[tile] = _run(specs, sinos, slices_per_item=1, tile=tile, **kw)[1]
        _, grouped = _run(specs, sinos, slices_per_item=k, tile=tile, **kw)
        for i in range(len(specs)):
            assert np.array_equal(grouped[i], singles[tile][i]), (k, tile, i)
    single = singles[0]
    # the data of each slice stays with its own image
    assert not np.array_equal(single[0], single[1])


@pytest.mark.parametrize("store,dtype,exact", [("f32", "f32", False), ("bf16", "f32", False),
                                               ("f32", "i16", False), ("f32", "f32", True)])
def test_plan_filter_batches_equal_single(store, dtype, exact):
    """Slices of one filter class are filtered by batched launches (up to 16 per
    launch): 18 slices of one geometry, interleaved with two of another, give the
    images of one plan per slice."""
    print(f"case store={store} dtype={dtype} exact={exact}", flush=True)
    specs, sinos = [], []
    for i in range(20):
        W, P = (64, 40) if i not in (5, 11) else (80, 40)
        s, fdeg, c = _sino(W, P, 700 + i)
        if dtype == "i16":
            s = phantom.to_int16(torch.from_numpy(s)).numpy()
        specs.append(fbp.SliceSpec(W, fdeg, c, sino_dtype=dtype))
        sinos.append(torch.from_numpy(s))
    kw = dict(store=store, exact_filter=exact)
    _, batch = _run(specs, sinos, **kw)
    for i in range(len(specs)):
        _, (one,) = _run([specs[i]], [sinos[i]], **kw)
        assert np.array_equal(batch[i], one), i


@pytest.mark.parametrize("store,rms_tol", [("f16", 1e-4), ("bf16", 2e-3)])
def test_plan_half_pairs_grouped(store, rms_tol):
    """FP16 pairs in LDS (lds_half) with slices of one geometry back-projected two,
    four or eight (default) per work item (packed rows, one item-wide
    power-of-two scale, FP16 dot products; eight on 64^2 and 32^2 tiles):
    every image matches the FP64 oracle to the store's tolerance, and the one
    slice per item path to a fraction of it."""
    print(f"case store={store}", flush=True)
    W, P = 256, 180
    specs, sinos, raw = [], [], []
    for i in range(11):                              # groups of 8 / 4+4 / 2x5, a pair, one left
        s, fdeg, c = _sino(W, P, 1300 + i, photons=1e4)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
        raw.append((s, fdeg, c))
    _, single = _run(specs, sinos, store=store, lds_half=True, slices_per_item=1)
    m = _circle(W)
    refs = []
    for s, fdeg, c in raw:
        g = O.geometry(W, 2, c)
        refs.append(O.recon_of64(s, fdeg, g, O.set_filter(g.ndim, 0)))
    for k, tile in ((0, 0), (2, 0), (4, 0), (8, 2)):
        _, grouped = _run(specs, sinos, store=store, lds_half=True, slices_per_item=k, tile=tile)
        for i, ref in enumerate(refs):
            rms, _ = _errors(grouped[i], ref, m)
            rms1, _ = _errors(single[i], ref, m)
            diff, _ = _errors(grouped[i], single[i].astype(np.float64), m)
            print(f"k={k} tile={tile} slice {i}: grouped rms {rms:.2e}, single rms {rms1:.2e}, "
                  f"grouped - single {diff:.2e}")
            assert rms <= rms_tol and diff <= 0.5 * rms_tol, (k, tile, i, rms, diff)


@pytest.mark.parametrize("store,rms_tol", [("f16", 1e-4), ("bf16", 2e-3)])
def test_plan_half_sixteen_per_item(store, rms_tol):
    """Sixteen FP16-pair slices per work item (32x64 tiles, eight waves, and 32x32
    tiles): 19 slices of one geometry form a sixteen-slice item, a pair and a
    single.  Every image matches the FP64 oracle to the store's tolerance and the
    one slice per item path to a fraction of it, also through graph replay."""
    print(f"case store={store}", flush=True)
    W, P = 200, 150
    specs, sinos, refs = [], [], []
    for i in range(19):
        s, fdeg, c = _sino(W, P, 1900 + i, photons=1e4)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
        g = O.geometry(W, 2, c)
        refs.append(O.recon_of64(s, fdeg, g, O.set_filter(g.ndim, 0)))
    m = _circle(W)
    _, single = _run(specs, sinos, store=store, lds_half=True, slices_per_item=1)
    for tile in (4, 2):
        plan = fbp.FbpPlan(specs, store=store, lds_half=True, slices_per_item=16, tile=tile, persistent=True)
        outs = [torch.empty(W, W, device="cuda") for _ in specs]
        for graph in (False, True, True):
            for o in outs:
                o.fill_(float("nan"))
            plan.execute([x.cuda() for x in sinos], outs, graph=graph)
            torch.cuda.synchronize()
            for i, ref in enumerate(refs):
                got = outs[i].cpu().numpy()
                rms, _ = _errors(got, ref, m)
                diff, _ = _errors(got, single[i].astype(np.float64), m)
                assert rms <= rms_tol and diff <= 0.5 * rms_tol, (tile, graph, i, rms, diff)
        print(f"tile {tile}: worst rms {max(_errors(o.cpu().numpy(), r, m)[0] for o, r in zip(outs, refs)):.2e}")
        plan.close()


@pytest.mark.parametrize("store,rms_tol,persistent", [("f16", 1e-4, True), ("bf16", 2e-3, False)])
def test_plan_mfma_thirty_two_per_item(store, rms_tol, persistent):
    """Thirty-two FP16 slices per work item on the matrix cores (BP_LINEAR_M32,
    32x32 tiles): 35 slices of one geometry form one such item, a pair and a
    single; the image side (200) leaves the last tile row and column partly
    outside and 150 angles a partial last chunk.  Every image matches the FP64
    oracle to the store's tolerance and the one slice per item FP16 path to a
    fraction of it (inside the reconstruction circle: the matrix kernel
    interpolates the zero-extended row, the others hold the last sample), also
    through graph replay."""
    print(f"case store={store} persistent={persistent}", flush=True)
    W, P = 200, 150
    specs, sinos, refs = [], [], []
    for i in range(35):
        s, fdeg, c = _sino(W, P, 2100 + i, photons=1e4)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
        g = O.geometry(W, 2, c)
        refs.append(O.recon_of64(s, fdeg, g, O.set_filter(g.ndim, 0)))
    m = _circle(W)
    _, single = _run(specs, sinos, store=store, lds_half=True, slices_per_item=1)
    plan = fbp.FbpPlan(specs, store=store, lds_half=True, slices_per_item=32, persistent=persistent)
    assert plan.info(0)["tile_w"] == 32
    outs = [torch.empty(W, W, device="cuda") for _ in specs]
    worst = 0.0
    for graph in (False, True, True):
        for o in outs:
            o.fill_(float("nan"))
        plan.execute([x.cuda() for x in sinos], outs, graph=graph)
        torch.cuda.synchronize()
        for i, ref in enumerate(refs):
            got = outs[i].cpu().numpy()
            rms, _ = _errors(got, ref, m)
            diff, _ = _errors(got, single[i].astype(np.float64), m)
            worst = max(worst, rms)
            assert rms <= rms_tol and diff <= 0.5 * rms_tol, (graph, i, rms, diff)
    print(f"worst rms {worst:.2e}")
    plan.close()


@pytest.mark.parametrize("persistent,graph", [(True, False), (False, True)])
def test_plan_half_groups_mixed_geometries_and_tiles(persistent, graph):
    """One FP16-pair plan holding two geometries on different tiles: nine slices
    on 64^2 tiles (an eight-slice item and a single) and five on 64x128 tiles
    (eight per item is not built there: a four-slice item and a single), so one
    execution packs rows of several kinds.  Every image equals the one slice per
    item FP16 path to a fraction of the FP16 tolerance, through plain and graph
    execution."""
    print(f"case persistent={persistent} graph={graph}", flush=True)
    specs, sinos, refs = [], [], []
    for i in range(9):
        s, fdeg, c = _sino(128, 90, 1500 + i, photons=1e4)
        specs.append(fbp.SliceSpec(128, fdeg, c, tile=0))
        sinos.append(torch.from_numpy(s))
    for i in range(5):
        s, fdeg, c = _sino(96, 60, 1600 + i, photons=1e4)
        specs.append(fbp.SliceSpec(96, fdeg, c, tile=1))
        sinos.append(torch.from_numpy(s))
    _, single = _run(specs, sinos, store="f16", lds_half=True, slices_per_item=1, persistent=persistent)
    plan = fbp.FbpPlan(specs, store="f16", lds_half=True, persistent=persistent)
    outs = [torch.empty(plan.out_n(s), plan.out_n(s), device="cuda") for s in range(len(specs))]
    for _ in range(2):                                   # a second execution repacks the rows
        plan.execute([x.cuda() for x in sinos], outs, graph=graph)
        torch.cuda.synchronize()
        for i, o in enumerate(outs):
            got = o.cpu().numpy()
            m = _circle(got.shape[0])
            diff, _ = _errors(got, single[i].astype(np.float64), m)
            assert diff <= 5e-5, (i, diff)
    plan.close()


def test_plan_slices_per_item_options():
    """slices_per_item: values other than 0, 1, 2, 4, 8, 16, 32 are rejected at plan
    creation; FP32 plans clamp 4, 8, 16 and 32 to pairs (the FP32 pair kernel), so
    their images equal the default's exactly."""
    specs, sinos = [], []
    for i in range(4):
        s, fdeg, c = _sino(96, 60, 1700 + i)
        specs.append(fbp.SliceSpec(96, fdeg, c))
        sinos.append(torch.from_numpy(s))
    for bad in (3, 64, -1):
        with pytest.raises(RuntimeError):
            fbp.FbpPlan(specs, slices_per_item=bad)
    _, ref = _run(specs, sinos)
    for k in (4, 8, 16, 32):
        _, got = _run(specs, sinos, slices_per_item=k)
        for i in range(len(specs)):
            assert np.array_equal(got[i], ref[i]), (k, i)


def test_plan_backproject_only_grouped_slices():
    """Filtered rows uploaded per slice and back-projected on their own
    (rvPlanSetFilteredRows + rvPlanBackproject) go through the same packing of
    grouped slices: four slices of one geometry as two FP32 pairs give exactly
    the one slice per item images, and each agrees with the FP64 linear
    back-projection of its rows."""
    W, P = 128, 90
    specs, rows = [], []
    for i in range(4):
        s, fdeg, c = _sino(W, P, 1800 + i)
        g = O.geometry(W, 2, c)
        rows.append(O.filter_rows(s, g, O.set_filter(g.ndim, 0), kind=1))
        specs.append(fbp.SliceSpec(W, fdeg, c))
    imgs = {}
    for k in (0, 1):
        plan = fbp.FbpPlan(specs, slices_per_item=k)
        for i in range(4):
            plan.set_filtered_rows(i, torch.from_numpy(rows[i]).cuda())
        outs = [torch.empty(W, W, device="cuda") for _ in range(4)]
        plan.backproject(outs)
        torch.cuda.synchronize()
        imgs[k] = [o.cpu().numpy() for o in outs]
        plan.close()
    for i in range(4):
        assert np.array_equal(imgs[0][i], imgs[1][i]), i
        ref = O.bp_rows_linear_f64(rows[i], specs[i].fdeg, g)
        rms, mx = _errors(imgs[0][i], ref, _circle(W))
        assert rms <= 1e-6 and mx <= 1e-5, (i, rms, mx)


def test_plan_normalize_gives_attenuation():
    """normalize = 1 multiplies each image by 2 pi / P (per slice, also in a
    batch of different P), so the output is attenuation per detector sample."""
    cases = [(96, 90), (96, 180)]
    specs, sinos = [], []
    for W, P in cases:
        s, fdeg, c = _sino(W, P, 60)          # one phantom, two projection counts
        specs.append(fbp.SliceSpec(W, fdeg, c, filter="ramp"))
        sinos.append(torch.from_numpy(s))
    _, raw = _run(specs, sinos, persistent=True)
    _, norm = _run(specs, sinos, persistent=True, normalize=True)
    for (W, P), r, n in zip(cases, raw, norm):
        np.testing.assert_allclose(n, r * np.float32(2 * np.pi / P), rtol=1e-6, atol=1e-7 * np.abs(r).max())
    # the two projection counts agree in physical units (inner disc, offset-free)
    m = _circle(96, 0.35)
    a, b = norm[0][m], norm[1][m]
    assert np.abs(a - b).mean() < 0.1 * np.abs(b).mean()


def test_plan_offset_ct_and_mixed_tiles():
    """An offset-CT slice inside a plan (GPU stitch -> filter -> BP) equals the
    stitch followed by an ordinary plan on the stitched rows with the centre remap
    of gazoDoc.cpp:3832-3837; slices with different tiles in one plan equal their
    single-slice plans."""
    from recview_amd import offset
    W, P = 96, 60
    for c_det in (0.8 * W + 0.3, 0.2 * W + 0.6):
        fdeg = (np.arange(2 * P + 1) * (180.0 / P)).astype(np.float32)
        ell = phantom.random_ellipses(5, count=6, axis_range=(0.1, 0.3))
        s360 = phantom.sinogram(ell, W, fdeg, c_det, scale=1.2 * W).float()
        _, (img_plan,) = _run([fbp.SliceSpec(W, fdeg, c_det, offset_ct=True)], [s360])
        stitched, ang = offset.stitch(s360.cuda(), fdeg, c_det - 0.5)
        c_user = offset.stitched_center(c_det - 0.5, W)
        _, (img_ref,) = _run([fbp.SliceSpec(2 * W, ang, c_user + 0.5)], [stitched.cpu()])
        assert img_plan.shape == (2 * W, 2 * W)
        assert np.array_equal(img_plan, img_ref)
    specs, sinos = [], []
    for i, (Wi, Pi, tile) in enumerate([(64, 30, 0), (96, 45, 2), (128, 60, 3), (80, 40, 1)]):
        s, fd, c = _sino(Wi, Pi, 90 + i)
        specs.append(fbp.SliceSpec(Wi, fd, c, tile=tile))
        sinos.append(torch.from_numpy(s))
    _, batch = _run(specs, sinos, persistent=True)
    for i in range(4):
        _, (one,) = _run([specs[i]], [sinos[i]])
        assert np.array_equal(batch[i], one)


def test_plan_graph_replay():
    """hipGraph replays (kernel nodes only) equal the eager result every time,
    with eager work between replays."""
    specs, sinos = [], []
    for i, (W, P) in enumerate([(64, 40), (128, 50)]):
        s, fdeg, c = _sino(W, P, 60 + i)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s).cuda())
    plan = fbp.FbpPlan(specs, persistent=True)
    outs = [torch.empty(plan.out_n(i), plan.out_n(i), device="cuda") for i in range(2)]
    plan.execute(sinos, outs)
    torch.cuda.synchronize()
    eager = [o.clone() for o in outs]
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        plan.execute(sinos, outs, graph=True)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()       # eager work between replays
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), f"replay {rep}"
        del junk
    # eager executions into other images between replays, on another stream: each
    # replay writes its own images again, and no eager result lands in them
    side = torch.cuda.Stream()
    others = [torch.empty_like(o) for o in outs]
    for rep in range(3):
        for o in outs + others:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            plan.execute(sinos, others, stream=side)
        plan.execute(sinos, outs, graph=True, stream=side)
        with torch.cuda.stream(side):
            plan.execute(sinos, others, stream=side)
        torch.cuda.synchronize()
        for o, x, e in zip(outs, others, eager):
            assert torch.equal(o, e), f"replay {rep} after eager executions"
            assert torch.equal(x, e), f"eager execution {rep} between replays"
    # a plan whose very first execution is a capture (nothing cached beforehand)
    spec = fbp.SliceSpec(96, phantom.angles_deg(30), 48.3)
    s, _, _ = _sino(96, 30, 61, c_off=0.3)
    plan2 = fbp.FbpPlan([spec], persistent=True)
    o2 = torch.empty(96, 96, device="cuda")
    plan2.execute([torch.from_numpy(s).cuda()], [o2], graph=True)
    torch.cuda.synchronize()
    _, (ref,) = _run([spec], [torch.from_numpy(s)])
    assert np.array_equal(o2.cpu().numpy(), ref)


def test_plan_tiles_agree():
    """Every tile configuration gives the same LUT16 integer image (bit-exact)
    and the same LINEAR image to FP32 rounding."""
    W, P = 160, 64
    s, fdeg, c = _sino(W, P, 77)
    rows = phantom.to_int16(torch.from_numpy(s))
    ref_i = ref_f = None
    for tile in range(5):
        _, (im,) = _run([fbp.SliceSpec(W, fdeg, c, sino_dtype="i16")], [rows], interp="lut16_i32",
                        tile=tile)
        _, (fl,) = _run([fbp.SliceSpec(W, fdeg, c)], [torch.from_numpy(s)], tile=tile)
        if ref_i is None:
            ref_i, ref_f = im, fl
        assert np.array_equal(im, ref_i)
        rms, mx = _errors(fl, ref_f.astype(np.float64), _circle(W))
        assert mx <= 1e-5


def test_plan_zoomed_roi():
    """Off-centre zoomed sub-region: pitch 1 with an integer origin equals the crop
    of the full grid; pitch 1/2 coincides with the full grid on even pixels."""
    W, P = 128, 90
    s, fdeg, c = _sino(W, P, 88)
    _, (full,) = _run([fbp.SliceSpec(W, fdeg, c)], [torch.from_numpy(s)])
    x0, y0 = 40, 24
    _, (roi1,) = _run([fbp.SliceSpec(W, fdeg, c, out_n=64, pix_zoom=1.0, x0=x0, y0=y0)],
                      [torch.from_numpy(s)])
    crop = full[y0:y0 + 64, x0:x0 + 64]
    rng = full.max() - full.min()
    assert np.abs(roi1 - crop).max() <= 1e-5 * rng
    _, (roi2,) = _run([fbp.SliceSpec(W, fdeg, c, out_n=64, pix_zoom=2.0, x0=x0, y0=y0)],
                      [torch.from_numpy(s)])
    assert np.abs(roi2[::2, ::2] - full[y0:y0 + 32, x0:x0 + 32]).max() <= 1e-5 * rng


def test_centre_sweep_filters_once_and_finds_axis():
    """Batched axis scan from one filtered sinogram: the rows are filtered at one
    centre and every centre back-projects them shifted by its integer centre
    offset, as the reference's kept-sinogram path does (cudaReconstHost.cpp:775-777).
    Each image equals FP64 filtering at that centre + FP64 back-projection with the
    offset; centres with the filtering centre's integer part equal their own plan bit
    for bit; the total variation is smallest at the true rotation axis."""
    W, P = 192, 180
    s, fdeg, c = _sino(W, P, 321, c_off=2.6)
    sino = torch.from_numpy(s).cuda()
    centres = [c + d for d in (-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0)]
    sweep = fbp.centre_sweep(sino, fdeg, centres).cpu().numpy()
    g_src = O.geometry(W, 2, centres[3])                 # the median integer part: 98.6
    px64 = O.filter_rows_f64(s, g_src, O.set_filter(g_src.ndim, 0))
    m = _circle(W)
    for k, ck in enumerate(centres):
        g = O.geometry(W, 2, ck)
        ref = O.bp_rows_linear_f64(px64, fdeg, g, icoff=g.ic - g_src.ic)
        rms, _ = _errors(sweep[k], ref, m)
        assert rms <= 2e-6, (k, rms)
        if g.ic == g_src.ic:
            _, (one,) = _run([fbp.SliceSpec(W, fdeg, ck)], [torch.from_numpy(s)], persistent=True)
            assert np.array_equal(sweep[k], one), k
    m4 = _circle(W, 0.4)
    tv = [np.abs(np.diff(img, axis=0))[m4[1:]].sum() + np.abs(np.diff(img, axis=1))[m4[:, 1:]].sum()
          for img in sweep]
    assert int(np.argmin(tv)) == 3
    # sharing rows needs the same detector and projections
    with pytest.raises(RuntimeError, match="rows_of"):
        fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c), fbp.SliceSpec(W + 8, fdeg, c, rows_from=0)])
    with pytest.raises(RuntimeError, match="rows_of"):
        fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c, rows_from=0)])


def test_plan_edge_cases():
    """Degenerate and ragged inputs: one projection, the narrowest detector the
    transform supports, a slice wider than tall batches, and clean errors."""
    # one projection: the image is that filtered row smeared along the ray
    s, fdeg, c = _sino(64, 1, 11)
    _, (img,) = _run([fbp.SliceSpec(64, fdeg, c)], [torch.from_numpy(s)])
    assert np.isfinite(img).all() and np.allclose(img, img[0:1, :], atol=1e-6 * np.abs(img).max() + 1e-12)
    # W = 4 (n = 8 point transform) and a mixed batch with it
    s4, f4, c4 = _sino(4, 6, 12)
    s9, f9, c9 = _sino(96, 7, 13)
    _, outs = _run([fbp.SliceSpec(4, f4, c4), fbp.SliceSpec(96, f9, c9)], [torch.from_numpy(s4), torch.from_numpy(s9)],
                   persistent=True)
    assert outs[0].shape == (4, 4) and np.isfinite(outs[0]).all() and np.isfinite(outs[1]).all()
    # errors are raised, not crashes: no slices, width below the supported range
    with pytest.raises(RuntimeError):
        fbp.FbpPlan([])
    with pytest.raises(RuntimeError):
        fbp.FbpPlan([fbp.SliceSpec(3, f4, 1.5)])
    # wrong number of sinograms for the plan
    plan = fbp.FbpPlan([fbp.SliceSpec(96, f9, c9)])
    with pytest.raises(Exception):
        plan.execute([], [])
    plan.close()


@pytest.mark.parametrize("W,interp", [(600, 2), (1000, 2), (2000, 2), (3000, 2), (600, 3), (1500, 3)])
def test_plan_filtered_rows_non_pow2(W, interp):
    """Widths that are not powers of two keep more than half of the transform
    (n/2 < N < n): every cropped column of both rows of a pair, in FP32 and in
    the 16-bit stores with their per-row scales, against the FP64 filter; an odd
    row count leaves the last transform with one row."""
    P = 13
    s, fdeg, c = _sino(W, P, 1700 + W + interp)
    g = O.geometry(W, interp, c)
    assert g.ndim // 2 < g.N < g.ndim
    ref64 = O.filter_rows_f64(s, g, O.set_filter(g.ndim, 0))
    rmax = np.abs(ref64).max(axis=1, keepdims=True)
    spec = fbp.SliceSpec(W, fdeg, c, iInterpolation=interp)
    for exact in (False, True):
        # FP32: rounding of the shared transform relative to the rows' magnitudes (a row
        # with a small maximum next to a large partner sees its partner's rounding);
        # a column left unwritten errs by the order of the row maximum itself
        for store, tol in (("f32", 1e-4), ("f16", 1e-3), ("bf16", 6e-3)):
            if exact and store == "f32":
                continue
            plan = fbp.FbpPlan([spec], store=store, exact_filter=exact)
            plan.filter([torch.from_numpy(s).cuda()])
            torch.cuda.synchronize()
            got = plan.filtered_rows_tensor(0).cpu().numpy()
            plan.close()
            err = np.abs(got - ref64) / rmax
            print(f"W={W} z={g.z} n={g.ndim} N={g.N} store={store} exact={exact}: "
                  f"max {err.max():.2e} of the row maximum (column {int(err.max(axis=0).argmax())})")
            assert err.max() <= tol, (W, store, exact)
            if store == "f32":
                # as accurate as the reference's own FP32 filter (CFft order) against FP64
                e_ref = np.abs(O.filter_rows(s, g, O.set_filter(g.ndim, 0), kind=1) - ref64).max()
                assert np.abs(got - ref64).max() <= 4 * e_ref, (np.abs(got - ref64).max(), e_ref)


def test_plan_pair_filter_every_transform_length():
    """The paired-row filter at every transform length it runs in LDS, 2^3 .. 2^14
    points (radix-16 passes, then a radix-8, -4 or -2 pass, or none), both rows of
    each pair against the FP64 filter and as accurate as the reference's own FP32
    filter; an odd row count leaves the last transform with one row."""
    P = 7
    for W in (5, 9, 20, 40, 70, 130, 260, 520, 1040, 2080, 4100, 8192):
        s, fdeg, c = _sino(W, P, 2300 + W)
        g = O.geometry(W, 2, c)
        ref64 = O.filter_rows_f64(s, g, O.set_filter(g.ndim, 0))
        rmax = np.abs(ref64).max(axis=1, keepdims=True)
        plan = fbp.FbpPlan([fbp.SliceSpec(W, fdeg, c)], store="f32")
        plan.filter([torch.from_numpy(s).cuda()])
        torch.cuda.synchronize()
        got = plan.filtered_rows_tensor(0).cpu().numpy()
        plan.close()
        err = np.abs(got - ref64) / rmax
        e_ref = np.abs(O.filter_rows(s, g, O.set_filter(g.ndim, 0), kind=1) - ref64).max()
        print(f"W={W} n={g.ndim}: max {err.max():.2e} of the row maximum, abs "
              f"{np.abs(got - ref64).max():.2e} vs reference FP32 {e_ref:.2e}")
        assert err.max() <= 1e-4, W
        assert np.abs(got - ref64).max() <= 4 * e_ref, W


@pytest.mark.parametrize("W,nsl,store,lds_half,rms_tol", [
    (1000, 8, "f16", True, 1e-4), (1000, 8, "bf16", True, 2e-3), (1000, 4, "f16", False, 1e-4),
    (2000, 2, "f16", True, 1e-4), (2000, 2, "bf16", False, 2e-3), (2000, 2, "f32", False, 2e-6),
    (3000, 2, "f16", True, 1e-4)])
def test_plan_non_pow2_grouped_vs_fp64(W, nsl, store, lds_half, rms_tol):
    """Whole pipeline at detector widths that are not powers of two, slices of one
    geometry grouped per work item (eight / two FP16 pairs, FP32 pairs), against
    the FP64 oracle inside the reconstruction circle."""
    P = 96
    specs, sinos = [], []
    for i in range(nsl):
        s, fdeg, c = _sino(W, P, 1800 + 31 * i + W, photons=1e4)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
    _, imgs = _run(specs, sinos, store=store, lds_half=lds_half)
    g = O.geometry(W, 2, c)
    mask = _circle(W)
    for i in range(nsl):
        ref = O.recon_of64(sinos[i].numpy(), fdeg, g, O.set_filter(g.ndim, 0))
        rms, mx = _errors(imgs[i], ref, mask)
        print(f"W={W} slice {i}/{nsl} store={store} lds_half={lds_half}: rms {rms:.2e} max {mx:.2e}")
        assert rms <= rms_tol, (W, i, store)


@pytest.mark.parametrize("store,lds_half", [("f32", False), ("f16", True)])
def test_plan_pack_then_backproject_packed(store, lds_half):
    """rvPlanPack + rvPlanBackprojectPacked (the two halves the benchmark times
    apart) give exactly the images of rvPlanBackproject: five slices of one
    geometry (grouped, packed rows) and one of another (single)."""
    specs, sinos = [], []
    for i in range(6):
        W, P = (128, 90) if i < 5 else (96, 60)
        s, fdeg, c = _sino(W, P, 2300 + i)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s).cuda())
    plan = fbp.FbpPlan(specs, store=store, lds_half=lds_half, persistent=True)
    plan.filter(sinos)
    a = [torch.empty(plan.out_n(i), plan.out_n(i), device="cuda") for i in range(6)]
    b = [torch.full_like(x, float("nan")) for x in a]
    plan.backproject(a)
    plan.pack()
    plan.backproject_packed(b)
    torch.cuda.synchronize()
    plan.close()
    for i in range(6):
        assert torch.equal(a[i], b[i]), i


def test_plan_geometry_tables_equal_per_chunk_geometry(monkeypatch):
    """FP32 pairs read their tile geometry from tables built at plan creation; with
    the tables switched off (RECVIEW_GEOMETRY_TABLES=0, the path plans take beyond
    the table budget) the kernel forms it per chunk.  Both give the same images bit
    for bit: four slices of one geometry (two pairs, or one item of four) on each
    tile configuration, persistent and one block per item, and a pair and a single
    of another geometry."""
    specs, sinos = [], []
    for i in range(7):
        W, P = (160, 97) if i < 4 else (96, 60)
        s, fdeg, c = _sino(W, P, 2400 + i)
        specs.append(fbp.SliceSpec(W, fdeg, c))
        sinos.append(torch.from_numpy(s))
    for tile, persistent, spi in [(t, p, k) for t in range(5) for p in (True, False) for k in (0, 4)]:
        kw = dict(tile=tile, persistent=persistent, slices_per_item=spi)
        monkeypatch.setenv("RECVIEW_GEOMETRY_TABLES", "0")
        _, per_chunk = _run(specs, sinos, **kw)
        monkeypatch.delenv("RECVIEW_GEOMETRY_TABLES")
        _, tables = _run(specs, sinos, **kw)
        for i in range(len(specs)):
            assert np.array_equal(per_chunk[i], tables[i]), (tile, persistent, spi, i)
