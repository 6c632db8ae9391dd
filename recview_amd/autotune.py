"""Per-shape autotuner of the back-projection launch: tile and angle chunk per mode.

The plan picks one back-projection kernel per work-item kind (the mode):

  linear4    FP32 filtered rows, four slices of one geometry per work item (default)
  linear2    FP32, two slices per work item (a class's pair left over by fours)
  linear     FP32, one slice per work item (slices without a partner of equal geometry)
  m32        FP16 filtered rows, 32 slices per work item on the matrix cores (32x32
             tiles; opt-in, slices_per_item = 32)
  h16        FP16 filtered rows and FP16 pairs in LDS, sixteen slices per work item
  h8         the same with eight slices per work item
  h4         the same with four slices per work item
  lut16_i32  the reference LUT16 in int32 (the drop-in semantics)

For every (mode, image side N, projections P) the tuner times each valid tile
configuration (64x64, 64x128, 32x32, 128x64, 32x64; eight-slice items exist on
64x64, 32x32 and 32x64, sixteen-slice items on 32x32 and 32x64, 32-slice items on
32x32 only) and each candidate chunk on a batch of at least 16384 64x64-tile items
(up to 64 slices: ten or more rounds of the resident blocks, so the last, partly
filled round does not decide the choice), back-projection stage only (pack
included), and keeps the fastest.  The table (tuned_tiles.json
next to this module) is keyed "mode:N:P"; `tile_for` / `chunk_for` return the
entry of the nearest tuned shape of the same mode (log distance), and `apply`
writes them into slice specs before a plan is built.
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuned_tiles.json")

MODES = ("linear4", "linear2", "linear", "m32", "h16", "h8", "h4", "lut16_i32")
TILES = {"linear4": (0, 1, 2, 3, 4), "linear2": (0, 1, 2, 3, 4), "linear": (0, 1, 2, 3, 4), "m32": (2,),
         "h16": (2, 4), "h8": (0, 2, 4), "h4": (0, 1, 2, 3, 4), "lut16_i32": (0, 1, 2, 3, 4)}
CHUNKS = {"linear4": (8, 12, 16), "linear2": (8, 16, 24, 32), "linear": (16, 32, 48), "m32": (8, 12, 16),
          "h16": (8, 12, 16), "h8": (16, 24, 32), "h4": (16, 24, 32), "lut16_i32": (16, 32, 48)}
# plan options of each mode (the slices of a tuning batch share one geometry, so
# they group as the mode says)
PLAN_KW = {
    "linear4": dict(interp="linear", store="f32", slices_per_item=4),
    "linear2": dict(interp="linear", store="f32", slices_per_item=2),
    "linear": dict(interp="linear", store="f32", slices_per_item=1),
    "m32": dict(interp="linear", store="f16", lds_half=True, slices_per_item=32),
    "h16": dict(interp="linear", store="f16", lds_half=True, slices_per_item=16),
    "h8": dict(interp="linear", store="f16", lds_half=True, slices_per_item=8),
    "h4": dict(interp="linear", store="f16", lds_half=True, slices_per_item=4),
    "lut16_i32": dict(interp="lut16_i32"),
}
GROUP = {"linear4": 4, "linear2": 2, "linear": 1, "m32": 32, "h16": 16, "h8": 8, "h4": 4, "lut16_i32": 1}


def mode_of(interp="linear", store="f32", lds_half=False, slices_per_item=0):
    """The mode a plan runs for slices of one geometry under these options."""
    if interp == "lut16_i32":
        return "lut16_i32"
    if interp != "linear":
        return None
    if store != "f32" and lds_half:
        return {4: "h4", 8: "h8", 32: "m32"}.get(slices_per_item, "h16")
    if store == "f32":
        return {1: "linear", 2: "linear2"}.get(slices_per_item, "linear4")
    return None


def load(path=None):
    path = path or TABLE
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _nearest(N, P, mode, table):
    best, dist = None, float("inf")
    for k, v in table.items():
        parts = k.split(":")
        if len(parts) != 3 or parts[0] != mode:
            continue
        d = abs(math.log(int(parts[1]) / N)) + abs(math.log(int(parts[2]) / P))
        if d < dist:
            best, dist = v, d
    return best


def tile_for(N, P, mode="linear2", table=None):
    """Tuned tile id for the nearest shape of this mode (0 when the mode is untuned)."""
    e = _nearest(N, P, mode, load() if table is None else table)
    return int(e["tile"]) if e else 0


def chunk_for(N, P, mode="linear2", table=None):
    """Tuned angles per chunk for the nearest shape of this mode (0 = the plan's default)."""
    e = _nearest(N, P, mode, load() if table is None else table)
    return int(e.get("chunk", 0)) if e else 0


def _class_mode(mode, count):
    """The kernel a class of `count` slices of one geometry runs under a plan mode
    (an FP16-pair plan runs a class of fewer than four on the FP32-pair kernels)."""
    if mode in ("m32", "h16", "h8", "h4"):
        top = GROUP[mode]
        for m in ("m32", "h16", "h8", "h4"):
            if GROUP[m] <= top and count >= GROUP[m]:
                return m
        return "linear2" if count >= 2 else "linear"
    if mode == "linear4":
        return "linear4" if count >= 4 else "linear2" if count >= 2 else "linear"
    if mode == "linear2":
        return "linear2" if count >= 2 else "linear"
    return mode


def apply(specs, sides, mode, table=None):
    """Sets tile and chunk of every slice spec from the table: sides[i] is the image
    side slice i reconstructs (2W for an offset-axis scan), its projection count the
    rows it reconstructs.  Slices of one geometry form a class; a class takes the
    entry of the kernel its size leads the plan to (32 FP16 slices per item on the
    matrix cores, sixteen, eight or four FP16 pairs, four or two FP32 slices, singles; an
    FP16-pair plan runs the pairs and singles of its classes smaller than four on
    the FP32-pair kernels).

    A plan runs one launch per (tile, kernel), with the angle chunk of the first
    slice it groups there, so a batch of several shapes gets one (tile, chunk) per
    kernel: the configuration with the least summed tuned time over the slices that
    kernel runs (each slice priced by its nearest entry's timing of every
    configuration).  For a batch of one shape this is that shape's best entry; an
    entry without the per-configuration timings keeps its own tile and chunk."""
    t = load() if table is None else table
    key = lambda s: (s.W, s.P, s.center, s.iInterpolation, s.offset_ct, s.out_n, s.tilt,
                     s.fdeg.tobytes())
    counts = {}
    for spec in specs:
        counts[key(spec)] = counts.get(key(spec), 0) + 1
    per_kernel = {}
    for spec, N in zip(specs, sides):
        P = (spec.P + 1) // 2 if spec.offset_ct else spec.P
        km = _class_mode(mode, counts[key(spec)])
        e = _nearest(N, P, km, t)
        if e:
            spec.tile = int(e["tile"])
            spec.chunk = int(e.get("chunk", 0))
            per_kernel.setdefault(km, []).append((spec, e))
    for members in per_kernel.values():
        grids = [e.get("all") for _, e in members]
        if not all(grids):
            continue
        common = set.intersection(*(set(g) for g in grids))
        if not common:
            continue
        best = min(sorted(common), key=lambda c: sum(g[c] for g in grids))
        tile, chunk = (int(v) for v in best.split("/"))
        for spec, _ in members:
            spec.tile, spec.chunk = tile, chunk
    return specs


def tune_shape(N, P, mode, reps=3, min_items=16384, device=0, log=None):
    """Times every (tile, chunk) of a mode for slices of side N with P projections."""
    import torch
    from . import fbp, phantom
    fdeg = phantom.angles_deg(P)
    group = GROUP[mode]
    tiles_per_slice = max(1, (N // 64) ** 2)
    items = -(-min_items // tiles_per_slice)                  # work items to fill the GPU
    nsl = group * max(1, min(64 // group, items))
    dev = torch.device("cuda", device)
    kw = PLAN_KW[mode]
    odt = torch.int32 if mode == "lut16_i32" else torch.float32
    outs = [torch.empty(N, N, device=dev, dtype=odt) for _ in range(nsl)]
    sino = (torch.rand(P, N, device=dev) * 0.01).contiguous()
    if mode == "lut16_i32":
        sino = (sino * 3000).to(torch.int16)
    res = {}
    for tile in TILES[mode]:
        for chunk in CHUNKS[mode]:
            specs = [fbp.SliceSpec(N, fdeg, N / 2 + 0.37, tile=tile, chunk=chunk,
                                   sino_dtype="i16" if mode == "lut16_i32" else "f32") for _ in range(nsl)]
            plan = fbp.FbpPlan(specs, persistent=True, tile=tile, device=device, **kw)
            plan.filter([sino] * nsl)
            plan.backproject(outs)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.backproject(outs)
            e1.record()
            torch.cuda.synchronize()
            res[(tile, chunk)] = e0.elapsed_time(e1) / reps / nsl
            plan.close()
    (bt, bc) = min(res, key=res.get)
    if log:
        log(f"  {mode} N={N} P={P}: tile {bt} chunk {bc} {res[(bt, bc)]:.4f} ms/slice")
    return {"tile": bt, "chunk": bc, "ms_per_slice": round(res[(bt, bc)], 5),
            "updates_per_s": N * N * P / (res[(bt, bc)] * 1e-3),
            "all": {f"{t}/{c}": round(v, 5) for (t, c), v in sorted(res.items())}}


def _log(msg):
    print(msg, flush=True)


def tune(shapes, modes=MODES, path=None, log=_log, min_items=16384):
    path = path or TABLE
    table = load(path)
    for mode in modes:
        for N, P in shapes:
            table[f"{mode}:{N}:{P}"] = tune_shape(N, P, mode, min_items=min_items, log=log)
            with open(path, "w") as f:          # written as it goes (a long sweep keeps its results)
                json.dump(table, f, indent=1, sort_keys=True)
    return table


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune the back-projection tile and chunk per mode and shape")
    ap.add_argument("--sides", default="512,1024,2048,4096,8192")
    ap.add_argument("--proj", default="450,900,1800,3600")
    ap.add_argument("--modes", default=",".join(MODES))
    ap.add_argument("--out", default=None, help="write the table here instead of the package")
    ap.add_argument("--min-items", type=int, default=16384,
                    help="64x64-tile work items a tuning batch holds at least (capped at 64 slices)")
    a = ap.parse_args()
    shapes = [(int(n), int(p)) for n in a.sides.split(",") for p in a.proj.split(",")]
    tune(shapes, modes=a.modes.split(","), path=a.out, min_items=a.min_items)
