"""Per-shape tile autotuner for the back-projection kernel.

For every (image side N, projections P) shape it times each tile configuration
(64x64, 64x128, 32x32, 128x64) on a batch large enough to fill the GPU, using
the BP stage alone (filtered rows uploaded once), and keeps the fastest.  The
table is stored next to this module (tuned_tiles.json) and looked up by
`tile_for`, nearest shape in log-space.  Plans then carry one tile per slice.
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuned_tiles.json")
TILES = (0, 1, 2, 3)
TILE_PIXELS = {0: 64 * 64, 1: 64 * 128, 2: 32 * 32, 3: 128 * 64}


def _key(N, P, interp, store):
    return f"{interp}:{store}:{N}:{P}"


def load():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def tile_for(N, P, interp="linear", store="f32", table=None):
    """Best tile id for a shape (nearest tuned shape of the same mode), or 0."""
    t = load() if table is None else table
    best, dist = 0, float("inf")
    for k, v in t.items():
        i, s, n, p = k.split(":")
        if i != interp or s != store:
            continue
        d = abs(math.log(int(n) / N)) + abs(math.log(int(p) / P))
        if d < dist:
            best, dist = v["tile"], d
    return best


def tune_shape(N, P, interp="linear", store="f32", reps=3, min_tiles=4096, device=0):
    """Times every tile configuration for slices of side N with P projections."""
    import numpy as np
    import torch
    from . import fbp, phantom
    fdeg = phantom.angles_deg(P)
    nsl = max(1, min(64, min_tiles // max(1, (N // 64) ** 2)))
    dev = torch.device("cuda", device)
    rows = torch.randn(P, N, device=dev) * 1e-3
    res = {}
    for tile in TILES:
        specs = [fbp.SliceSpec(N, fdeg, N / 2 + 0.37, tile=tile) for _ in range(nsl)]
        plan = fbp.FbpPlan(specs, interp=interp, store=store, persistent=True, tile=tile, device=device)
        for s in range(nsl):
            plan.set_filtered_rows(s, rows)
        outs = [torch.empty(N, N, device=dev) for _ in range(nsl)]
        plan.backproject(outs)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.backproject(outs)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[tile] = ms / nsl
        plan.close()
        del outs
    best = min(res, key=res.get)
    return {"tile": best, "ms_per_slice": {str(k): round(v, 4) for k, v in res.items()},
            "updates_per_s": N * N * P / (res[best] * 1e-3)}


def tune(shapes, interp="linear", store="f32", save=True, log=print):
    table = load()
    for N, P in shapes:
        r = tune_shape(N, P, interp, store)
        table[_key(N, P, interp, store)] = r
        log(f"tuned N={N} P={P}: tile {r['tile']} {r['ms_per_slice']}")
    if save:
        with open(TABLE, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return table


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="tune the back-projection tile per shape")
    ap.add_argument("--sides", default="512,1024,2048,4096,8192")
    ap.add_argument("--proj", default="450,900,1800,3600")
    ap.add_argument("--out", default=None, help="write the table here instead of the package")
    a = ap.parse_args()
    if a.out:
        TABLE = a.out
    shapes = [(int(n), int(p)) for n in a.sides.split(",") for p in a.proj.split(",")]
    tune(shapes)
