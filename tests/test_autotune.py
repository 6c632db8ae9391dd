"""Autotuner lookups (CPU): the table is keyed by back-projection mode, and a
slice gets the tile and chunk of the nearest tuned shape of its own mode."""
from recview_amd import autotune, fbp

TABLE = {
    "linear2:2048:1800": {"tile": 1, "chunk": 16},
    "linear2:512:450": {"tile": 2, "chunk": 32},
    "h8:2048:1800": {"tile": 0, "chunk": 24},
    "h8:512:450": {"tile": 2, "chunk": 16},
    "lut16_i32:2048:1800": {"tile": 3, "chunk": 48},
    "linear:f32:2048:1800": {"tile": 2},           # a key of another layout is ignored
}


def test_lookup_per_mode():
    assert autotune.tile_for(2048, 1800, "linear2", TABLE) == 1
    assert autotune.chunk_for(2048, 1800, "linear2", TABLE) == 16
    assert autotune.tile_for(2048, 1800, "h8", TABLE) == 0
    assert autotune.chunk_for(2048, 1800, "h8", TABLE) == 24
    assert autotune.tile_for(2048, 1800, "lut16_i32", TABLE) == 3
    # nearest shape in log distance within the mode
    assert autotune.tile_for(600, 500, "linear2", TABLE) == 2
    assert autotune.chunk_for(4096, 3600, "h8", TABLE) == 24
    # an untuned mode keeps the plan's defaults
    assert autotune.tile_for(2048, 1800, "h4", TABLE) == 0
    assert autotune.chunk_for(2048, 1800, "h4", TABLE) == 0


def test_mode_of_plan_options():
    assert autotune.mode_of("linear", "f32") == "linear4"
    assert autotune.mode_of("linear", "f32", slices_per_item=4) == "linear4"
    assert autotune.mode_of("linear", "f32", slices_per_item=2) == "linear2"
    assert autotune.mode_of("linear", "f32", slices_per_item=1) == "linear"
    assert autotune.mode_of("linear", "f16", lds_half=True) == "h16"
    assert autotune.mode_of("linear", "f16", lds_half=True, slices_per_item=32) == "m32"
    assert autotune.mode_of("linear", "f16", lds_half=True, slices_per_item=8) == "h8"
    assert autotune.mode_of("linear", "bf16", lds_half=True, slices_per_item=4) == "h4"
    assert autotune.mode_of("lut16_i32") == "lut16_i32"
    assert autotune.mode_of("lut16_f32") is None


def test_apply_sets_specs():
    import numpy as np
    fd = np.arange(450, dtype=np.float32) * 0.4
    specs = ([fbp.SliceSpec(2048, np.arange(1800, dtype=np.float32) * 0.1, 1024.37) for _ in range(8)] +
             [fbp.SliceSpec(256, np.arange(901, dtype=np.float32) * 0.4, 60.0, offset_ct=True) for _ in range(8)] +
             [fbp.SliceSpec(512, fd, 256.37) for _ in range(8)])
    autotune.apply(specs, [2048] * 8 + [512] * 16, "h8", TABLE)
    assert (specs[0].tile, specs[0].chunk) == (0, 24)
    assert (specs[8].tile, specs[8].chunk) == (2, 16)      # 451 rows after stitching
    assert (specs[16].tile, specs[16].chunk) == (2, 16)
    assert specs[0].to_c().chunk == 24


def test_apply_picks_the_kernel_of_each_class_size():
    """Classes of one geometry take the entry of the kernel their size leads to:
    32 FP16 slices (matrix cores), sixteen, eight, four; FP32 pairs, a single (also
    in an FP16-pair plan, whose classes of fewer than four run the FP32 kernels)."""
    import numpy as np
    table = dict(TABLE)
    table["h4:2048:1800"] = {"tile": 3, "chunk": 32}
    table["linear:2048:1800"] = {"tile": 2, "chunk": 48}
    fd = np.arange(1800, dtype=np.float32) * 0.1
    mk = lambda c, n: [fbp.SliceSpec(2048, fd, c) for _ in range(n)]
    table["m32:2048:1800"] = {"tile": 2, "chunk": 12}
    table["h16:2048:1800"] = {"tile": 2, "chunk": 16}
    for mode in ("m32", "h16"):
        for n, want in ((40, (2, 12)), (16, (2, 16)), (9, (0, 24)), (5, (3, 32)), (3, (1, 16)), (1, (2, 48))):
            if mode == "h16" and n == 40:
                want = (2, 16)                     # the plan's limit is sixteen
            specs = autotune.apply(mk(1024.37, n), [2048] * n, mode, table)
            assert (specs[0].tile, specs[0].chunk) == want, (mode, n)
    specs = autotune.apply(mk(1024.37, 2) + mk(1030.0, 1), [2048] * 3, "linear2", table)
    assert [(s.tile, s.chunk) for s in specs] == [(1, 16), (1, 16), (2, 48)]
    # FP32 default (four per item): a class of four or more takes the four-slice entry
    table["linear4:2048:1800"] = {"tile": 0, "chunk": 8}
    specs = autotune.apply(mk(1024.37, 5) + mk(1030.0, 3) + mk(1040.0, 1), [2048] * 9, "linear4", table)
    assert [(s.tile, s.chunk) for s in specs] == [(0, 8)] * 5 + [(1, 16)] * 3 + [(2, 48)]


def test_apply_one_configuration_per_kernel_in_a_mixed_batch():
    """Slices of several shapes that run one kernel share its launch, so they get the
    (tile, chunk) of least summed tuned time; each shape alone keeps its own best."""
    import numpy as np
    table = {
        # 2048-wide: tile 1 chunk 16 best by a little; 512-wide: tile 3 chunk 8 best
        "linear2:2048:1800": {"tile": 1, "chunk": 16,
                              "all": {"0/16": 0.62, "1/16": 0.60, "3/8": 0.70, "0/8": 0.63}},
        "linear2:512:450": {"tile": 3, "chunk": 8,
                            "all": {"0/16": 0.013, "1/16": 0.014, "3/8": 0.011, "0/8": 0.012}},
        "linear:2048:1800": {"tile": 2, "chunk": 48},          # no timings: keeps its entry
    }
    big = lambda n: [fbp.SliceSpec(2048, np.arange(1800, dtype=np.float32) * 0.1, 1024.37)
                     for _ in range(n)]
    small = lambda n: [fbp.SliceSpec(512, np.arange(450, dtype=np.float32) * 0.4, 256.37)
                       for _ in range(n)]
    specs = autotune.apply(big(4), [2048] * 4, "linear2", table)
    assert {(s.tile, s.chunk) for s in specs} == {(1, 16)}
    specs = autotune.apply(small(4), [512] * 4, "linear2", table)
    assert {(s.tile, s.chunk) for s in specs} == {(3, 8)}
    # mixed: 4 x (0.62, 0.60, 0.70, 0.63) + 20 x (0.013, 0.014, 0.011, 0.012)
    #      = 2.74 (0/16), 2.68 (1/16), 3.02 (3/8), 2.76 (0/8) -> 1/16 for every pair slice
    specs = autotune.apply(big(4) + small(20), [2048] * 4 + [512] * 20, "linear2", table)
    assert {(s.tile, s.chunk) for s in specs} == {(1, 16)}
    # with many small slices the small shape's choice wins: 4 x 0.70 + 400 x 0.011 = 7.2
    # against 4 x 0.60 + 400 x 0.014 = 8.0 and 4 x 0.62 + 400 x 0.013 = 7.68
    specs = autotune.apply(big(4) + small(400), [2048] * 4 + [512] * 400, "linear2", table)
    assert {(s.tile, s.chunk) for s in specs} == {(3, 8)}
    # a single (class of one) runs the singles kernel, whose entry has no timings
    one = [fbp.SliceSpec(2048, np.arange(1800, dtype=np.float32) * 0.1, 1030.0)]
    specs = autotune.apply(big(2) + one, [2048] * 3, "linear2", table)
    assert [(s.tile, s.chunk) for s in specs] == [(1, 16), (1, 16), (2, 48)]


def test_shipped_table_is_per_mode():
    t = autotune.load()
    assert t, "tuned table missing"
    modes = {k.split(":")[0] for k in t}
    assert {"linear2", "h16", "h8", "lut16_i32"} <= modes
    for k, v in t.items():
        assert len(k.split(":")) == 3 and v["tile"] in (0, 1, 2, 3, 4) and v["chunk"] > 0
