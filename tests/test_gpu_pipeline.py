"""The whole reference pipeline on the GPU, on the reference's own 450-frame
angle schedule (docs/output450.log): scan log -> default frame exclusions ->
raw detector frames (synthetic: white, sample and dark counts of an analytic
phantom) -> sinogram generation (CudaSinogramHost) -> filtered
back-projection (CudaReconstHost) -> output stage -> 16-bit TIFF, each stage
checked bit-exact against the oracle's restatement."""
import os

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import logfile as LF
from recview_amd import output, phantom, reconst, sinogram

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_log_to_tif(tmp_path):
    log = LF.load_output_log(os.path.join(GOLD, "output450.log"))
    excl, _, _ = LF.frames_to_exclude(log)
    binc = LF.apply_exclusions(log, excl)
    W, dark = 256, 100
    c_user = W / 2 - 0.5 + 0.37                      # RECONST_INFO.center = c_user + 0.5
    ell = phantom.random_ellipses(450, count=8, mu_range=(0.0, 2.0 / W))
    p = phantom.sinogram(ell, W, log.fdeg, c_user + 0.5, scale=W).numpy()
    rng = np.random.default_rng(450)
    flat = rng.uniform(30000, 40000, W)
    frames = np.empty((len(log.fdeg) + 1, 1, W))
    for i in range(len(log.fdeg)):
        if binc[i] & LF.BINC_SAMPLE:
            frames[i, 0] = flat * np.exp(-p[i]) + dark
        else:
            frames[i, 0] = flat + dark
    frames[-1, 0] = dark
    frames = np.round(frames).astype(np.uint16)
    fexp = log.fexp
    fdeg = np.concatenate([log.fdeg, [0]]).astype(np.float32)
    binc_all = np.concatenate([binc, [0]]).astype(np.int8)

    sino, st = sinogram.generate(frames, binc_all, np.concatenate([fexp, [0]]), fdeg)
    ref_sino, rst = O.sinogram(frames, binc_all, np.concatenate([fexp, [0]]), fdeg)
    assert st == 0 and rst == 0 and np.array_equal(sino, ref_sino)

    rows = sino[:-1, 0]                               # frames x W, int16
    res = reconst.deconv_back_proj(rows, log.fdeg, c_user, binc=binc, normalise=False)
    use = (binc & LF.BINC_SAMPLE) != 0
    g = O.geometry(W, 2, c_user + 0.5)
    ref = O.recon_rint(rows[use], log.fdeg[use], g, O.set_filter(g.ndim, 0), formula=1, rule=1)
    assert res["nSinogr"] == int(use.sum())
    assert np.array_equal(res["image"], ref)

    u16, lo, hi = output.output_stage(torch.from_numpy(res["image"]).cuda(), c_user, 2)
    torch.cuda.synchronize()
    ref16, rlo, rhi = O.output_stage(ref, 1, c_user, 1)
    got16 = u16.cpu().numpy().view(np.uint16)
    assert (lo, hi) == (rlo, rhi) and np.array_equal(got16.astype(np.int32), ref16)
    desc = output.image_description(res["nSinogr"], 0.5, 1.0, lo, c_user, 0)
    output.write_tif16(tmp_path / "rec00000.tif", got16.astype(np.int32), desc)
    back, d, _ = output.read_tif16(tmp_path / "rec00000.tif")
    assert np.array_equal(back, got16) and d == desc
