"""RecviewOutputStage (GPU) bit-exact against the oracle's restatement of the
caller-side output stage (gazoDoc.cpp:3993-4021), and a slice written as
RecView's 16-bit TIFF."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import output

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,interp,center,lo,hi", [(256, 2, 128.37, -5000, 90000), (255, 3, 64.2, 10, 400),
                                                   (512, 1, 540.0, -70000, -100), (128, 2, 64.0, 7, 7)])
def test_output_stage_bit_exact(N, interp, center, lo, hi):
    rng = np.random.default_rng(N + interp)
    img = rng.integers(lo, hi + 1, (N, N)).astype(np.int32)
    ibin = 4 if interp == 0 else (2 if interp == 1 else 1)
    z = 2 ** (interp - 2) if interp > 2 else 1
    ref, rmin, rmax = O.output_stage(img, z, center, ibin)
    got, gmin, gmax = output.output_stage(torch.from_numpy(img).cuda(), center, interp)
    torch.cuda.synchronize()
    g = got.cpu().numpy().view(np.uint16).astype(np.int32)
    assert (gmin, gmax) == (rmin, rmax)
    assert np.array_equal(g, ref)
    # float images with the same integer values take the same path
    gf, fmin, fmax = output.output_stage(torch.from_numpy(img.astype(np.float32)).cuda(), center, interp)
    torch.cuda.synchronize()
    if abs(lo) < 2 ** 24 and abs(hi) < 2 ** 24:
        assert np.array_equal(gf.cpu().numpy().view(np.uint16).astype(np.int32), ref)
        assert (fmin, fmax) == (rmin, rmax)


def test_slice_to_tif(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.normal(0, 1, (128, 128)).astype(np.float32)
    u16, lo, hi = output.output_stage(torch.from_numpy(img).cuda(), 64.5, 2)
    torch.cuda.synchronize()
    a = u16.cpu().numpy().view(np.uint16)
    desc = output.image_description(900, 0.5, 1.0, lo, 64.5, 0)
    output.write_tif16(tmp_path / "slice.tif", a.astype(np.int32), desc)
    back, d, _ = output.read_tif16(tmp_path / "slice.tif")
    assert np.array_equal(back, a) and d == desc
    assert a.max() <= 65534 and a[0, 0] == 0
