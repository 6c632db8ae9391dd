"""The C++ host harness (recview_amd/host/recview_harness.cpp): RecView's
calling sequence through the C ABI only -- no Python, no HIP header -- gives the
oracle's R-int image bit for bit and a 16-bit TIFF equal to the oracle's output
stage."""
import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import pyoracle as O
from recview_amd import build, output, phantom

pytestmark = pytest.mark.gpu


def test_cpp_host_harness(tmp_path):
    if not os.path.exists(build.HARNESS):
        pytest.skip("harness not built")
    W, P = 256, 180
    c_det = W / 2 + 0.37
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(77, count=8)
    sino = phantom.to_int16(phantom.sinogram(ell, W, fdeg, c_det, scale=W)).numpy()
    sino.tofile(tmp_path / "s.i16")
    fdeg.tofile(tmp_path / "a.f32")
    r = subprocess.run([build.HARNESS, str(tmp_path / "s.i16"), str(P), str(W), str(c_det - 0.5),
                        str(tmp_path / "a.f32"), str(tmp_path / "o.i32"), str(tmp_path / "o.tif")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "o.i32", np.int32).reshape(W, W)
    g = O.geometry(W, 2, c_det)
    ref = O.recon_rint(sino, fdeg, g, O.set_filter(g.ndim, 0), formula=1, rule=1)
    assert np.array_equal(img, ref)
    ref16, lo, hi = O.output_stage(ref, 1, c_det - 0.5, 1)
    back, desc, _ = output.read_tif16(tmp_path / "o.tif")
    assert np.array_equal(back.astype(np.int32), ref16)
    assert desc.startswith(f"{P}\t0.500000\t")
