"""Offset-CT stitch (gazoDoc.cpp:3536-3665): oracle pinned against the analytic
180-degree sinogram of a double-width detector; GPU kernel bit-exact with the
oracle (int16 rows) over both axis halves, both dCent branches, binning and trim.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from recview_amd import phantom


def _scan360(W, P360, c_det, seed, fov_scale):
    """P360 + 1 frames over 0..360 deg inclusive (as the reference's offset-CT logs,
    docs/output1800offset.log, end at or beyond 360 deg)."""
    fdeg = (np.arange(P360 + 1) * (360.0 / P360)).astype(np.float32)
    ell = phantom.random_ellipses(seed, count=6, axis_range=(0.1, 0.3), mu_range=(0.002, 0.01))
    sino = phantom.sinogram(ell, W, fdeg, c_det, scale=fov_scale)
    return ell, fdeg, sino


@pytest.mark.parametrize("c_frac,side", [(0.3, "right"), (0.7, "left"), (0.2, "left")])
def test_oracle_stitch_matches_wide_detector(c_frac, side):
    """The stitched row k must measure the line integral at t = k - c_wide with
    c_wide = c_det (axis right of centre) or W + c_det (left): a wrong mirror
    index, interpolation weight or placement fails this."""
    W, P360 = 128, 240
    c_det = (int(0.8 * W) if side == "right" else int(0.2 * W)) + c_frac
    ell, fdeg, sino = _scan360(W, P360, c_det, 11, fov_scale=1.2 * W)
    rows = phantom.to_int16(sino).numpy()
    st = O.offset_stitch(rows, fdeg, np.ones(P360 + 1, np.int8), c_det)
    assert st.shape == (P360 // 2 + 1, 2 * W)
    c_wide = c_det if side == "right" else W + c_det
    ref = phantom.sinogram(ell, 2 * W, fdeg[:st.shape[0]], c_wide, scale=1.2 * W).numpy() * 3000
    d = np.abs(st.astype(np.float64) - ref)
    covered = ref > 0
    assert np.percentile(d[covered], 95) < 0.02 * ref.max()
    assert np.median(d[covered]) < 0.005 * ref.max()


def test_oracle_stitch_right_half_upper_fraction_quirk():
    """Axis right of centre with frac(c) >= 0.5 (gazoDoc.cpp:3600-3606): the
    reference interpolates between mirror samples idx-1 and idx at weight
    2 - 2 dCent, i.e. reads the partner view at q = 2 iCent + 1 - 2 dCent - k, so
    the mirrored part measures t = k - (iCent + 1 - 3 dCent) instead of
    k - (iCent + dCent).  Kept for parity; pinned here by that closed form."""
    W, P360 = 128, 240
    iCent, dCent = 102, 0.7
    c_det = iCent + dCent
    ell, fdeg, sino = _scan360(W, P360, c_det, 11, fov_scale=1.2 * W)
    rows = phantom.to_int16(sino).numpy()
    st = O.offset_stitch(rows, fdeg, np.ones(P360 + 1, np.int8), c_det)
    g = W - 1 - iCent
    mirror = slice(iCent + g + 1, 2 * W)          # beyond the cross-fade band
    ang = fdeg[:st.shape[0]]
    quirk = phantom.sinogram(ell, 2 * W, ang, iCent + 1 - 3 * dCent, scale=1.2 * W).numpy() * 3000
    true = phantom.sinogram(ell, 2 * W, ang, c_det, scale=1.2 * W).numpy() * 3000
    dq = np.abs(st[:, mirror] - quirk[:, mirror])
    dt = np.abs(st[:, mirror] - true[:, mirror])
    cov = quirk[:, mirror] > 0
    assert np.median(dq[cov]) < 0.005 * quirk.max()
    assert np.median(dt[cov]) > 2 * np.median(dq[cov])


def test_oracle_stitch_direct_part_exact():
    """Samples on the axis side of the cross-fade are copied unchanged (:3580-3582)."""
    W, P360 = 64, 40
    c_det = 50.25
    _, fdeg, sino = _scan360(W, P360, c_det, 3, fov_scale=1.2 * W)
    rows = phantom.to_int16(sino).numpy()
    st = O.offset_stitch(rows, fdeg, np.ones(P360 + 1, np.int8), c_det)
    iCent = int(c_det)
    g = W - 1 - iCent
    assert np.array_equal(st[:, :iCent - g + 1], rows[:st.shape[0], :iCent - g + 1])


@pytest.mark.gpu
@pytest.mark.parametrize("W,P360,c_det,bin_,trim", [
    (128, 120, 0.8 * 128 + 0.3, 1, 0),      # right half, dCent < 0.5
    (128, 120, 0.8 * 128 + 0.8, 1, 0),      # right half, dCent >= 0.5
    (100, 90, 0.25 * 100 + 0.2, 1, 0),      # left half, dCent < 0.5
    (100, 90, 0.25 * 100 + 0.6, 1, 0),      # left half, dCent >= 0.5
    (256, 60, 200.4, 2, 0),                 # binning 2
    (256, 60, 190.7, 2, 4),                 # binning 2 with trim
    (96, 72, 40.1, 1, 0),                   # left half near the centre
])
def test_gpu_stitch_bit_exact(W, P360, c_det, bin_, trim):
    import torch
    from recview_amd import offset
    print(f"case W={W} P={P360} c={c_det} bin={bin_} trim={trim}", flush=True)
    # rows of the binned detector (width W / bin), axis at c_det in raw pixels
    _, fdeg, sino = _scan360(W // bin_, P360, (c_det - 0.5) / bin_ + 0.5, 21 + W, fov_scale=1.2 * W / bin_)
    rows = phantom.to_int16(sino).numpy()
    binc = np.ones(P360 + 1, np.int8)
    binc[[5, 17]] = 0                        # skipped frames leave zero rows / shift partners
    ref = O.offset_stitch(rows, fdeg, binc, c_det, bin_=bin_, trim=trim)
    got, ang = offset.stitch(torch.from_numpy(rows).cuda(), fdeg, c_det - 0.5, binc=binc,
                             binning=bin_, trim=trim)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(ang, fdeg[:ref.shape[0]])
    # float rows: same arithmetic without the (short) truncations
    gotf, _ = offset.stitch(torch.from_numpy(rows.astype(np.float32)).cuda(), fdeg, c_det - 0.5,
                            binc=binc, binning=bin_, trim=trim)
    assert np.abs(gotf.cpu().numpy() - ref).max() <= 2.0
