"""Pins the CPU oracle (oracle/oracle.c) -- the checker of every GPU parity test.

The reference ships no tests, golden vectors or fixtures for this path (its
only data files are angle logs, see tests/test_golden.py), so each oracle
function is pinned here by mathematics the reference source fixes:
closed forms derived from cfft.cpp, brute-force DFT / convolution on tiny
sizes, hand-computed integer cases, invariants, and an analytic phantom whose
reconstruction must return its attenuation values.  Each test says which
plausible mistake it would catch.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from recview_amd import phantom

RNG = np.random.default_rng(20241202)


# ----------------------------------------------------------------- CFft --

@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 512])
def test_fft1rev_closed_form(n):
    """FFT1Rev = sum_s f[s] exp(-2 pi i k (s - C)/n), C = n/2 - 1 (cfft.cpp:193-277).
    Catches a wrong stage order, twiddle sign, or the reversed-input / reversed-
    output convention of the nmin < 0 branch."""
    x = (RNG.standard_normal(n) + 1j * RNG.standard_normal(n)).astype(np.complex64)
    C = n / 2 - 1
    s = np.arange(n)
    ref = (x[None, :].astype(np.complex128) * np.exp(-2j * np.pi * s[:, None] * (s[None, :] - C) / n)).sum(1)
    got, _ = O.fft1rev(x)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max() * math.log2(n)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 512])
def test_fft1_closed_form(n):
    """FFT1 = (1/n) sum_k g[k] exp(+2 pi i k (q - C)/n) (cfft.cpp:279-369, 412-429)."""
    x = (RNG.standard_normal(n) + 1j * RNG.standard_normal(n)).astype(np.complex64)
    C = n / 2 - 1
    s = np.arange(n)
    ref = (x[None, :].astype(np.complex128) * np.exp(2j * np.pi * s[None, :] * (s[:, None] - C) / n)).sum(1) / n
    got, _ = O.fft1(x)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max() * math.log2(n)


def test_fft_roundtrip_and_twiddles():
    """FFT1(FFT1Rev(x)) = x exactly up to rounding; twiddle 0 is exactly (1, -0)."""
    n = 256
    x = (RNG.standard_normal(n) + 1j * RNG.standard_normal(n)).astype(np.complex64)
    F, _ = O.fft1rev(x)
    y, _ = O.fft1(F.astype(np.complex64))
    assert np.abs(y - x).max() < 1e-5
    wr, wi = O.cfft_twiddles(8)
    assert wr[0] == 1.0 and wi[0] == 0.0
    # w[k] = exp(-2 pi i k / n): quarter turn is -i
    assert abs(wr[64]) < 1e-6 and abs(wi[64] + 1.0) < 1e-6


# -------------------------------------------------------------- filters --

def test_set_filter_shapes():
    """SetFilter (gazoDoc.cpp:3747-3812): ramp k/n below n*cutoff and 0 above;
    Hann = ramp*(1+cos(pi k/ndimh))/2; Hamming = ramp*(0.54+0.46cos);
    Parzen branches meet at ndimh/2; Butterworth is not truncated."""
    n = 1024
    k = np.arange(n)
    ramp = O.set_filter(n, 2, 0.5)
    nh = 512
    assert np.array_equal(ramp[nh:], np.zeros(n - nh, np.float32))
    np.testing.assert_allclose(ramp[:nh], k[:nh] / n, rtol=1e-6)
    hann = O.set_filter(n, 0, 0.5)
    np.testing.assert_allclose(hann[:nh], k[:nh] / n * (1 + np.cos(np.pi * k[:nh] / nh)) / 2, rtol=2e-6, atol=1e-9)
    ham = O.set_filter(n, 1, 0.5)
    np.testing.assert_allclose(ham[:nh], k[:nh] / n * (0.54 + 0.46 * np.cos(np.pi * k[:nh] / nh)), rtol=2e-6)
    par = O.set_filter(n, 3, 0.5)
    f = 0.5
    assert abs((1 - 6 * f * f * (1 - f)) - 2 * (1 - f) ** 3) < 1e-12   # continuity of the two branches
    np.testing.assert_allclose(par[nh // 2], (nh // 2) / n * 0.25, rtol=1e-5)
    but = O.set_filter(n, 4, 0.25, 6.0)
    np.testing.assert_allclose(but, k / n / np.sqrt(1 + (k / 256.0) ** 6), rtol=5e-6, atol=1e-9)
    assert but[-1] > 0                                       # not zeroed above the cutoff
    ones = O.set_filter(n, 99)
    assert np.all(ones == 1)
    cut = O.set_filter(n, 2, 0.3)
    assert cut[int(n * 0.3) - 1] > 0 and cut[int(n * 0.3)] == 0


# ------------------------------------------------------------- geometry --

def test_geometry_special_cases():
    """cudaReconstHost.cpp:635-658: n = 2^(floor(log2 W)+1+zoom) -- 4096 for W = 2048
    (the comment at :642), less than 2x padding for non powers of two."""
    g = O.geometry(2048, 2, 1024.5)
    assert (g.ndim, g.N, g.z, g.h, g.ihoffset) == (4096, 2048, 1, 1024, 2048 - 1 - 1024)
    g = O.geometry(1000, 2, 500.0)
    assert g.ndim == 1024 and g.N == 1000
    g = O.geometry(1024, 3, 512.25)           # x2 zoom
    assert g.z == 2 and g.N == 2048 and g.ndim == 4096
    # fcenter = (h + frac*z)*16 for zoom, (h + frac)*16 without (cudaReconst.cu:220-221)
    assert g.fcenter16 == np.float32((1024 + 0.25 * 2) * 16)
    g = O.geometry(1024, 2, 512.25)
    assert g.fcenter16 == np.float32((512 + 0.25) * 16) and g.ic == 512


# ------------------------------------------------------------- row prep --

def test_row_prep_integer_zoom_truncates_toward_zero():
    """(s[k]*(z-j))/z + (s[k+1]*j)/z in C integer arithmetic (cudaReconstHost.cpp:748-749):
    for s = [-7, 5], z = 4, j = 1: -21/4 -> -5, 5/4 -> 1, sum -4 (floor would give -5)."""
    W = 2
    g = O.geometry(W, 4, 0.5)             # zoom x4: z = 4, ic = 0
    assert g.z == 4
    s = np.array([-7, 5], np.int16)
    re = O.row_prep(s, g, 0)
    idx0 = (0 - g.ic) * g.z + g.ndim // 2 - 1
    # j = 1: -21/4 + 5/4 = -5 + 1; j = 2: -14/4 + 10/4 = -3 + 2; j = 3: -7/4 + 15/4 = -1 + 3
    assert re[idx0] == -7 and re[idx0 + 1] == -4 and re[idx0 + 2] == -1 and re[idx0 + 3] == 2
    assert re[idx0 + 4] == 5
    # replicate padding on both sides
    assert np.all(re[:idx0] == -7) and np.all(re[idx0 + 5:] == 5)


def test_row_prep_placement_and_cudafft():
    """Sample k lands at (k - ic)*z + n/2 - 1 (host) or + n/2 (CUDAFFT, zero padded)."""
    W = 10
    s = np.arange(1, W + 1, dtype=np.int16) * 10
    g = O.geometry(W, 2, 4.3)
    re = O.row_prep(s, g, 0)
    for k in range(W):
        assert re[(k - g.ic) + g.ndim // 2 - 1] == s[k]
    assert re[0] == s[0] and re[-1] == s[-1]
    rz = O.row_prep(s, g, 2)
    for k in range(W):
        assert rz[(k - g.ic) + g.ndim // 2] == s[k]
    assert rz[0] == 0 and rz[-1] == 0


# --------------------------------------------------- filter = convolution --

def _h_re(filt):
    n = len(filt)
    d = np.arange(n)
    k = np.arange(n)
    return (filt[None, :].astype(np.float64) * np.cos(2 * np.pi * np.outer(d, k) / n)).sum(1) / n


def test_filter_rows_is_circular_convolution():
    """Closed form of FFT1Rev / x H / FFT1 / .re: y = f (*) h_re with
    h_re[d] = (1/n) sum_k H[k] cos(2 pi k d / n), no index shift.  A dropped
    1/n, a shift of the crop, or a conjugated twiddle fails this."""
    W, P = 24, 3
    sino = RNG.standard_normal((P, W)).astype(np.float32)
    g = O.geometry(W, 2, 11.6)
    for fid in (0, 2, 4):
        filt = O.set_filter(g.ndim, fid, 0.5, 4.0)
        px = O.filter_rows(sino, g, filt, kind=1)
        h = _h_re(filt)
        n = g.ndim
        for r in range(P):
            f = O.row_prep(sino[r], g, 1).astype(np.float64)
            y = np.array([sum(f[s] * h[(i - s) % n] for s in range(n)) for i in range(n)])
            ref = y[g.ihoffset:g.ihoffset + g.N]
            assert np.abs(px[r] - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
        pd = O.filter_rows_f64(sino, g, filt)          # O-f64 filter: same closed form
        h = _h_re(filt)
        for r in range(P):
            f = O.row_prep(sino[r], g, 1).astype(np.float64)
            y = np.array([sum(f[s] * h[(i - s) % n] for s in range(n)) for i in range(n)])
            np.testing.assert_allclose(pd[r], y[g.ihoffset:g.ihoffset + g.N], atol=1e-11)


def test_ramp_kills_constant_rows_and_is_linear():
    """A constant row with replicate padding is constant on the whole circle, so
    a filter with H[0] = 0 returns zeros; filtering is linear."""
    W = 64
    g = O.geometry(W, 2, 32.0)
    filt = O.set_filter(g.ndim, 0)
    const = np.full((1, W), 1234, np.int16)
    assert np.abs(O.filter_rows(const, g, filt, kind=0)).max() < 1e-3
    a = RNG.standard_normal((2, W)).astype(np.float32)
    fa = O.filter_rows(a, g, filt, kind=1)
    fs = O.filter_rows((2 * a[0] - 3 * a[1])[None], g, filt, kind=1)
    np.testing.assert_allclose(fs[0], 2 * fa[0] - 3 * fa[1], atol=1e-5)


# ------------------------------------------------------------------ LUT --

def test_lut16_formulas_by_hand():
    """general.cpp:2131-2136 vs cudaReconst.cu:237-239 on px = [1.005, 2.0]:
    entry 16j + s = (int)(p0 + slope*s); the last sample has slope 0."""
    px = np.array([1.005, 2.0], np.float32)
    cpu = O.lut16_i32(px, 0)
    cuda = O.lut16_i32(px, 1)
    p0 = np.float32(px[0]) * np.float32(100)
    sl_cpu = (px[1] - px[0]) / np.float32(16) * np.float32(100)
    sl_cuda = (px[1] * np.float32(100) - p0) / np.float32(16)
    for s in range(16):
        assert cpu[s] == int(np.float32(p0 + np.float32(sl_cpu * np.float32(s))))
        assert cuda[s] == int(np.float32(p0 + np.float32(sl_cuda * np.float32(s))))
    assert np.all(cpu[16:] == 200) and np.all(cuda[16:] == 200)
    assert cpu[0] == 100 and cpu[15] == int(100.5 + 99.5 * 15 / 16)
    lf = O.lut16_f32(px)
    np.testing.assert_allclose(lf[:16], px[0] + (px[1] - px[0]) * np.arange(16) / 16, rtol=1e-6)


# ------------------------------------------------------------------ BP --

def test_bp_rules_and_point_geometry():
    """One projection of a row that is 1 at a single sample back-projects onto
    the line u = (ix-h)cos - (iy-h)sin + fcenter/16 = k (cudaReconst.cu:26-28).
    Rule 0 masks t < 0 as a float (projx64.asm:261-266); rule 1 truncates first,
    so t in (-1, 0) reads entry 0 (general.cpp:2152-2156)."""
    N = 16
    g = O.geometry(N, 2, 8.0)
    lut = np.zeros(16 * N, np.int32)
    lut[16 * 5:16 * 6] = 7                      # sample 5
    for deg in (0.0, 30.0, 90.0, 135.0):
        c, s, o = O.bp_params(deg, 0.0, g)
        img = np.zeros((N, N), np.int32)
        O.bp_lut_i32(img, lut, N, c, s, o, 0)
        th = np.float64(np.float32(deg)) * 0.0174532925199
        iy, ix = np.nonzero(img)
        u = (ix - 8) * math.cos(th) - (iy - 8) * math.sin(th) + g.fcenter16 / 16.0
        assert len(ix) > 0 and np.all((u > 5 - 1e-4) & (u < 6 + 1e-4))
    # t in (-1, 0): rule 1 reads lut[0], rule 0 does not
    lut = np.arange(16 * N, dtype=np.int32) + 1
    img0 = np.zeros((N, N), np.int32)
    img1 = np.zeros((N, N), np.int32)
    O.bp_lut_i32(img0, lut, N, 1.0, 0.0, -0.5, 0)     # t = ix - 0.5
    O.bp_lut_i32(img1, lut, N, 1.0, 0.0, -0.5, 1)
    assert img0[0, 0] == 0 and img1[0, 0] == 1
    assert np.array_equal(img0[:, 1:], img1[:, 1:])


def test_bp_brute_force_tiny():
    """orc_bp_lut_i32 against an independent per-pixel loop (N = 12, random LUT)."""
    N = 12
    g = O.geometry(N, 2, 6.3)
    lut = RNG.integers(-1000, 1000, 16 * N).astype(np.int32)
    for deg, rule in ((17.0, 0), (101.0, 1), (163.0, 0)):
        c, s, o = O.bp_params(deg, 0.5, g)
        img = np.zeros((N, N), np.int32)
        O.bp_lut_i32(img, lut, N, c, s, o, rule)
        ref = np.zeros((N, N), np.int64)
        for iy in range(N):
            fy = np.float32(np.float32(iy) * np.float32(s)) + np.float32(o)
            for ix in range(N):
                t = np.float32(np.float32(ix) * np.float32(c)) + np.float32(fy)
                if rule == 0:
                    if 0 <= t < 16 * N:
                        ref[iy, ix] += lut[int(t)]
                else:
                    i0 = int(t)
                    if 0 <= i0 < 16 * N:
                        ref[iy, ix] += lut[i0]
        assert np.array_equal(img, ref)


def _phantom(W, P, seed, c_off=0.37):
    c_det = W / 2 + c_off
    fdeg = phantom.angles_deg(P)
    ell = phantom.random_ellipses(seed, count=3, axis_range=(0.15, 0.3), mu_range=(0.005, 0.02))
    sino = phantom.sinogram(ell, W, fdeg, c_det, scale=W).numpy().astype(np.float32)
    return ell, fdeg, c_det, sino


def test_reconstruction_returns_attenuation():
    """FBP of an analytic phantom returns mu per sample at pixel (ix, iy) =
    (X + h, h - Y): pins orientation, the half-Ram-Lak normalisation (one-sided
    filter + .re, restored by 2 pi / P, gazoDoc.cpp:4000-4001) and the centre."""
    W, P = 96, 180
    ell, fdeg, c_det, sino = _phantom(W, P, 77)
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 2, 0.5)           # ramp
    truth = phantom.image(ell, W, W).numpy()
    rec = O.recon_of64(sino, fdeg, g, filt) * 2 * math.pi / P
    rf = O.recon_rf32(sino, fdeg, g, filt) * 2 * math.pi / P
    yy, xx = np.mgrid[0:W, 0:W]
    inner = ((xx - W / 2) ** 2 + (yy - W / 2) ** 2) < (0.4 * W) ** 2
    # away from edges the reconstruction matches the attenuation map
    from scipy.ndimage import binary_erosion
    core = binary_erosion(truth > 0, iterations=3) & inner
    bg = binary_erosion(truth == 0, iterations=3) & inner
    assert core.sum() > 50
    # The frequency-sampled ramp (H[0] = 0) leaves a small uniform offset; after
    # removing it, inside and outside the objects agree with the truth.
    offset = np.median(rec[bg])
    assert abs(offset) < 0.15 * truth[core].mean()
    assert abs(np.median(rec[core] - truth[core]) - offset) < 0.05 * truth[core].mean()
    assert np.percentile(np.abs(rec[core] - truth[core] - offset), 90) < 0.25 * truth[core].mean()
    assert np.abs(rec - rf).max() < 0.02 * truth.max()   # R-f32 (LUT16) close to O-f64


def test_rint_matches_rf32_and_mirror_invariance():
    """R-int = (x3000 input, x100 LUT, truncations) tracks R-f32 * 3e5; a scan
    with angles + 180 deg and mirrored rows reconstructs the same image."""
    W, P = 64, 90
    ell, fdeg, c_det, sino = _phantom(W, P, 5, c_off=0.0)
    g = O.geometry(W, 2, c_det)
    filt = O.set_filter(g.ndim, 0)
    rows = phantom.to_int16(__import__("torch").from_numpy(sino)).numpy()
    ri = O.recon_rint(rows, fdeg, g, filt, formula=0, rule=0).astype(np.float64)
    rf = O.recon_rf32(rows.astype(np.float32), fdeg, g, filt)
    rng_ = rf.max() - rf.min()
    assert np.abs(ri - 100 * rf).max() < 1e-2 * 100 * rng_
    # mirror: row reversed about the detector centre c_det = W/2 (samples k <-> W-k)
    m = np.zeros_like(sino)
    m[:, 1:] = sino[:, ::-1][:, :-1]
    a = O.recon_of64(sino, fdeg, g, filt)
    b = O.recon_of64(m, (fdeg + 180.0).astype(np.float32), g, filt)
    inner = (np.add.outer((np.arange(W) - W / 2) ** 2, (np.arange(W) - W / 2) ** 2) < (0.4 * W) ** 2)
    assert np.abs(a - b)[inner].max() < 1e-3 * np.abs(a).max()


def test_output_stage_by_hand():
    """gazoDoc.cpp:3993-4021: imax starts at 0, values scaled to 0..65534, zero
    outside the circle of radius min(z c, N - z c)."""
    N = 8
    img = np.arange(N * N, dtype=np.int32).reshape(N, N) - 10
    out, imin, imax = O.output_stage(img, 1, 3.0)
    assert imin == -10 and imax == 53
    assert out[0, 0] == 0                      # outside radius 3
    v = int(np.float32(np.float32(img[4, 4] + 10) * np.float32(65534)) / np.float32(63))
    assert out[4, 4] == v
    out2, imin2, imax2 = O.output_stage(-np.ones((N, N), np.int32), 1, 4.0)
    assert imax2 == 0 and imin2 == -1          # imax never starts below 0


def test_bp_points_equals_grid_and_reproduces_linear_rows():
    """The point-wise FP64 back-projection equals the pinned grid version on
    integer points, and at fractional (zoomed) points reproduces the exact
    back-projection of rows that are linear in the detector coordinate."""
    W, P = 64, 37
    g = O.geometry(W, 2, W / 2 + 0.37)
    rng = np.random.default_rng(9)
    fdeg = np.sort(rng.uniform(0, 180, P)).astype(np.float32)
    px = rng.standard_normal((P, g.N)).astype(np.float32)
    grid = O.bp_rows_linear_f64(px, fdeg, g)
    ys, xs = np.mgrid[0:g.N, 0:g.N]
    pts = O.bp_points_linear_f64(px.astype(np.float64), fdeg, g, xs, ys)
    np.testing.assert_allclose(pts.reshape(g.N, g.N), grid, rtol=1e-12, atol=1e-12)
    # rows p_i[k] = a_i + b_i k: lerp is exact, so the sum is a closed form
    a, b = rng.standard_normal(P), rng.standard_normal(P) * 0.01
    lin = a[:, None] + b[:, None] * np.arange(g.N)[None, :]
    x = 20.0 + rng.uniform(0, 20, 50)
    y = 20.0 + rng.uniform(0, 20, 50)
    got = O.bp_points_linear_f64(lin, fdeg, g, x, y)
    th = fdeg.astype(np.float64) * 0.0174532925199
    h = g.N / 2
    u = (x[None, :] - h) * np.cos(th)[:, None] - (y[None, :] - h) * np.sin(th)[:, None] + g.fcenter16 / 16.0
    ok = (u >= 0) & (u < g.N - 1)
    ref = ((a[:, None] + b[:, None] * u) * ok).sum(axis=0)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10)


# ---------------------------------------------------------------- sinogram generation
def _frames(W, values, dark=100):
    """uint16 frames [n][1][W]: given rows plus a dark frame of constant `dark`."""
    f = [np.asarray(v, np.float64) for v in values] + [np.full(W, dark)]
    return np.round(np.stack(f)[:, None, :]).astype(np.uint16)


def test_sinogram_recovers_line_integrals():
    """sample = dark + I0 exp(-p) gives round(3000 p) (count rounding <= 1 unit)."""
    W, dark, I0 = 64, 100, 60000.0
    p = np.linspace(0.0, 2.9, W)
    fr = _frames(W, [np.full(W, I0 + dark), I0 * np.exp(-p) + dark, np.full(W, I0 + dark)], dark)
    out, st = O.sinogram(fr, [0, 1, 0], [0, 1, 2], [0, 0, 0])
    assert st == 0
    assert np.abs(out[1, 0] - 3000 * p).max() <= 1.0
    assert np.array_equal(out[0], fr[0].view(np.int16)) and np.array_equal(out[2], fr[2].view(np.int16))


def test_sinogram_incident_interpolated_in_exposure_time():
    """Incident at t=0 (1000 counts) and t=2 (3000): a sample at t=1 sees 2000."""
    W, dark = 8, 50
    fr = _frames(W, [np.full(W, 1000 + dark), np.full(W, 2000 * np.exp(-0.5) + dark), np.full(W, 3000 + dark)], dark)
    out, _ = O.sinogram(fr, [0, 1, 0], [0.0, 1.0, 2.0], [0, 0, 0])
    assert np.all(np.abs(out[1, 0] - 1500) <= 1)
    # equal exposure times: t0 = 0 -> the earlier incident frame alone
    out2, _ = O.sinogram(fr, [0, 1, 0], [5.0, 1.0, 5.0], [0, 0, 0])
    assert np.all(np.abs(out2[1, 0] - np.round(3000 * np.log(1000 / (2000 * np.exp(-0.5))))) <= 1)


def test_sinogram_floors_zero_cases_and_zernike():
    W, dark = 6, 100
    inc = np.array([5000, 5000, 5000, 40, 0, 5000]) + dark        # incident - dark = 0 at j=4
    smp = np.array([2000, dark + 19, dark + 20, 30, 3000, 1000]) + 0
    smp[0] += dark
    smp[5] += dark
    fr = _frames(W, [inc, smp, inc], dark)
    out, _ = O.sinogram(fr, [0, 1, 0], [0, 1, 2], [0, 0, 0])
    row = out[1, 0]
    assert row[1] == 0                      # sample - dark = 19 < 20
    assert row[2] == int(np.log(5000 / 20) * 3000 + 0.5)
    assert row[3] == 0                      # sample below dark
    assert row[4] == 0                      # incident - dark <= 0
    out8, _ = O.sinogram(fr, [0, 1, 0], [0, 1, 2], [0, 0, 0], is8bit=True)
    assert out8[1, 0, 1] == int(np.log(5000 / 19) * 3000 + 0.5)   # 8-bit floor is 3
    z, _ = O.sinogram(fr, [0, 1, 0], [0, 1, 2], [0, 0, 0], flags=2)
    assert z[1, 0, 0] == int(5000 * 1000 / 2000 + 0.5) and z[1, 0, 5] == int(5000 * 1000 / 1000 + 0.5)


def test_sinogram_frame_selection_skip_drift():
    W, dark = 16, 10
    base = np.full(W, 4000.0 + dark)
    smp = 4000 * np.exp(-np.linspace(0.1, 1.5, W)) + dark
    fr = _frames(W, [base, smp, smp, smp, smp, base], dark)
    binc = [0, 1, 1 | 4, 1, 1, 0]
    ref, _ = O.sinogram(fr, [0, 1, 1, 1, 1, 0], np.arange(6), np.zeros(6))
    out, _ = O.sinogram(fr, binc, np.arange(6), np.zeros(6))
    assert (out[2] == 0).all()                              # SKIP
    assert np.array_equal(out[1], ref[1])
    # start/step: only i % 2 == 1 converted, the others untouched
    sel, _ = O.sinogram(fr, [0, 1, 1, 1, 1, 0], np.arange(6), np.zeros(6), start=1, step=2)
    assert np.array_equal(sel[1], ref[1]) and np.array_equal(sel[3], ref[3])
    assert np.array_equal(sel[2], fr[2].view(np.int16)) and np.array_equal(sel[4], fr[4].view(np.int16))
    # drift list: shift by -(piDrift[i]) / binning, truncated toward zero, zero fill
    pd = np.array([0, 3, -3, 5, -5, 0])
    dr, _ = O.sinogram(fr, [0, 1, 1, 1, 1, 0], np.arange(6), np.zeros(6), flags=4, pidrift=pd, binning=2)
    for i, s in ((1, -1), (3, -2), (4, 2)):                 # -3/2 -> -1, -5/2 -> -2, 5/2 -> 2
        exp = np.zeros(W, np.int16)
        if s >= 0:
            exp[s:] = ref[i, 0, :W - s]
        else:
            exp[:W + s] = ref[i, 0, -s:]
        assert np.array_equal(dr[i, 0], exp), i
    # drift parameters: frames strictly inside (drStart, drEnd) are zeroed with drOmit
    om, _ = O.sinogram(fr, [0, 1, 1, 1, 1, 0], np.arange(6), np.zeros(6), flags=8, drift=(1, 4, 0.0, 0.0, 1))
    assert (om[2] == 0).all() and (om[3] == 0).all() and np.array_equal(om[1], ref[1])


def test_sinogram_incident_search_quirks():
    """i0 skips NOUSE incident frames; the forward search for i1 does not."""
    W, dark = 4, 0
    a, b, c = np.full(W, 1000.0), np.full(W, 4000.0), np.full(W, 9000.0)
    smp = np.full(W, 500.0)
    fr = _frames(W, [a, smp, b, c], dark)
    # frame 2 (b) is an incident frame marked NOUSE: i1 for frame 1 is still b
    out, _ = O.sinogram(fr, [0, 1, 2, 0], [0.0, 1.0, 2.0, 3.0], np.zeros(4))
    assert out[1, 0, 0] == int(np.log(2500 / 500) * 3000 + 0.5)
    # no incident frame at all -> error status
    _, st = O.sinogram(_frames(W, [smp, smp], dark), [1, 1], [0, 1], [0, 0])
    assert st == -1


@pytest.mark.parametrize("isa", [0, 1, 2])
def test_cpu_baseline_image_equals_reference_integer_path(isa):
    """The timed CPU baseline (projx64 restated with intrinsics, row bands, per-thread
    filtering, partial-image sum) computes the reference CPU path's image: equal bit
    for bit to R-int with the CPU LUT formula and the double-precision centre, the
    float-mask index rule for AVX2 / AVX-512 and truncation for the scalar loop.  An
    odd thread count leaves bands of unequal height; a width that is not a multiple
    of 16 exercises the masked tail of the vector loops."""
    if isa > O.cpu_isa():
        pytest.skip("host lacks the instruction set")
    W, P = 120, 60
    ell, fdeg, c_det, sino = _phantom(W, P, 91, c_off=0.37)
    rows = phantom.to_int16(__import__("torch").from_numpy(sino)).numpy()
    g = O.geometry(W, 2, c_det, float_center=False)
    filt = O.set_filter(g.ndim, 0, 0.5, 6.0)
    secs, img = O.cpu_baseline(rows, fdeg, g, filt, 3, isa=isa)
    ref = O.recon_rint(rows, fdeg, g, filt, formula=0, rule=0 if isa > 0 else 1)
    assert secs > 0
    assert np.array_equal(img, ref)
    assert np.abs(ref).max() > 1000
