"""The synthetic-input generators (CPU): the ray-driven projector used for the
C2 texture follows the analytic sinogram's geometry."""
import numpy as np
import torch

from recview_amd import phantom


def test_ray_projector_matches_analytic_sinogram():
    """Projecting the point-sampled image of ellipses reproduces their analytic
    line integrals up to the pixel discretisation; a mirrored axis, a swapped
    sine or a centre off by a sample would differ by the order of the peak."""
    W = 128
    fdeg = phantom.angles_deg(30)
    c = W / 2 + 0.37
    ell = phantom.random_ellipses(5, count=4, axis_range=(0.15, 0.3), mu_range=(0.01, 0.02))
    ana = phantom.sinogram(ell, W, fdeg, c, scale=W).float()
    ray = phantom.texture_sinogram(phantom.image(ell, W, W), W, fdeg, c)
    peak = float(ana.max())
    assert float((ray - ana).abs().mean()) < 0.005 * peak
    assert float((ray - ana).abs().max()) < 0.05 * peak
    # a point at pixel (ix, iy) projects to t = (ix - h) cos + (h - iy) sin
    img = torch.zeros(W, W, dtype=torch.float64)
    img[40, 90] = 1.0
    row = phantom.texture_sinogram(img, W, np.array([30.0], np.float32), c)[0]
    th = np.deg2rad(30.0)
    t = (90 - W / 2) * np.cos(th) + (W / 2 - 40) * np.sin(th)
    assert abs(float((row * (torch.arange(W) - c)).sum() / row.sum()) - t) < 0.05
