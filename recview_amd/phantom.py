"""Seeded synthetic phantoms and their parallel-beam sinograms.

Inputs for tests and benchmarks (there are no real datasets offline).  All
random parameters come from a SplitMix64 stream, so a seed gives the same
phantom on any machine; line integrals are analytic (ellipses / ellipsoid
slices), evaluated in float64 with torch on whatever device is given.

Geometry convention (shared with the reconstruction): detector sample k of a
row measures the line integral at t = k - c_det, where c_det is the detector
position of the rotation axis (RECONST_INFO.center = c_det, i.e. the user
centre c_det - 0.5, gazoDoc.cpp:3935).  Image pixel (ix, iy) of an N x N slice
sits at X = ix - N/2, Y = N/2 - iy, and projects to t = X cos th + Y sin th.
"""
import math

import numpy as np
import torch

LOG_SCALE = 3000          # constants.h:39


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014): deterministic across platforms."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def random_ellipses(seed, count=10, fov=1.0, axis_range=(0.05, 0.4), mu_range=(0.0, 0.02),
                    margin=0.95):
    """Ellipses (cx, cy, a, b, phi, mu) inside a disc of radius fov/2.

    Semi-axes are fractions of the FOV diameter; the centre is drawn so that the
    whole ellipse stays within margin * radius.  mu is per detector sample.
    """
    rng = SplitMix64(seed)
    R = fov / 2.0
    out = []
    for _ in range(count):
        a = rng.uniform(*axis_range) * fov
        b = rng.uniform(*axis_range) * fov
        amax = max(a, b)
        rmax = max(margin * R - amax, 0.0)
        r = rmax * math.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2 * math.pi)
        phi = rng.uniform(0.0, math.pi)
        mu = rng.uniform(*mu_range)
        out.append((r * math.cos(ang), r * math.sin(ang), a, b, phi, mu))
    return np.array(out, dtype=np.float64)


def random_ellipsoids(seed, count=40, fov=1.0, axis_range=(0.05, 0.4), mu_range=(0.0, 0.02)):
    """Ellipsoids (cx, cy, cz, a, b, c, phi, mu) rotated about the z axis, z in [-1/2, 1/2]."""
    rng = SplitMix64(seed)
    R = fov / 2.0
    out = []
    for _ in range(count):
        a = rng.uniform(*axis_range) * R
        b = rng.uniform(*axis_range) * R
        c = rng.uniform(*axis_range) * R
        rmax = max(0.95 * R - max(a, b), 0.0)
        r = rmax * math.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2 * math.pi)
        cz = rng.uniform(-0.5 + c, 0.5 - c) * fov if c < 0.5 else 0.0
        phi = rng.uniform(0.0, math.pi)
        mu = rng.uniform(*mu_range)
        out.append((r * math.cos(ang), r * math.sin(ang), cz, a, b, c, phi, mu))
    return np.array(out, dtype=np.float64)


def ellipsoid_slice(ellipsoids, z):
    """Ellipses cut from the ellipsoids by the plane at height z (FOV units)."""
    out = []
    for cx, cy, cz, a, b, c, phi, mu in ellipsoids:
        q = 1.0 - ((z - cz) / c) ** 2
        if q <= 0:
            continue
        s = math.sqrt(q)
        out.append((cx, cy, a * s, b * s, phi, mu))
    return np.array(out, dtype=np.float64).reshape(-1, 6)


def angles_deg(P, span=180.0):
    """Evenly spaced projection angles (degrees, float32 like RECONST_INFO.fdeg)."""
    return (np.arange(P, dtype=np.float64) * (span / P)).astype(np.float32)


def sinogram(ellipses, W, fdeg, c_det, scale, device="cpu", dtype=torch.float64):
    """Analytic parallel-beam sinogram [P, W] of ellipses given in FOV units.

    scale converts FOV units to detector samples (e.g. W for a FOV spanning the
    detector); mu is per sample.  Row i, sample k: integral at t = k - c_det
    along direction fdeg[i].
    """
    dev = torch.device(device)
    th = torch.as_tensor(np.asarray(fdeg, np.float64) * (math.pi / 180.0), device=dev, dtype=dtype)
    t = torch.arange(W, device=dev, dtype=dtype) - c_det
    cth, sth = torch.cos(th)[:, None], torch.sin(th)[:, None]
    out = torch.zeros(len(fdeg), W, device=dev, dtype=dtype)
    for cx, cy, a, b, phi, mu in np.asarray(ellipses, np.float64).reshape(-1, 6):
        cx, cy, a, b = cx * scale, cy * scale, a * scale, b * scale
        t0 = cx * cth + cy * sth
        ca = torch.cos(th - phi)[:, None]
        sa = torch.sin(th - phi)[:, None]
        s2 = (a * ca) ** 2 + (b * sa) ** 2
        d2 = s2 - (t[None, :] - t0) ** 2
        out += torch.where(d2 > 0, (2.0 * mu * a * b) * torch.sqrt(d2.clamp_min(0)) / s2,
                           torch.zeros((), device=dev, dtype=dtype))
    return out


def image(ellipses, N, scale, device="cpu", dtype=torch.float64):
    """Point-sampled attenuation image of ellipses on the N x N reconstruction grid."""
    return image_roi(ellipses, N, scale, N / 2.0, 0.0, 0.0, 1.0, device=device, dtype=dtype)


def image_roi(ellipses, n_out, scale, h_ref, x0, y0, zoom, device="cpu", dtype=torch.float64):
    """Attenuation image on an off-centre zoomed grid: pixel (ix, iy) centre at
    reference-grid coordinates (x0 + ix/zoom, y0 + iy/zoom) of a grid of half-size h_ref."""
    dev = torch.device(device)
    xs = x0 + torch.arange(n_out, device=dev, dtype=dtype) / zoom - h_ref
    ys = h_ref - (y0 + torch.arange(n_out, device=dev, dtype=dtype) / zoom)
    X = xs[None, :].expand(n_out, n_out)
    Y = ys[:, None].expand(n_out, n_out)
    out = torch.zeros(n_out, n_out, device=dev, dtype=dtype)
    for cx, cy, a, b, phi, mu in np.asarray(ellipses, np.float64).reshape(-1, 6):
        cx, cy, a, b = cx * scale, cy * scale, a * scale, b * scale
        dx, dy = X - cx, Y - cy
        u = dx * math.cos(phi) + dy * math.sin(phi)
        v = -dx * math.sin(phi) + dy * math.cos(phi)
        out += torch.where((u / a) ** 2 + (v / b) ** 2 <= 1.0, torch.full((), mu, device=dev, dtype=dtype),
                           torch.zeros((), device=dev, dtype=dtype))
    return out


def texture_sinogram(img, W, fdeg, c_det, batch=16):
    """Ray-driven parallel-beam projection of a pixel image on the N x N
    reconstruction grid (pixel (ix, iy) at X = ix - N/2, Y = N/2 - iy): for row i and
    sample k the line X cos + Y sin = k - c_det, sampled every pixel along the ray
    with bilinear interpolation and summed (a discrete forward projector for
    structure with no analytic line integral, such as a noise texture)."""
    import torch.nn.functional as F
    N = img.shape[0]
    dev, dt = img.device, torch.float32
    h = N / 2.0
    t = torch.arange(W, device=dev, dtype=dt) - float(c_det)
    s = torch.arange(N, device=dev, dtype=dt) - (N - 1) / 2.0
    th = np.asarray(fdeg, np.float64) * (math.pi / 180.0)
    out = torch.empty(len(th), W, device=dev, dtype=dt)
    src = img.to(dt)[None, None]
    for a0 in range(0, len(th), batch):
        c = torch.as_tensor(np.cos(th[a0:a0 + batch]), device=dev, dtype=dt)[:, None, None]
        sn = torch.as_tensor(np.sin(th[a0:a0 + batch]), device=dev, dtype=dt)[:, None, None]
        X = t[None, None, :] * c - s[None, :, None] * sn
        Y = t[None, None, :] * sn + s[None, :, None] * c
        gx = (X + h) * (2.0 / (N - 1)) - 1.0
        gy = (h - Y) * (2.0 / (N - 1)) - 1.0
        A = X.shape[0]
        grid = torch.stack((gx, gy), dim=-1).reshape(1, A * N, W, 2)
        v = F.grid_sample(src, grid, mode="bilinear", padding_mode="zeros", align_corners=True)
        out[a0:a0 + A] = v.reshape(A, N, W).sum(dim=1)
    return out


def poisson(sino, photons, generator=None):
    """-ln(max(k,1)/I0) with k ~ Poisson(I0 exp(-p)) (float32 out)."""
    lam = photons * torch.exp(-sino.double())
    k = torch.poisson(lam, generator=generator).clamp_min(1.0)
    return (-torch.log(k / photons)).float()


def to_int16(sino):
    """(short)(p * LOG_SCALE + 0.5), general.cpp:2502, clipped to the int16 range."""
    v = torch.trunc(sino.double() * LOG_SCALE + 0.5)
    return v.clamp(-32768, 32767).to(torch.int16)
