"""Sparse-view restoration with the native U-Net: training pairs, training, evaluation.

Pairs come from the FBP engine itself: a seeded phantom (random ellipses plus a
fine texture of small ellipses) is projected analytically over P angles; the
full-angle FBP (all P rows) is the target and the sparse-view FBP (every 8th
row, P/8 angles) the input.  Both reconstructions run through one FbpPlan
(LINEAR interpolation, FP32), so the network learns to remove exactly the
streaks this engine's sparse-view FBP produces.  Both images are normalised
to attenuation per detector sample (2 pi / projections used), the units of the
phantom.  Full-angle FBP stays the fallback and the error reference for the
restored slice.

  python -m recview_amd.learned train --out weights.npz   (GPU)
  python -m recview_amd.learned eval --weights weights.npz
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import fbp, phantom
from .unet import Unet

SPARSE = 8
W_DEFAULT = 2048
P_DEFAULT = 1800
FILTER = "shepp-logan"


def phantom_params(seed, W):
    """(ellipses, texture) of one slice; attenuation per detector sample."""
    ells = phantom.random_ellipses(seed, count=10, mu_range=(0.0, 2.0 / W))
    rng = phantom.SplitMix64(seed ^ 0x5DEECE66D)
    tex = []
    amp = 0.02 * 2.0 / W
    for _ in range(400):
        a = rng.uniform(0.003, 0.012)
        b = rng.uniform(0.003, 0.012)
        r = 0.45 * math.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2 * math.pi)
        phi = rng.uniform(0.0, math.pi)
        mu = amp * rng.uniform(-1.0, 1.0)
        tex.append((r * math.cos(ang), r * math.sin(ang), a, b, phi, mu))
    return ells, np.array(tex, dtype=np.float64)


def texture_sinogram(tex, W, fdeg, c_det, scale, device):
    """Analytic sinogram of many small ellipses, evaluated only on each one's
    detector footprint (a window of samples around its projected centre)."""
    dev = torch.device(device)
    th = torch.as_tensor(np.asarray(fdeg, np.float64) * (math.pi / 180.0), device=dev)
    T = torch.as_tensor(tex, device=dev)
    cx, cy, a, b, phi, mu = (T[:, i:i + 1] * (scale if i < 4 else 1.0) for i in range(6))
    c, s = torch.cos(th)[None, :], torch.sin(th)[None, :]
    t0 = cx * c + cy * s                                        # [E, P]
    ca, sa = torch.cos(th[None, :] - phi), torch.sin(th[None, :] - phi)
    s2 = (a * ca) ** 2 + (b * sa) ** 2
    hw = torch.sqrt(s2)
    K = int(2 * float(hw.max()) + 3)
    k0 = torch.floor(t0 + c_det - hw).long()                    # first sample of the window
    ks = k0[..., None] + torch.arange(K, device=dev)            # [E, P, K]
    tt = ks.double() - c_det - t0[..., None]
    d2 = s2[..., None] - tt ** 2
    val = torch.where(d2 > 0, (2.0 * mu * a * b)[..., None] * torch.sqrt(d2.clamp_min(0)) / s2[..., None],
                      torch.zeros((), device=dev, dtype=torch.float64))
    valid = (ks >= 0) & (ks < W)
    rows = torch.arange(len(fdeg), device=dev)[None, :, None].expand_as(ks)
    out = torch.zeros(len(fdeg) * W, device=dev, dtype=torch.float64)
    out.index_add_(0, (rows * W + ks.clamp(0, W - 1))[valid], val[valid])
    return out.view(len(fdeg), W)


def slice_sinogram(seed, W, P, device):
    ells, tex = phantom_params(seed, W)
    fdeg = phantom.angles_deg(P)
    c_det = W / 2.0
    sino = phantom.sinogram(ells, W, fdeg, c_det, scale=W, device=device)
    sino += texture_sinogram(tex, W, fdeg, c_det, W, device)
    return sino.float(), fdeg, c_det


def slice_image(seed, W, device):
    ells, tex = phantom_params(seed, W)
    img = phantom.image(ells, W, W, device=device)
    img += phantom.image(tex, W, W, device=device)
    return img.float()


class PairMaker:
    """Full- and sparse-view FBP of `batch` slices per call, one plan reused."""

    def __init__(self, W=W_DEFAULT, P=P_DEFAULT, batch=16, device=0):
        self.W, self.P, self.batch = W, P, batch
        self.dev = torch.device("cuda", device)
        fdeg = phantom.angles_deg(P)
        c = W / 2.0
        specs = [fbp.SliceSpec(W, fdeg, c, filter=FILTER) for _ in range(batch)]
        specs += [fbp.SliceSpec(W, fdeg[::SPARSE], c, filter=FILTER) for _ in range(batch)]
        self.plan = fbp.FbpPlan(specs, interp="linear", store="f32", persistent=True, normalize=True,
                                device=device)
        self.N = self.plan.out_n(0)
        self.full = [torch.empty(self.N, self.N, device=self.dev) for _ in range(batch)]
        self.sparse = [torch.empty(self.N, self.N, device=self.dev) for _ in range(batch)]

    def make(self, seeds):
        assert len(seeds) == self.batch
        sinos = [slice_sinogram(s, self.W, self.P, self.dev)[0].contiguous() for s in seeds]
        sparse = [s[::SPARSE].contiguous() for s in sinos]
        self.plan.execute(sinos + sparse, self.full + self.sparse)
        torch.cuda.synchronize()
        return self.full, self.sparse


def random_crops(n, N, crop, rng):
    """Top-left corners of n crops fully inside the reconstruction circle."""
    R = N / 2.0 - crop / math.sqrt(2.0) - 2
    out = []
    while len(out) < n:
        x, y = rng.uniform(-R, R, 2)
        if x * x + y * y <= R * R:
            out.append((int(N / 2 + y - crop / 2), int(N / 2 + x - crop / 2)))
    return out


def make_dataset(n_crops, seed0=1, crop=512, crops_per_slice=8, W=W_DEFAULT, P=P_DEFAULT, batch=16,
                 device=0, log=print):
    """(inputs, targets) float32 [n_crops, crop, crop] on the device."""
    pm = PairMaker(W, P, batch, device)
    X = torch.empty(n_crops, crop, crop, device=pm.dev)
    Y = torch.empty(n_crops, crop, crop, device=pm.dev)
    rng = np.random.default_rng(seed0)
    k, s = 0, seed0
    t0 = time.time()
    while k < n_crops:
        seeds = list(range(s, s + batch))
        s += batch
        full, sparse = pm.make(seeds)
        for f, sp in zip(full, sparse):
            for (y, x) in random_crops(crops_per_slice, pm.N, crop, rng):
                if k >= n_crops:
                    break
                X[k] = sp[y:y + crop, x:x + crop]
                Y[k] = f[y:y + crop, x:x + crop]
                k += 1
        if log and (s - seed0) % (batch * 40) == 0:
            log(f"dataset: {k}/{n_crops} crops ({time.time() - t0:.1f} s)")
    pm.plan.close()
    return X, Y


def train(net, X, Y, scale, steps, batch=16, lr=5e-4, lr_min=2e-5, log=print, log_every=100, seed=0):
    """Adam with a cosine learning-rate schedule over random mini-batches."""
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    n = X.shape[0]
    hist = []
    t0 = time.time()
    xb = torch.empty(batch, *X.shape[1:], device=X.device)
    yb = torch.empty_like(xb)
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(X.device)
        torch.index_select(X, 0, idx, out=xb)
        torch.index_select(Y, 0, idx, out=yb)
        # random flips/transposes keep the streak statistics and enlarge the set
        k = it % 8
        xa, ya = xb, yb
        if k & 1:
            xa, ya = xa.flip(-1), ya.flip(-1)
        if k & 2:
            xa, ya = xa.flip(-2), ya.flip(-2)
        if k & 4:
            xa, ya = xa.transpose(-1, -2), ya.transpose(-1, -2)
        xa, ya = xa.contiguous(), ya.contiguous()
        cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / max(1, steps)))
        want = (it % log_every == 0) or it == steps - 1
        loss = net.train_step(xa, ya, scale, cur, want_loss=want)
        if want:
            hist.append((it, loss))
            if log:
                log(f"step {it} loss {loss:.4e} lr {cur:.2e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return hist, time.time() - t0


def psnr(img, ref, mask, rng_):
    d = (img - ref)[mask]
    mse = float((d.double() ** 2).mean())
    return 10.0 * math.log10(rng_ ** 2 / mse)


def ssim(img, ref, mask, rng_):
    """Mean SSIM (Gaussian window 11, sigma 1.5, K1 0.01, K2 0.03) over the mask."""
    import torch.nn.functional as F
    x = img[None, None].double()
    y = ref[None, None].double()
    r = torch.arange(11, dtype=torch.float64, device=img.device) - 5
    gk = torch.exp(-r ** 2 / (2 * 1.5 ** 2))
    gk = gk / gk.sum()
    k2 = (gk[:, None] * gk[None, :])[None, None]
    f = lambda t: F.conv2d(t, k2, padding=5)
    mx, my = f(x), f(y)
    sxx, syy, sxy = f(x * x) - mx * mx, f(y * y) - my * my, f(x * y) - mx * my
    c1, c2 = (0.01 * rng_) ** 2, (0.03 * rng_) ** 2
    m = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
    return float(m[0, 0][mask].mean())


def evaluate(net, scale, seeds, W=W_DEFAULT, P=P_DEFAULT, device=0):
    """PSNR / SSIM of sparse FBP, sparse FBP + network and full FBP, against the
    phantom and against full-angle FBP, inside 0.9 of the reconstruction circle."""
    pm = PairMaker(W, P, len(seeds), device)
    full, sparse = pm.make(list(seeds))
    N = pm.N
    yy, xx = torch.meshgrid(torch.arange(N, device=pm.dev), torch.arange(N, device=pm.dev), indexing="ij")
    mask = ((xx - N / 2 + 0.5) ** 2 + (yy - N / 2 + 0.5) ** 2) <= (0.9 * N / 2) ** 2
    res = []
    for s, f, sp in zip(seeds, full, sparse):
        truth = slice_image(s, W, pm.dev)
        restored = net.forward(sp.contiguous(), scale)
        rng_ = float(truth[mask].max() - truth[mask].min())
        rf = float(f[mask].max() - f[mask].min())
        res.append({
            "seed": s,
            "psnr_phantom": {"sparse": psnr(sp, truth, mask, rng_), "restored": psnr(restored, truth, mask, rng_),
                             "full": psnr(f, truth, mask, rng_)},
            "psnr_vs_full": {"sparse": psnr(sp, f, mask, rf), "restored": psnr(restored, f, mask, rf)},
            "ssim_vs_full": {"sparse": ssim(sp, f, mask, rf), "restored": ssim(restored, f, mask, rf)},
        })
    pm.plan.close()
    return res


def summarize(res):
    keys = lambda d: {k: float(np.mean([r[d][k] for r in res])) for k in res[0][d]}
    return {d: keys(d) for d in ("psnr_phantom", "psnr_vs_full", "ssim_vs_full")}


def main(argv=None):
    ap = argparse.ArgumentParser(description="train / evaluate the sparse-view restoration U-Net")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--crops", type=int, default=20000)
    t.add_argument("--steps", type=int, default=2500)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=5e-4)
    t.add_argument("--width", type=int, default=W_DEFAULT)
    t.add_argument("--proj", type=int, default=P_DEFAULT)
    t.add_argument("--out", required=True)
    t.add_argument("--eval-seeds", type=int, default=4)
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--seeds", type=int, default=4)
    a = ap.parse_args(argv)
    log = lambda m: (print(m), sys.stdout.flush())
    if a.cmd == "train":
        scale = a.width / 2.0
        X, Y = make_dataset(a.crops, W=a.width, P=a.proj, log=log)
        log(f"dataset: input*scale mean {float(X.mean()) * scale:.3f} std {float(X.std()) * scale:.3f}; "
            f"target-input rms {float((Y - X).pow(2).mean().sqrt()) * scale:.4f}")
        net = Unet(seed=12345)
        hist, secs = train(net, X, Y, scale, a.steps, a.batch, a.lr, log=log)
        del X, Y
        torch.cuda.empty_cache()
        res = evaluate(net, scale, range(1_000_001, 1_000_001 + a.eval_seeds), a.width, a.proj)
        summ = summarize(res)
        log(json.dumps({"train_seconds": secs, "steps": a.steps, "batch": a.batch, "crops": a.crops,
                        "final_loss": hist[-1][1], "eval": summ}))
        net.save(a.out, scale, meta={"train_steps": a.steps, "train_crops": a.crops,
                                     "loss_history": np.array(hist, np.float64)})
    else:
        net = Unet.load(a.weights)
        res = evaluate(net, net.scale, range(1_000_001, 1_000_001 + a.seeds))
        log(json.dumps(summarize(res)))


if __name__ == "__main__":
    main()
