"""CPU FP64 restatement of the learned restoration stage -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's baseline leg use this module,
as the checker of the GPU network (recview_amd/csrc/rv_unet.{cpp,hip}); no
product path imports it.  The reference has no learned stage, so nothing here
follows a reference file: the network is defined by the C ABI
(include/recview_hip.h, Part 3) and restated here in numpy float64.

Network (levels L, base channels C, C_l = C << l), NHWC tensors:
  e_0 = relu(conv(relu(conv_in(x))))                      (conv_in: 1 -> C_0)
  e_l = relu(conv(relu(conv(avgpool2(e_{l-1})))))         l = 1 .. L-1
  d_l = relu(conv(relu(conv(concat[up2(d_{l+1} or e_{L-1}), e_l]))))  l = L-2 .. 0
  y   = x + conv_out(d_0)                                 (conv_out: C_0 -> 1)
x is the input image times `scale`; the loss is mean((y - scale*target)^2).

`rnd` reproduces the GPU's storage precision: every stored activation and
activation gradient is rounded to bfloat16 (round to nearest even), the MFMA
layers use bfloat16-rounded weights, the two edge layers FP32 weights; with
`rnd=None` the whole computation is exact float64 (used for finite differences).

How this restatement is pinned (tests/test_unet_oracle.py):
  * conv3x3 forward against torch.nn.functional.conv2d in float64;
  * every parameter gradient against central finite differences of the loss;
  * bf16_round against torch's float32 -> bfloat16 conversion.
"""
import numpy as np

EDGE_IN, CONV, EDGE_OUT = 0, 1, 2


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def layout(base, levels):
    """[(kind, cin, cout, w_offset, b_offset)] in parameter-vector order."""
    C = lambda l: base << l
    specs = [(EDGE_IN, 1, C(0)), (CONV, C(0), C(0))]
    for l in range(1, levels):
        specs += [(CONV, C(l - 1), C(l)), (CONV, C(l), C(l))]
    for l in range(levels - 2, -1, -1):
        specs += [(CONV, C(l + 1) + C(l), C(l)), (CONV, C(l), C(l))]
    specs.append((EDGE_OUT, C(0), 1))
    out, off = [], 0
    for kind, cin, cout in specs:
        w = off
        off += 9 * cin * cout
        out.append((kind, cin, cout, w, off))
        off += cout
    return out


def param_count(base, levels):
    k, cin, cout, w, b = layout(base, levels)[-1]
    return b + cout


def conv3x3(x, w, b=None):
    """y[n,i,j,o] = b[o] + sum_{ky,kx,c} w[o,ky,kx,c] x[n,i+ky-1,j+kx-1,c] (zero padding)."""
    B, H, W, Cin = x.shape
    xp = np.zeros((B, H + 2, W + 2, Cin))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((B, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += xp[:, ky:ky + H, kx:kx + W, :] @ w[:, ky, kx, :].T
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(g, x):
    """(dW[o,ky,kx,c], db[o]) of conv3x3 given dY = g and input x."""
    B, H, W, Cin = x.shape
    xp = np.zeros((B, H + 2, W + 2, Cin))
    xp[:, 1:-1, 1:-1] = x
    g2 = g.reshape(-1, g.shape[-1])
    dw = np.zeros((g.shape[-1], 3, 3, Cin))
    for ky in range(3):
        for kx in range(3):
            dw[:, ky, kx, :] = g2.T @ xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, Cin)
    return dw, g2.sum(axis=0)


def conv3x3_dgrad(g, w):
    """dX of conv3x3 given dY = g."""
    B, H, W, _ = g.shape
    gp = np.zeros((B, H + 2, W + 2, w.shape[3]))
    for ky in range(3):
        for kx in range(3):
            gp[:, ky:ky + H, kx:kx + W, :] += g @ w[:, ky, kx, :]
    return gp[:, 1:-1, 1:-1]


def avgpool2(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up2_bwd(g):
    B, H, W, C = g.shape
    return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


class UnetRef:
    def __init__(self, base=32, levels=4):
        self.base, self.levels = base, levels
        self.layers = layout(base, levels)
        self.n = param_count(base, levels)

    def C(self, l):
        return self.base << l

    def unpack(self, params, rnd):
        ws = []
        for kind, cin, cout, w, b in self.layers:
            W = np.asarray(params[w:w + 9 * cin * cout], np.float64).reshape(cout, 3, 3, cin)
            if kind == CONV:
                W = rnd(W)
            ws.append((W, np.asarray(params[b:b + cout], np.float64)))
        return ws

    def forward(self, params, x, scale, rnd=bf16_round):
        """x: [B, H, W] image; returns (y_normalised, cache)."""
        rnd = rnd or _ident
        L = self.levels
        ws = self.unpack(params, rnd)
        li = 0
        xn = np.asarray(x, np.float64) * scale
        c = {"x": xn, "ea": [None] * L, "e": [None] * L, "pool": [None] * L, "cat": [None] * L,
             "da": [None] * L, "d": [None] * L, "ws": ws}
        for l in range(L):
            if l == 0:
                inp = xn[..., None]
            else:
                c["pool"][l - 1] = rnd(avgpool2(c["e"][l - 1]))
                inp = c["pool"][l - 1]
            W, b = ws[li]
            c["ea"][l] = rnd(np.maximum(conv3x3(inp, W, b), 0))
            W, b = ws[li + 1]
            c["e"][l] = rnd(np.maximum(conv3x3(c["ea"][l], W, b), 0))
            li += 2
        for l in range(L - 2, -1, -1):
            src = c["e"][L - 1] if l == L - 2 else c["d"][l + 1]
            c["cat"][l] = np.concatenate([up2(src), c["e"][l]], axis=-1)
            W, b = ws[li]
            c["da"][l] = rnd(np.maximum(conv3x3(c["cat"][l], W, b), 0))
            W, b = ws[li + 1]
            c["d"][l] = rnd(np.maximum(conv3x3(c["da"][l], W, b), 0))
            li += 2
        W, b = ws[li]
        y = xn + conv3x3(c["d"][0], W, b)[..., 0]
        return y, c

    def loss_grad(self, params, x, target, scale, rnd=bf16_round):
        """(loss, flat gradient) of mean((y - scale*target)^2)."""
        r = rnd or _ident
        L = self.levels
        y, c = self.forward(params, x, scale, rnd)
        ws = c["ws"]
        t = np.asarray(target, np.float64) * scale
        n = y.size
        loss = float(np.mean((y - t) ** 2))
        dy = 2.0 * (y - t) / n
        G = np.zeros(self.n)
        idx = {}
        li = 0
        for l in range(L):
            idx[("enc", l)] = (li, li + 1)
            li += 2
        for l in range(L - 2, -1, -1):
            idx[("dec", l)] = (li, li + 1)
            li += 2
        idx["out"] = li

        def put(i, dw, db):
            kind, cin, cout, w, b = self.layers[i]
            G[w:w + 9 * cin * cout] = dw.reshape(-1)
            G[b:b + cout] = db

        mask = lambda a: (a > 0).astype(np.float64)
        # output edge layer
        dw, db = conv3x3_wgrad(dy[..., None], c["d"][0])
        put(idx["out"], dw, db)
        g = r(mask(c["d"][0]) * conv3x3_dgrad(dy[..., None], ws[idx["out"]][0]))
        g_up = None
        for l in range(0, L - 1):
            ia, ib = idx[("dec", l)]
            if l > 0:
                g = g_up
            dw, db = conv3x3_wgrad(g, c["da"][l])
            put(ib, dw, db)
            g_da = r(mask(c["da"][l]) * conv3x3_dgrad(g, ws[ib][0]))
            dw, db = conv3x3_wgrad(g_da, c["cat"][l])
            put(ia, dw, db)
            c.setdefault("g_cat", [None] * L)[l] = r(mask(c["cat"][l]) * conv3x3_dgrad(g_da, ws[ia][0]))
            g_up = r(up2_bwd(c["g_cat"][l][..., :self.C(l + 1)]))
        # bottleneck (g_up is the gradient of e_{L-1})
        l = L - 1
        ia, ib = idx[("enc", l)]
        dw, db = conv3x3_wgrad(g_up, c["ea"][l])
        put(ib, dw, db)
        g_ea = r(mask(c["ea"][l]) * conv3x3_dgrad(g_up, ws[ib][0]))
        dw, db = conv3x3_wgrad(g_ea, c["pool"][l - 1])
        put(ia, dw, db)
        g_pool = r(conv3x3_dgrad(g_ea, ws[ia][0]))
        for l in range(L - 2, -1, -1):
            ia, ib = idx[("enc", l)]
            skip = c["g_cat"][l][..., self.C(l + 1):]
            g_e = r((skip + 0.25 * up2(g_pool)) * mask(c["e"][l]))
            dw, db = conv3x3_wgrad(g_e, c["ea"][l])
            put(ib, dw, db)
            g_ea = r(mask(c["ea"][l]) * conv3x3_dgrad(g_e, ws[ib][0]))
            inp = c["x"][..., None] if l == 0 else c["pool"][l - 1]
            dw, db = conv3x3_wgrad(g_ea, inp)
            put(ia, dw, db)
            if l > 0:
                g_pool = r(conv3x3_dgrad(g_ea, ws[ia][0]))
        return loss, G


def init_params(base, levels, seed):
    """He-normal weights, zero biases (any deterministic init serves the tests)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(base, levels))
    for kind, cin, cout, w, b in layout(base, levels):
        sd = np.sqrt(2.0 / (9 * cin)) if kind != EDGE_OUT else 0.1 * np.sqrt(1.0 / (9 * cin))
        p[w:w + 9 * cin * cout] = rng.standard_normal(9 * cin * cout) * sd
        p[b:b + cout] = rng.standard_normal(cout) * 0.05
    return p
