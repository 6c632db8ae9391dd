// rv_unet.cpp -- native runtime of the learned restoration stage: a residual
// U-Net over reconstructed slices, trained and run with the kernels of
// rv_unet.hip.  The reference has no learned stage; this is the restoration
// step that turns a sparse-view (1/8 of the angles) FBP slice into an estimate
// of the full-angle FBP slice, with plain full-angle FBP as fallback.
//
// Network (levels L, base channels C, C_l = C << l):
//   encoder level 0: edge conv 1 -> C_0 + ReLU, conv C_0 -> C_0 + ReLU (= e_0)
//   encoder level l: 2x2 average pool of e_{l-1}, conv C_{l-1} -> C_l + ReLU,
//                    conv C_l -> C_l + ReLU (= e_l); level L-1 is the bottleneck
//   decoder level l: concat[nearest 2x upsample of the level below, e_l],
//                    conv (C_{l+1} + C_l) -> C_l + ReLU, conv C_l -> C_l + ReLU
//   output:          y = x + edge conv C_0 -> 1 (no activation)
// All convolutions are 3x3 with zero padding.  Inputs are scaled by `scale`
// before the network and outputs divided by it.  Parameters are one FP32
// vector, layer by layer (weights [Cout][3][3][Cin], then biases [Cout]) in the
// order: edge-in, enc0b, enc1a, enc1b, ..., dec(L-2)a, dec(L-2)b, ..., dec0a,
// dec0b, edge-out.  The concat of level l is stored as one NHWC tensor with
// the upsampled channels first, so e_l is written straight into its slice.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "recview_hip.h"
#include "rv_common.h"
#include "rv_unet.h"

namespace rv {

enum { K_EDGE_IN = 0, K_CONV = 1, K_EDGE_OUT = 2 };

struct ULayer {
    int kind, cin, cout;
    long long w, b;                 // offsets into the parameter vector
    uint16_t* wf = nullptr;         // forward fragments
    uint16_t* wb = nullptr;         // data-gradient fragments
};

struct Unet {
    int device = 0, base = 32, levels = 4;
    std::vector<ULayer> layers;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
    long long np = 0;
    int step = 0;
    // per-shape buffers
    int B = 0, H = 0, W = 0;
    bool have_grads = false;
    std::vector<void*> act, grad;
    float *x = nullptr, *y = nullptr, *t = nullptr, *dy = nullptr, *lpart = nullptr, *part = nullptr;
    std::vector<uint16_t*> ea, cat, pool, da, d;
    uint16_t* bott = nullptr;
    std::vector<uint16_t*> g_ea, g_cat, g_pool, g_da, g_d, g_e;
    uint16_t* g_bott = nullptr;

    int C(int l) const { return base << l; }
    int cat_c(int l) const { return C(l + 1) + C(l); }
    int li_in() const { return 0; }
    int li_enc(int l, int k) const { return 2 * l + k; }   // k: 0 a (edge-in at l = 0), 1 b
    int li_dec(int l, int k) const { return 2 * levels + 2 * (levels - 2 - l) + k; }
    int li_out() const { return 2 * levels + 2 * (levels - 1); }
};

static bool ok(hipError_t e, const char* what) { return hip_ok(e, what); }

static void free_list(std::vector<void*>& v) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

static void* dalloc(std::vector<void*>& owner, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) p = nullptr;
    owner.push_back(p);   // a null entry marks the failure for ensure_shape
    return p;
}

// SplitMix64 + Box-Muller, deterministic He-normal initialisation.
struct Rng {
    unsigned long long s;
    unsigned long long next() {
        s += 0x9E3779B97F4A7C15ULL;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

static bool repack(Unet* u, hipStream_t s) {
    for (auto& L : u->layers) {
        if (L.kind != K_CONV) continue;
        if (!ok(launch_pack_conv_weights(u->P + L.w, L.cin, L.cout, 0, L.wf, s), "pack") ||
            !ok(launch_pack_conv_weights(u->P + L.w, L.cin, L.cout, 1, L.wb, s), "pack"))
            return false;
    }
    return true;
}

static void release_shape(Unet* u) {
    free_list(u->act);
    free_list(u->grad);
    u->B = u->H = u->W = 0;
    u->have_grads = false;
}

static size_t part_floats(const Unet* u, int B, int H, int W) {
    size_t m = 0;
    for (int i = 0; i < (int)u->layers.size(); i++) {
        const ULayer& L = u->layers[i];
        int lev = 0;
        for (int l = 0; l < u->levels; l++)
            if (i == u->li_enc(l, 0) || i == u->li_enc(l, 1) ||
                (l < u->levels - 1 && (i == u->li_dec(l, 0) || i == u->li_dec(l, 1))))
                lev = l;
        const int h = H >> lev, w = W >> lev;
        if (L.kind == K_CONV) {
            const int ns = wgrad_splits(L.cin, L.cout, B, h, w);
            m = std::max<size_t>(m, (size_t)ns * (9ull * L.cin * L.cout + L.cout));
        } else {
            m = std::max<size_t>(m, (size_t)edge_blocks(B, H, W) * (10ull * u->base + 1));
        }
    }
    return m;
}

static bool ensure_shape(Unet* u, int B, int H, int W, bool grads) {
    const int L = u->levels;
    if (B <= 0 || H <= 0 || W <= 0 || H % (1 << (L - 1)) || W % (1 << (L - 1))) {
        set_message("rvUnet: batch dimensions must be positive and H, W divisible by 2^(levels-1)");
        return false;
    }
    if (u->B == B && u->H == H && u->W == W && (u->have_grads || !grads)) return true;
    release_shape(u);
    auto px = [&](int l) { return (size_t)B * (H >> l) * (W >> l); };
    auto bf = [&](std::vector<void*>& o, int l, int c) {
        return (uint16_t*)dalloc(o, px(l) * c * sizeof(uint16_t));
    };
    const size_t n0 = px(0);
    u->x = (float*)dalloc(u->act, n0 * 4);
    u->y = (float*)dalloc(u->act, n0 * 4);
    u->ea.assign(L, nullptr);
    u->cat.assign(L, nullptr);
    u->pool.assign(L, nullptr);
    u->da.assign(L, nullptr);
    u->d.assign(L, nullptr);
    for (int l = 0; l < L; l++) {
        u->ea[l] = bf(u->act, l, u->C(l));
        if (l < L - 1) {
            u->cat[l] = bf(u->act, l, u->cat_c(l));
            u->pool[l] = bf(u->act, l + 1, u->C(l));
            u->da[l] = bf(u->act, l, u->C(l));
            u->d[l] = bf(u->act, l, u->C(l));
        }
    }
    u->bott = bf(u->act, L - 1, u->C(L - 1));
    if (grads) {
        u->t = (float*)dalloc(u->grad, n0 * 4);
        u->dy = (float*)dalloc(u->grad, n0 * 4);
        u->lpart = (float*)dalloc(u->grad, 1024 * 4);
        u->part = (float*)dalloc(u->grad, part_floats(u, B, H, W) * 4);
        u->g_ea.assign(L, nullptr);
        u->g_cat.assign(L, nullptr);
        u->g_pool.assign(L, nullptr);
        u->g_da.assign(L, nullptr);
        u->g_d.assign(L, nullptr);
        u->g_e.assign(L, nullptr);
        for (int l = 0; l < L; l++) {
            u->g_ea[l] = bf(u->grad, l, u->C(l));
            if (l < L - 1) {
                u->g_cat[l] = bf(u->grad, l, u->cat_c(l));
                u->g_pool[l] = bf(u->grad, l + 1, u->C(l));
                u->g_da[l] = bf(u->grad, l, u->C(l));
                u->g_d[l] = bf(u->grad, l, u->C(l));
                u->g_e[l] = bf(u->grad, l, u->C(l));
            }
        }
        u->g_bott = bf(u->grad, L - 1, u->C(L - 1));
    }
    for (void* p : u->act)
        if (!p) goto oom;
    for (void* p : u->grad)
        if (!p) goto oom;
    u->B = B;
    u->H = H;
    u->W = W;
    u->have_grads = grads;
    return true;
oom:
    release_shape(u);
    set_message("rvUnet: out of device memory for this batch shape");
    return false;
}

static bool conv_fwd(Unet* u, int li, const uint16_t* in, int in_stride, int in_off, uint16_t* out,
                     int out_stride, int out_off, int lev, hipStream_t s, uint16_t* pool_out = nullptr,
                     const uint16_t* up = nullptr, int Cup = 0) {
    const ULayer& L = u->layers[li];
    ConvArgs a{};
    a.in = in; a.in_stride = in_stride; a.in_off = in_off; a.Cin = L.cin;
    a.wfrag = L.wf; a.bias = u->P + L.b;
    a.out = out; a.out_stride = out_stride; a.out_off = out_off; a.Cout = L.cout;
    a.mask = nullptr; a.mask_stride = 0; a.mask_off = 0;
    a.relu = 1;
    a.B = u->B; a.H = u->H >> lev; a.W = u->W >> lev;
    a.up = up; a.Cup = Cup; a.pool_out = pool_out;
    return ok(launch_conv3x3(a, s), "conv3x3");
}

// dX = conv(dY, flipped W^T) masked by (mask > 0); dY contiguous with L.cout channels.
static bool conv_bwd(Unet* u, int li, const uint16_t* g, uint16_t* out, int out_stride, int out_off,
                     const uint16_t* mask, int mask_stride, int mask_off, int lev, hipStream_t s) {
    const ULayer& L = u->layers[li];
    ConvArgs a{};
    a.in = g; a.in_stride = L.cout; a.in_off = 0; a.Cin = L.cout;
    a.wfrag = L.wb; a.bias = nullptr;
    a.out = out; a.out_stride = out_stride; a.out_off = out_off; a.Cout = L.cin;
    a.mask = mask; a.mask_stride = mask_stride; a.mask_off = mask_off;
    a.relu = 0;
    a.B = u->B; a.H = u->H >> lev; a.W = u->W >> lev;
    return ok(launch_conv3x3(a, s), "conv3x3 data gradient");
}

static bool conv_wgrad(Unet* u, int li, const uint16_t* g, const uint16_t* x, int x_stride, int x_off,
                       int lev, hipStream_t s) {
    const ULayer& L = u->layers[li];
    WgradArgs a{};
    a.dy = g; a.dy_stride = L.cout; a.dy_off = 0; a.Cout = L.cout;
    a.x = x; a.x_stride = x_stride; a.x_off = x_off; a.Cin = L.cin;
    a.B = u->B; a.H = u->H >> lev; a.W = u->W >> lev;
    a.nsplit = wgrad_splits(L.cin, L.cout, a.B, a.H, a.W);
    const long long nw = 9LL * L.cin * L.cout;
    a.part = u->part;
    a.partb = u->part + (long long)a.nsplit * nw;
    return ok(launch_conv3x3_wgrad(a, s), "conv3x3 weight gradient") &&
           ok(launch_sum_splits(a.part, a.nsplit, (int)nw, u->G + L.w, s), "sum") &&
           ok(launch_sum_splits(a.partb, a.nsplit, L.cout, u->G + L.b, s), "sum");
}

// materialize: write the upsampled half of every concat (the backward pass
// reads it); inference lets the decoder convolution read the half-resolution
// tensor directly instead.
static bool forward(Unet* u, const float* in, float scale, hipStream_t s, bool materialize) {
    const int Lv = u->levels;
    const long long n0 = (long long)u->B * u->H * u->W;
    if (!ok(launch_scale(in, scale, u->x, n0, s), "scale")) return false;
    const ULayer& Li = u->layers[u->li_in()];
    // e_l into the skip slice of cat[l] (or the bottleneck tensor)
    auto e_ptr = [&](int l) { return l < Lv - 1 ? u->cat[l] : u->bott; };
    auto e_stride = [&](int l) { return l < Lv - 1 ? u->cat_c(l) : u->C(l); };
    auto e_off = [&](int l) { return l < Lv - 1 ? u->C(l + 1) : 0; };
    // e_0 and, below the bottleneck, its 2x2 average pool come out of one launch;
    // inference with 32 base channels also computes the edge convolution in that
    // launch (training keeps its output for the backward pass)
    const char* fused_env = std::getenv("RECVIEW_UNET_FUSED");
    const bool fuse = !materialize && u->C(0) == 32 && !(fused_env && std::strcmp(fused_env, "0") == 0);
    if (fuse) {
        const ULayer& L = u->layers[u->li_enc(0, 1)];
        ConvArgs a{};
        a.in = nullptr; a.in_stride = 0; a.in_off = 0; a.Cin = L.cin;
        a.wfrag = L.wf; a.bias = u->P + L.b;
        a.out = e_ptr(0); a.out_stride = e_stride(0); a.out_off = e_off(0); a.Cout = L.cout;
        a.relu = 1;
        a.B = u->B; a.H = u->H; a.W = u->W;
        a.pool_out = Lv > 1 ? u->pool[0] : nullptr;
        if (!ok(launch_edge_conv3x3_c32(a, u->x, u->P + Li.w, u->P + Li.b, s), "edge conv + conv3x3"))
            return false;
    } else if (!ok(launch_edge_in(u->x, u->P + Li.w, u->P + Li.b, u->C(0), u->ea[0], u->C(0), 0, u->B,
                                  u->H, u->W, s), "edge conv") ||
               !conv_fwd(u, u->li_enc(0, 1), u->ea[0], u->C(0), 0, e_ptr(0), e_stride(0), e_off(0), 0, s,
                         Lv > 1 ? u->pool[0] : nullptr)) {
        return false;
    }
    for (int l = 1; l < Lv; l++) {
        if (!conv_fwd(u, u->li_enc(l, 0), u->pool[l - 1], u->C(l - 1), 0, u->ea[l], u->C(l), 0, l, s) ||
            !conv_fwd(u, u->li_enc(l, 1), u->ea[l], u->C(l), 0, e_ptr(l), e_stride(l), e_off(l), l, s,
                      l < Lv - 1 ? u->pool[l] : nullptr))
            return false;
    }
    for (int l = Lv - 2; l >= 0; l--) {
        const uint16_t* src = l == Lv - 2 ? u->bott : u->d[l + 1];
        if (materialize) {
            if (!ok(launch_upsample2(src, u->C(l + 1), u->cat[l], u->cat_c(l), 0, u->B, u->H >> l,
                                     u->W >> l, s), "upsample") ||
                !conv_fwd(u, u->li_dec(l, 0), u->cat[l], u->cat_c(l), 0, u->da[l], u->C(l), 0, l, s))
                return false;
        } else if (!conv_fwd(u, u->li_dec(l, 0), u->cat[l], u->cat_c(l), 0, u->da[l], u->C(l), 0, l, s,
                             nullptr, src, u->C(l + 1))) {
            return false;
        }
        if (!conv_fwd(u, u->li_dec(l, 1), u->da[l], u->C(l), 0, u->d[l], u->C(l), 0, l, s))
            return false;
    }
    const ULayer& Lo = u->layers[u->li_out()];
    return ok(launch_edge_out(u->d[0], u->C(0), u->P + Lo.w, u->P + Lo.b, u->x, u->y, u->B, u->H, u->W,
                              s), "edge conv");
}

// Gradients of the loss whose derivative w.r.t. the normalised output is u->dy.
static bool backward(Unet* u, hipStream_t s) {
    const int Lv = u->levels;
    const int B = u->B, H = u->H, W = u->W;
    const int nb = edge_blocks(B, H, W);
    const ULayer& Lo = u->layers[u->li_out()];
    if (!ok(launch_edge_out_wgrad(u->dy, u->d[0], u->C(0), u->part, nb, B, H, W, s), "edge grad") ||
        !ok(launch_sum_splits(u->part, nb, 9 * u->C(0) + 1, u->G + Lo.w, s), "sum") ||
        !ok(launch_edge_out_bwd(u->dy, u->P + Lo.w, u->d[0], u->C(0), u->g_d[0], B, H, W, s), "edge grad"))
        return false;
    for (int l = 0; l <= Lv - 2; l++) {
        const int db = u->li_dec(l, 1), dA = u->li_dec(l, 0);
        if (!conv_wgrad(u, db, u->g_d[l], u->da[l], u->C(l), 0, l, s) ||
            !conv_bwd(u, db, u->g_d[l], u->g_da[l], u->C(l), 0, u->da[l], u->C(l), 0, l, s) ||
            !conv_wgrad(u, dA, u->g_da[l], u->cat[l], u->cat_c(l), 0, l, s) ||
            !conv_bwd(u, dA, u->g_da[l], u->g_cat[l], u->cat_c(l), 0, u->cat[l], u->cat_c(l), 0, l, s))
            return false;
        uint16_t* tgt = l == Lv - 2 ? u->g_bott : u->g_d[l + 1];
        if (!ok(launch_upsample2_bwd(u->g_cat[l], u->cat_c(l), 0, u->C(l + 1), tgt, B, H >> l, W >> l, s),
                "upsample grad"))
            return false;
    }
    {
        const int l = Lv - 1;
        const int eb = u->li_enc(l, 1), eA = u->li_enc(l, 0);
        if (!conv_wgrad(u, eb, u->g_bott, u->ea[l], u->C(l), 0, l, s) ||
            !conv_bwd(u, eb, u->g_bott, u->g_ea[l], u->C(l), 0, u->ea[l], u->C(l), 0, l, s) ||
            !conv_wgrad(u, eA, u->g_ea[l], u->pool[l - 1], u->C(l - 1), 0, l, s) ||
            !conv_bwd(u, eA, u->g_ea[l], u->g_pool[l - 1], u->C(l - 1), 0, nullptr, 0, 0, l, s))
            return false;
    }
    for (int l = Lv - 2; l >= 0; l--) {
        if (!ok(launch_pool_bwd_add(u->g_cat[l], u->cat_c(l), u->C(l + 1), u->g_pool[l], u->cat[l],
                                    u->cat_c(l), u->C(l + 1), u->C(l), u->g_e[l], B, H >> l, W >> l, s),
                "pool grad"))
            return false;
        const int eb = u->li_enc(l, 1);
        if (!conv_wgrad(u, eb, u->g_e[l], u->ea[l], u->C(l), 0, l, s) ||
            !conv_bwd(u, eb, u->g_e[l], u->g_ea[l], u->C(l), 0, u->ea[l], u->C(l), 0, l, s))
            return false;
        if (l > 0) {
            const int eA = u->li_enc(l, 0);
            if (!conv_wgrad(u, eA, u->g_ea[l], u->pool[l - 1], u->C(l - 1), 0, l, s) ||
                !conv_bwd(u, eA, u->g_ea[l], u->g_pool[l - 1], u->C(l - 1), 0, nullptr, 0, 0, l, s))
                return false;
        } else {
            const ULayer& Li = u->layers[u->li_in()];
            if (!ok(launch_edge_in_wgrad(u->g_ea[0], u->C(0), u->x, u->part, nb, B, H, W, s), "edge grad") ||
                !ok(launch_sum_splits(u->part, nb, 10 * u->C(0), u->G + Li.w, s), "sum"))
                return false;
        }
    }
    return true;
}

static bool loss_and_grad(Unet* u, const float* in, const float* target, float scale, float* loss,
                          hipStream_t s) {
    const long long n0 = (long long)u->B * u->H * u->W;
    if (!forward(u, in, scale, s, true) || !ok(launch_scale(target, scale, u->t, n0, s), "scale") ||
        !ok(launch_mse(u->y, u->t, u->dy, u->lpart, n0, s), "loss") || !backward(u, s))
        return false;
    if (loss) {
        const int nb = loss_blocks(n0);
        std::vector<float> h(nb);
        if (!ok(hipMemcpyAsync(h.data(), u->lpart, nb * 4, hipMemcpyDeviceToHost, s), "loss copy") ||
            !ok(hipStreamSynchronize(s), "sync"))
            return false;
        double acc = 0;
        for (float v : h) acc += v;
        *loss = (float)(acc / (double)n0);
    }
    return true;
}

}  // namespace rv

using namespace rv;

struct rv_unet_opaque : Unet {};

extern "C" {

int rvUnetCreate(rv_unet_t* out, int device, int base_channels, int levels, unsigned long long seed) {
    if (!out) return -1;
    *out = nullptr;
    if (levels < 2 || levels > 6 || (base_channels != 32 && base_channels != 64)) {
        set_message("rvUnetCreate: levels must be 2..6 and base_channels 32 or 64");
        return -1;
    }
    if (!ok(hipSetDevice(device), "hipSetDevice")) return -1;
    auto* u = new rv_unet_opaque();
    u->device = device;
    u->base = base_channels;
    u->levels = levels;
    auto add = [&](int kind, int cin, int cout) {
        ULayer L;
        L.kind = kind;
        L.cin = cin;
        L.cout = cout;
        L.w = u->np;
        u->np += 9LL * cin * cout;
        L.b = u->np;
        u->np += cout;
        u->layers.push_back(L);
    };
    add(K_EDGE_IN, 1, u->C(0));
    add(K_CONV, u->C(0), u->C(0));
    for (int l = 1; l < levels; l++) {
        add(K_CONV, u->C(l - 1), u->C(l));
        add(K_CONV, u->C(l), u->C(l));
    }
    for (int l = levels - 2; l >= 0; l--) {
        add(K_CONV, u->cat_c(l), u->C(l));
        add(K_CONV, u->C(l), u->C(l));
    }
    add(K_EDGE_OUT, u->C(0), 1);
    std::vector<float> h(u->np, 0.f);
    Rng rng{seed};
    for (auto& L : u->layers) {
        const double sd = L.kind == K_EDGE_OUT ? 0.1 * std::sqrt(1.0 / (9.0 * L.cin))
                                               : std::sqrt(2.0 / (9.0 * L.cin));
        for (long long i = 0; i < 9LL * L.cin * L.cout; i++) h[L.w + i] = (float)(sd * rng.normal());
    }
    bool good = ok(hipMalloc(&u->P, u->np * 4), "hipMalloc") && ok(hipMalloc(&u->G, u->np * 4), "hipMalloc") &&
                ok(hipMalloc(&u->M, u->np * 4), "hipMalloc") && ok(hipMalloc(&u->V, u->np * 4), "hipMalloc");
    for (auto& L : u->layers)
        if (good && L.kind == K_CONV)
            good = ok(hipMalloc(&L.wf, 9LL * L.cin * L.cout * 2), "hipMalloc") &&
                   ok(hipMalloc(&L.wb, 9LL * L.cin * L.cout * 2), "hipMalloc");
    good = good && ok(hipMemcpy(u->P, h.data(), u->np * 4, hipMemcpyHostToDevice), "upload") &&
           ok(hipMemset(u->M, 0, u->np * 4), "memset") && ok(hipMemset(u->V, 0, u->np * 4), "memset") &&
           ok(hipMemset(u->G, 0, u->np * 4), "memset") && repack(u, nullptr) &&
           ok(hipDeviceSynchronize(), "sync");
    if (!good) {
        rvUnetDestroy(u);
        return -1;
    }
    *out = u;
    return 0;
}

long long rvUnetParamCount(rv_unet_t u) { return u ? u->np : -1; }

int rvUnetGetParams(rv_unet_t u, float* host, long long n) {
    if (!u || !host || n != u->np) return -1;
    if (!ok(hipSetDevice(u->device), "hipSetDevice")) return -1;
    return ok(hipMemcpy(host, u->P, n * 4, hipMemcpyDeviceToHost), "download") ? 0 : -1;
}

int rvUnetSetParams(rv_unet_t u, const float* host, long long n) {
    if (!u || !host || n != u->np) return -1;
    if (!ok(hipSetDevice(u->device), "hipSetDevice")) return -1;
    u->step = 0;
    return ok(hipMemcpy(u->P, host, n * 4, hipMemcpyHostToDevice), "upload") &&
                   ok(hipMemset(u->M, 0, n * 4), "memset") && ok(hipMemset(u->V, 0, n * 4), "memset") &&
                   repack(u, nullptr) && ok(hipDeviceSynchronize(), "sync")
               ? 0
               : -1;
}

int rvUnetForward(rv_unet_t u, const float* d_in, float* d_out, int B, int H, int W, float scale,
                  rv_stream_t stream) {
    if (!u || !d_in || !d_out || !(scale > 0.f)) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (!ok(hipSetDevice(u->device), "hipSetDevice") || !ensure_shape(u, B, H, W, u->have_grads && u->B == B &&
                                                                                 u->H == H && u->W == W))
        return -1;
    return forward(u, d_in, scale, s, false) &&
                   ok(launch_scale(u->y, 1.f / scale, d_out, (long long)B * H * W, s), "scale")
               ? 0
               : -1;
}

int rvUnetLossGrad(rv_unet_t u, const float* d_in, const float* d_target, int B, int H, int W,
                   float scale, float* loss, float* host_grads, rv_stream_t stream) {
    if (!u || !d_in || !d_target || !(scale > 0.f)) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (!ok(hipSetDevice(u->device), "hipSetDevice") || !ensure_shape(u, B, H, W, true)) return -1;
    if (!loss_and_grad(u, d_in, d_target, scale, loss, s)) return -1;
    if (host_grads &&
        (!ok(hipMemcpyAsync(host_grads, u->G, u->np * 4, hipMemcpyDeviceToHost, s), "download") ||
         !ok(hipStreamSynchronize(s), "sync")))
        return -1;
    return 0;
}

int rvUnetTrainStep(rv_unet_t u, const float* d_in, const float* d_target, int B, int H, int W,
                    float scale, float lr, float* loss, rv_stream_t stream) {
    if (!u || !d_in || !d_target || !(scale > 0.f) || !(lr >= 0.f)) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (!ok(hipSetDevice(u->device), "hipSetDevice") || !ensure_shape(u, B, H, W, true)) return -1;
    if (!loss_and_grad(u, d_in, d_target, scale, loss, s)) return -1;
    u->step++;
    return ok(launch_adam(u->P, u->G, u->M, u->V, u->np, lr, 0.9f, 0.999f, 1e-8f, u->step, s), "adam") &&
                   repack(u, s)
               ? 0
               : -1;
}

void rvUnetDestroy(rv_unet_t u) {
    if (!u) return;
    (void)hipSetDevice(u->device);
    (void)hipDeviceSynchronize();
    release_shape(u);
    for (auto& L : u->layers) {
        if (L.wf) (void)hipFree(L.wf);
        if (L.wb) (void)hipFree(L.wb);
    }
    if (u->P) (void)hipFree(u->P);
    if (u->G) (void)hipFree(u->G);
    if (u->M) (void)hipFree(u->M);
    if (u->V) (void)hipFree(u->V);
    delete u;
}

// Single-layer entry points (building blocks and their tests).
int rvConv3x3Bf16(const unsigned short* d_in, int in_stride, int in_off, int Cin, const float* d_w,
                  const float* d_bias, unsigned short* d_out, int out_stride, int out_off, int Cout,
                  const unsigned short* d_mask, int mask_stride, int mask_off, int relu, int data_gradient,
                  int B, int H, int W, rv_stream_t stream) {
    hipStream_t s = (hipStream_t)stream;
    const int cin_k = data_gradient ? Cout : Cin, cout_k = data_gradient ? Cin : Cout;
    // data_gradient: d_in holds dY (Cout channels of the forward layer), d_out dX (Cin channels).
    if (!d_in || !d_w || !d_out || cin_k % 32 || cout_k % 32) {
        set_message("rvConv3x3Bf16: channel counts must be multiples of 32");
        return -1;
    }
    uint16_t* wf = nullptr;
    if (!ok(hipMalloc(&wf, 9LL * Cin * Cout * 2), "hipMalloc")) return -1;
    ConvArgs a{};
    a.in = d_in; a.in_stride = in_stride; a.in_off = in_off; a.Cin = cin_k;
    a.wfrag = wf; a.bias = data_gradient ? nullptr : d_bias;
    a.out = d_out; a.out_stride = out_stride; a.out_off = out_off; a.Cout = cout_k;
    a.mask = d_mask; a.mask_stride = mask_stride; a.mask_off = mask_off;
    a.relu = relu; a.B = B; a.H = H; a.W = W;
    const bool good = ok(launch_pack_conv_weights(d_w, Cin, Cout, data_gradient, wf, s), "pack") &&
                      ok(launch_conv3x3(a, s), "conv3x3") && ok(hipStreamSynchronize(s), "sync");
    (void)hipFree(wf);
    return good ? 0 : -1;
}

int rvConv3x3WgradBf16(const unsigned short* d_dy, int dy_stride, int dy_off, int Cout, const unsigned short* d_x,
                       int x_stride, int x_off, int Cin, float* d_dw, float* d_db, int B, int H, int W,
                       rv_stream_t stream) {
    hipStream_t s = (hipStream_t)stream;
    if (!d_dy || !d_x || !d_dw || Cin % 32 || Cout % 32) {
        set_message("rvConv3x3WgradBf16: channel counts must be multiples of 32");
        return -1;
    }
    WgradArgs a{};
    a.dy = d_dy; a.dy_stride = dy_stride; a.dy_off = dy_off; a.Cout = Cout;
    a.x = d_x; a.x_stride = x_stride; a.x_off = x_off; a.Cin = Cin;
    a.B = B; a.H = H; a.W = W;
    a.nsplit = wgrad_splits(Cin, Cout, B, H, W);
    const long long nw = 9LL * Cin * Cout;
    float* part = nullptr;
    if (!ok(hipMalloc(&part, (size_t)a.nsplit * (nw + Cout) * 4), "hipMalloc")) return -1;
    a.part = part;
    a.partb = part + (long long)a.nsplit * nw;
    bool good = ok(launch_conv3x3_wgrad(a, s), "wgrad") &&
                ok(launch_sum_splits(a.part, a.nsplit, (int)nw, d_dw, s), "sum");
    if (good && d_db) good = ok(launch_sum_splits(a.partb, a.nsplit, Cout, d_db, s), "sum");
    good = good && ok(hipStreamSynchronize(s), "sync");
    (void)hipFree(part);
    return good ? 0 : -1;
}

}  // extern "C"
