// rv_common.cpp -- geometry, filter design, twiddle cache and error plumbing.
#include "rv_common.h"
#include "recview_hip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace rv {

// constants.h:3,6-8
static constexpr double kDegToRad = 0.0174532925199;
static constexpr double kPi = 3.1415926535897;
static constexpr double kLog2 = 0.69314718055994530941723212145818;

Geom geometry(int ixdim, int iInterpolation, double center, bool float_center)
{
    Geom g{};
    g.W = ixdim;
    g.izoom = (iInterpolation > 2) ? (iInterpolation - 2) : 0;        // cudaReconstHost.cpp:636
    g.z = (int)std::pow(2.0, g.izoom);                                  // :637
    g.ndimp = (ixdim > 0) ? (int)(std::log((double)ixdim) / kLog2) + 1 + g.izoom : 0;  // :638
    g.n = (int)std::pow(2.0, g.ndimp);                                  // :639
    g.N = ixdim * g.z;
    g.h = g.N / 2;
    const float fC = (float)center;
    if (float_center) {
        g.ic = (int)fC;                                                 // :655
        g.fcenter16 = (g.z == 1) ? (g.h + fC - (int)fC) * 16            // cudaReconst.cu:220-221
                                 : (g.h + (fC - (int)fC) * g.z) * 16;
    } else {
        g.ic = (int)center;                                             // general.cpp:2109
        g.fcenter16 = (g.z == 1) ? (g.h + (float)center - (int)center) * 16   // general.cpp:2034
                                 : (g.h + ((float)center - (int)center) * g.z) * 16;
    }
    g.ihoffset = g.n / 2 - 1 - g.h;                                     // :654
    return g;
}

void set_filter(int ndim, int ifilter, double cutoff, double order, float* f)
{
    const int ndimh = (int)(ndim * cutoff);                             // gazoDoc.cpp:3761
    const float forder = (float)order;
    const float revndim = 1.f / ndim;
    switch (ifilter) {
    case 0:                                                             // Hann :3767-3777
        for (int i = 0; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = 0.5f * revndim * i * (float)(1 + std::cos(kPi * frac));
        }
        for (int i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 1:                                                             // Hamming :3778-3784
        for (int i = 0; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * (float)(0.54 + 0.46 * std::cos(kPi * frac));
        }
        for (int i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 2:                                                             // Ramp :3785-3788
        for (int i = 0; i < ndimh; i++) f[i] = revndim * i;
        for (int i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 3:                                                             // Parzen :3789-3799
        for (int i = 0; i < ndimh / 2; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * (1 - 6 * frac * frac * (1 - frac));
        }
        for (int i = ndimh / 2; i < ndimh; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i * 2 * (1 - frac) * (1 - frac) * (1 - frac);
        }
        for (int i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    case 4:                                                             // Butterworth :3800-3805
        for (int i = 0; i < ndim; i++) {
            const float frac = (float)i / ndimh;
            f[i] = revndim * i / (float)std::sqrt(1 + std::pow(frac, forder));
        }
        break;
    case 5:                                                             // Shepp-Logan (new)
        for (int i = 0; i < ndimh; i++) {
            const double x = kPi * (double)i / (2.0 * ndimh);
            const double s = (i == 0) ? 1.0 : std::sin(x) / x;
            f[i] = (float)((double)i / ndim * s);
        }
        for (int i = ndimh; i < ndim; i++) f[i] = 0;
        break;
    default:
        for (int i = 0; i < ndim; i++) f[i] = 1;
        break;
    }
}

double theta_rad(float fdeg, float tilt)
{
    return (fdeg + tilt) * kDegToRad;                                   // cudaReconstHost.cpp:776
}

void lut_params(float fdeg, float tilt, const Geom& g, int iCenterOffset, float* fcos, float* fsin,
                float* foffset)
{
    const double th = theta_rad(fdeg, tilt);
    const float c = (float)(std::cos(th) * 16);                        // cudaReconst.cu:216
    const float s = (float)(-std::sin(th) * 16);                       // :217
    *fcos = c;
    *fsin = s;
    *foffset = g.fcenter16 - g.h * (c + s) + 16 * iCenterOffset;       // :222
}

// ------------------------------------------------------------ twiddles
static std::mutex g_tw_mu;
static std::map<std::pair<int, int>, float*> g_tw;

const float* device_twiddles(int dev, int np)
{
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, np);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    const int n = 1 << np;
    std::vector<float> h((size_t)n);                                   // n/2 (re, im) pairs
    float p0 = (float)6.283185307178;                                   // cfft.cpp:403-404
    p0 *= -1;
    for (int i = 0; i < n / 2; i++) {
        const float a = p0 * ((float)i / (float)n);                     // :407
        h[2 * i] = (float)std::cos(a);
        h[2 * i + 1] = (float)std::sin(a);
    }
    float* d = nullptr;
    if (hipMalloc(&d, sizeof(float) * n) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_tw[key] = d;
    return d;
}

const float* device_twiddles_full(int dev, int np)
{
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(dev, np | 0x100);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    const int n = 1 << np;
    std::vector<float> h(2 * (size_t)n);
    for (int k = 0; k < n; k++) {
        const double a = -6.283185307179586476925 * k / n;
        h[2 * k] = (float)std::cos(a);
        h[2 * k + 1] = (float)std::sin(a);
    }
    float* d = nullptr;
    if (hipMalloc(&d, sizeof(float) * 2 * n) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_tw[key] = d;
    return d;
}

void drop_device_caches(int dev)
{
    std::lock_guard<std::mutex> lk(g_tw_mu);
    for (auto it = g_tw.begin(); it != g_tw.end();) {
        if (it->first.first == dev) it = g_tw.erase(it);   // memory already gone with the reset
        else ++it;
    }
}

// -------------------------------------------------------------- errors
static std::mutex g_err_mu;
static std::vector<int> g_err;
static thread_local std::string g_msg;

void log_code(int code)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    if (g_err.size() >= 1000) g_err.erase(g_err.begin());              // 1000 entries, cerror.cpp:35-43
    g_err.push_back(code);
}

void set_message(const std::string& m) { g_msg = m; }

bool hip_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    g_msg = buf;
    return false;
}

int current_device()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

int device_cu_count(int dev)
{
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) v = 1;
    cache[dev] = v;
    return v;
}

}  // namespace rv

extern "C" {

int RecviewGetErrors(int* codes, int maxn)
{
    std::lock_guard<std::mutex> lk(rv::g_err_mu);
    const int n = (int)rv::g_err.size();
    const int m = n < maxn ? n : maxn;
    for (int i = 0; i < m; i++) codes[i] = rv::g_err[n - m + i];
    return m;
}

void RecviewClearErrors(void)
{
    std::lock_guard<std::mutex> lk(rv::g_err_mu);
    rv::g_err.clear();
}

const char* RecviewLastMessage(void) { return rv::g_msg.c_str(); }

void RecviewSetFilter(int ndim, int ifilter, double cutoff, double order, float* out)
{
    rv::set_filter(ndim, ifilter, cutoff, order, out);
}

int RecviewGeometry(int ixdim, int iInterpolation, int* N, int* zoom)
{
    const rv::Geom g = rv::geometry(ixdim, iInterpolation, ixdim / 2.0, true);
    if (N) *N = g.N;
    if (zoom) *zoom = g.z;
    return g.n;
}

}  // extern "C"
