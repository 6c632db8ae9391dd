// rv_stitch.hip -- offset-CT stitch as a GPU pre-pass.
//
// A 360-degree scan with the rotation axis off the detector centre is turned
// into a 180-degree sinogram of double width: each frame i below 180 deg is
// merged with its partner at >= 180 deg further, mirrored about the axis with
// sub-pixel interpolation and a linear cross-fade band (gazoDoc.cpp:3536-3665).
// The reference walks each output row sequentially and carries the last valid
// value (`lastDens`) across positions whose mirror index leaves the detector;
// here every output sample is one thread, and the carry is resolved by
// evaluating the last valid position in the reference's loop order.
// Int16 rows reproduce the reference's (short) truncations; float rows use the
// same arithmetic without them.
#include "rv_internal.h"

namespace rv {

template <typename T>
__device__ __forceinline__ T to_out(double v)
{
    if constexpr (sizeof(T) == 2) return (T)(int)v;     // (short)(double): truncation
    else return (T)v;
}

template <typename T>
__device__ T stitch_value(const StitchJob& J, const T* s1, const T* s2, int o)
{
    const int W = J.W, bin = J.bin, trim = J.trim;
    const int ixlen = W / bin;
    if (J.center > W / 2.) {
        // axis in the right half (gazoDoc.cpp:3575-3616)
        const double fc = (J.center + J.deltac + trim) / bin;
        const int iCent = (int)fc;
        const double dCent = fc - iCent;
        const int g = W / bin - 1 - iCent;
        const int k = o + trim / bin;
        if (k <= iCent - g) return s1[k];
        const int kend = (2 * W - trim) / bin;           // exclusive
        if (k >= kend) return (T)0;
        const bool lowhalf = dCent < 0.5;
        const int lo = lowhalf ? 0 : 1, hi = lowhalf ? ixlen - 2 : ixlen - 1;
        const int kstart = iCent - g + 1;
        const int kv0 = (2 * iCent - hi > kstart) ? 2 * iCent - hi : kstart;   // first valid k
        const int kv1 = 2 * iCent - lo;                                         // last valid k
        if (k < kv0 || kv1 < kv0) return (T)0;
        const int ke = (k < kv1) ? k : kv1;
        const int idx = 2 * iCent - ke;
        double v;
        if (lowhalf) {
            const T ir0 = s2[idx + 1], ir1 = s2[idx];
            v = (double)to_out<T>((double)ir0 + ((double)ir1 - (double)ir0) * (1 - 2 * dCent));
        } else {
            const T ir0 = s2[idx], ir1 = s2[idx - 1];
            v = (double)to_out<T>((double)ir1 + ((double)ir0 - (double)ir1) * (2 - 2 * dCent));
        }
        if (ke <= iCent + g) {
            const double ratio = (double)(ke - (iCent - g)) / (g * 2 + 1);
            v = (double)to_out<T>(ratio * v + (1. - ratio) * (double)s1[ke]);
        }
        return (T)v;
    } else {
        // axis in the left half (gazoDoc.cpp:3617-3660)
        const double fc = (J.center + J.deltac) / bin;
        const int iCent = (int)(fc + 0.5);
        const double dCent = fc - (int)fc;
        const int g = iCent;
        const int shift = (W - trim) / bin;
        const int k = o - shift;
        if (k >= iCent + g) return (k <= (W - 1 - trim) / bin) ? s1[k] : (T)0;
        const bool lowhalf = dCent < 0.5;
        const int kstart = iCent + g - 1;                               // loop runs downwards
        const int kstop = lowhalf ? 2 * iCent + 1 - W / bin : 2 * iCent - W / bin;
        if (k < kstop) return (T)0;
        // idx = 2 iCent - k (+0 or -1); valid for 0 <= idx <= ixlen - 2
        const int off = lowhalf ? 0 : 1;
        const int kv_hi = 2 * iCent - off;                              // idx = 0
        const int kv_lo = 2 * iCent - off - (ixlen - 2);                // idx = ixlen - 2
        const int top = (kv_hi < kstart) ? kv_hi : kstart;             // first valid in loop order
        if (k > top) return (T)0;
        if (top < kv_lo) return (T)0;
        const int ke = (k > kv_lo) ? k : kv_lo;
        const int idx = 2 * iCent - off - ke;
        double v;
        if (lowhalf) {
            const T ir1 = s2[idx], ir2 = s2[idx + 1];
            v = (double)to_out<T>((double)ir1 + ((double)ir2 - (double)ir1) * (2 * dCent));
        } else {
            const T ir0 = s2[idx], ir1 = s2[idx + 1];
            v = (double)to_out<T>((double)ir0 + ((double)ir1 - (double)ir0) * (2 * dCent - 1));
        }
        if (ke >= iCent - g) {
            const double ratio = (double)((iCent + g) - ke) / (g * 2 + 1);
            v = (double)to_out<T>(ratio * v + (1. - ratio) * (double)s1[ke]);
        }
        return (T)v;
    }
}

template <typename T>
__global__ void stitch_kernel(StitchJob J)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (o >= J.out_width) return;
    T* out = (T*)J.out + (long long)i * J.out_width;
    const int p = J.pair[i];
    if (p < 0) { out[o] = (T)0; return; }
    const int stride = J.W / J.bin;                     // rows are already binned
    const T* s1 = (const T*)J.sino + (long long)i * stride;
    const T* s2 = (const T*)J.sino + (long long)p * stride;
    out[o] = stitch_value<T>(J, s1, s2, o);
}

void launch_stitch(const StitchJob& J, hipStream_t stream)
{
    if (J.nout <= 0) return;
    const dim3 grid((J.out_width + 255) / 256, J.nout);
    if (J.is_float) hipLaunchKernelGGL(stitch_kernel<float>, grid, dim3(256), 0, stream, J);
    else hipLaunchKernelGGL(stitch_kernel<int16_t>, grid, dim3(256), 0, stream, J);
}

}  // namespace rv
