// rv_output.hip -- caller-side output stage on the GPU (gazoDoc.cpp:3993-4021):
// min / max over the whole square (imax starts at 0, imin at INT_MAX), then
// every pixel inside the circle of radius min(z*c/bin, N - z*c/bin) becomes
// (int)((float)(v - imin) * 65534 / (imax - imin)) and everything outside 0.
// Three launches: per-block min/max, one-block final reduction, normalise.
#include "recview_hip.h"
#include "rv_common.h"

#include <climits>
#include <cfloat>
#include <map>
#include <mutex>

namespace rv {

template <typename T>
struct Lim;
template <>
struct Lim<int> {
    static __device__ int lo() { return INT_MAX; }
};
template <>
struct Lim<float> {
    static __device__ float lo() { return FLT_MAX; }
};

template <typename T>
__global__ void minmax_kernel(const T* img, long long n, T* part_min, T* part_max)
{
    __shared__ T smin[256], smax[256];
    T mn = Lim<T>::lo(), mx = (T)0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const T v = img[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = smin[threadIdx.x + s] < smin[threadIdx.x] ? smin[threadIdx.x + s] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + s] > smax[threadIdx.x] ? smax[threadIdx.x + s] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_min[blockIdx.x] = smin[0];
        part_max[blockIdx.x] = smax[0];
    }
}

// In: nparts partial minima / maxima.  Out: range[0] = min, range[1] = max
// (max bumped by one unit when equal to the min, gazoDoc.cpp:3999).
template <typename T>
__global__ void minmax_final_kernel(const T* part_min, const T* part_max, int nparts, T* range)
{
    __shared__ T smin[256], smax[256];
    T mn = Lim<T>::lo(), mx = (T)0;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        mn = part_min[i] < mn ? part_min[i] : mn;
        mx = part_max[i] > mx ? part_max[i] : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = smin[threadIdx.x + s] < smin[threadIdx.x] ? smin[threadIdx.x + s] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + s] > smax[threadIdx.x] ? smax[threadIdx.x + s] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        range[0] = smin[0];
        range[1] = smax[0] == smin[0] ? smax[0] + (T)1 : smax[0];
    }
}

template <typename T>
__global__ void normalise_kernel(const T* img, int N, int irad2, const T* range, uint16_t* out)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)N * N) return;
    const int r = (int)(i / N), c = (int)(i % N);
    const int h = N / 2;
    const int i2 = (r - h) * (r - h), j2 = (c - h) * (c - h);
    uint16_t v = 0;
    if (i2 + j2 <= irad2) {
        const T mn = range[0], mx = range[1];
        const float d = (float)(img[i] - mn);
        v = (uint16_t)(int)(d * (float)(65535 - 1) / (float)(mx - mn));
    }
    out[i] = v;
}

template <typename T>
static void run(const T* img, int N, int irad2, uint16_t* out, T* part, T* range, hipStream_t s)
{
    const long long n = (long long)N * N;
    const int nparts = (int)std::min<long long>(kOutputParts, (n + 255) / 256);
    T* pmin = part;
    T* pmax = part + kOutputParts;
    hipLaunchKernelGGL(minmax_kernel<T>, dim3(nparts), dim3(256), 0, s, img, n, pmin, pmax);
    hipLaunchKernelGGL(minmax_final_kernel<T>, dim3(1), dim3(256), 0, s, pmin, pmax, nparts, range);
    hipLaunchKernelGGL(normalise_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, img, N,
                       irad2, range, out);
}

void launch_output_stage_parts(const void* img, int is_float, int N, int irad2, uint16_t* out, void* part,
                               void* range, hipStream_t s)
{
    if (is_float)
        run<float>((const float*)img, N, irad2, out, (float*)part, (float*)range, s);
    else
        run<int>((const int*)img, N, irad2, out, (int*)part, (int*)range, s);
}

void launch_output_stage(const void* img, int is_float, int N, int irad2, uint16_t* out, void* scratch,
                         hipStream_t s)
{
    launch_output_stage_parts(img, is_float, N, irad2, out, scratch, (char*)scratch + 2 * kOutputParts * 4, s);
}

}  // namespace rv

// ------------------------------------------------ allocation-free output stage
// The partial minima and maxima of RecviewOutputStageNoAlloc live in this
// array: 1024 + 1024 four-byte values (8 KiB) on every GPU the library runs on,
// held until the process ends.  The partials are capped at 1024 whatever the
// image side, so the array serves every image the entry point accepts.  Host
// code never passes the array's name to a kernel: it passes the device address
// hipGetSymbolAddress returns for the current device, kept per device.
static __device__ __align__(256) unsigned int scratch[2 * rv::kOutputParts];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;

// The device address of `scratch` on the current device (looked up on the first
// call on that device), or nullptr with the message set.
void* scratch_on_current_device()
{
    int dev = 0;
    if (!rv::hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* ptr = nullptr;
    if (!rv::hip_ok(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
        return nullptr;
    g_scratch_ptr[dev] = ptr;
    return ptr;
}

}  // namespace

extern "C" int RecviewOutputStageNoAlloc(const void* d_img, int is_float, int N, double center,
                                         int iInterpolation, unsigned short* d_out, void* d_range,
                                         rv_stream_t stream)
{
    if (!d_img || !d_out || !d_range || N <= 0 || iInterpolation < 0 || iInterpolation > 5) {
        rv::set_message("RecviewOutputStageNoAlloc: bad arguments");
        return -1;
    }
    if (N > RV_OUTPUT_NOALLOC_MAX_SIDE) {
        rv::set_message("RecviewOutputStageNoAlloc: image side N exceeds 65536");
        return -1;
    }
    void* part = scratch_on_current_device();
    if (!part) return -1;
    // every partial the final reduction reads is written by the first launch of
    // this call, so nothing needs clearing
    rv::launch_output_stage_parts(d_img, is_float, N, rv::output_irad2(N, center, iInterpolation), d_out,
                                  part, d_range, (hipStream_t)stream);
    return rv::hip_ok(hipGetLastError(), "output stage") ? 0 : -1;
}
