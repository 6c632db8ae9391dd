// rv_output.hip -- caller-side output stage on the GPU (gazoDoc.cpp:3993-4021):
// min / max over the whole square (imax starts at 0, imin at INT_MAX), then
// every pixel inside the circle of radius min(z*c/bin, N - z*c/bin) becomes
// (int)((float)(v - imin) * 65534 / (imax - imin)) and everything outside 0.
// Three launches: per-block min/max, one-block final reduction, normalise.
#include "rv_internal.h"

#include <climits>
#include <cfloat>

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
static void run(const T* img, int N, int irad2, uint16_t* out, T* scratch, hipStream_t s)
{
    const long long n = (long long)N * N;
    const int nparts = (int)std::min<long long>(1024, (n + 255) / 256);
    T* pmin = scratch;
    T* pmax = scratch + 1024;
    T* range = scratch + 2048;
    hipLaunchKernelGGL(minmax_kernel<T>, dim3(nparts), dim3(256), 0, s, img, n, pmin, pmax);
    hipLaunchKernelGGL(minmax_final_kernel<T>, dim3(1), dim3(256), 0, s, pmin, pmax, nparts, range);
    hipLaunchKernelGGL(normalise_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, img, N,
                       irad2, range, out);
}

void launch_output_stage(const void* img, int is_float, int N, int irad2, uint16_t* out, void* scratch,
                         hipStream_t s)
{
    if (is_float)
        run<float>((const float*)img, N, irad2, out, (float*)scratch, s);
    else
        run<int>((const int*)img, N, irad2, out, (int*)scratch, s);
}

}  // namespace rv
