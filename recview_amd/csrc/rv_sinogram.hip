// rv_sinogram.hip -- sinogram generation on the GPU: dark / incident (flat)
// correction with the incident interpolated in exposure time, -ln(I/I0) x
// LOG_SCALE (or the Zernike ratio), the pixel-count floor and the drift shift,
// with the production semantics of GenerateSinogramThread (general.cpp:2377-2532).
// One block row per converted (frame, multiplex) row; the host resolves the
// incident frames, weights and shifts (the sequential part) into a job table.
#include "rv_internal.h"

namespace rv {

// out_float = 0: int16 rows as the reference; 1: float rows, -ln(I/I0) (or the
// Zernike ratio) without the x3000 scale and rounding.  out_full = 1 writes row
// frame * mult + k (the frames' layout), 0 writes row j of the job list.
__global__ void sinogram_kernel(const uint16_t* frames, int nframes, int mult, int W,
                                const SinoJob* jobs, void* out, int pixmin, int zernike, int out_float,
                                int out_full)
{
    const SinoJob J = jobs[blockIdx.y];
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= W) return;
    const size_t orow = out_full ? (size_t)J.frame * mult + J.k : (size_t)blockIdx.y;
    int16_t v = 0;
    float fv = 0.f;
    const int src = j - J.shift;   // general.cpp:2501-2513: out[j] = row[j - shift], zero fill
    if (J.action == 1 && src >= 0 && src < W) {
        const uint16_t* dark = frames + ((size_t)(nframes - 1) * mult + J.k) * W;
        const uint16_t* a0 = frames + ((size_t)J.i0 * mult + J.k) * W;
        const uint16_t* a1 = frames + ((size_t)J.i1 * mult + J.k) * W;
        const uint16_t* row = frames + ((size_t)J.frame * mult + J.k) * W;
        // unsigned short + float arithmetic in float, kept as double (general.cpp:2483)
        const float fa = (float)((int)a1[src] - (int)a0[src]) * J.t0;
        const float fb = (float)a0[src] + fa;
        const double inc = (double)(fb - (float)dark[src]);
        if (inc > 0) {
            const int is = (int)row[src] - (int)dark[src];
            if (is >= pixmin) {
                if (out_float)
                    fv = zernike ? (float)(inc / is) : (float)log(inc / is);
                else
                    v = zernike ? (int16_t)(int)(inc * 1000 / is + 0.5)
                                : (int16_t)(int)(log(inc / is) * 3000 + 0.5);
            }
        }
    }
    if (out_float)
        ((float*)out)[orow * W + j] = fv;
    else
        ((int16_t*)out)[orow * W + j] = v;
}

void launch_sinogram(const uint16_t* frames, int nframes, int mult, int W, const SinoJob* jobs,
                     int njobs, void* out, int pixmin, int zernike, int out_float, int out_full,
                     hipStream_t stream)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(sinogram_kernel, dim3((W + 255) / 256, njobs), dim3(256), 0, stream, frames,
                       nframes, mult, W, jobs, out, pixmin, zernike, out_float, out_full);
}

}  // namespace rv
