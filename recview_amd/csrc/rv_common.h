// rv_common.h -- host-side helpers shared by the compatibility entry points
// (rv_compat.cpp) and the native plan API (rv_plan.cpp).
#pragma once
#include "rv_internal.h"
#include <string>

namespace rv {

// Geometry of cudaReconstHost.cpp:635-658 (float_center) or
// general.cpp:2021-2035 (double center).
struct Geom {
    int W, izoom, z, ndimp, n, N, h, ic;
    float fcenter16;   // fcenter in 1/16-sample units
    int ihoffset;      // n/2 - 1 - h
};
Geom geometry(int ixdim, int iInterpolation, double center, bool float_center);

// gazoDoc.cpp:3747-3812 (+ 5 = Shepp-Logan).
void set_filter(int ndim, int ifilter, double cutoff, double order, float* out);

// Per-projection LUT16 parameters, cudaReconst.cu:216-222.
void lut_params(float fdeg, float tilt, const Geom& g, int iCenterOffset, float* fcos,
                float* fsin, float* foffset);
double theta_rad(float fdeg, float tilt);

// CFft twiddles (cfft.cpp:403-409) for 2^np points, cached per device.
const float* device_twiddles(int dev, int np);
// exp(-2 pi i k / n) for k < n (interleaved re, im; computed in double) for the
// standard-convention transforms of the paired filter, cached per device.
const float* device_twiddles_full(int dev, int np);
void drop_device_caches(int dev);

// Error reporting: the ring of reference error codes and a thread-local message.
void log_code(int code);
void set_message(const std::string& m);
bool hip_ok(hipError_t e, const char* what);
int current_device();

// Persistent-grid sizing.
int bp_blocks_per_cu(int mode, int tile, int chunk);
int device_cu_count(int dev);
void set_kernel_lds_limits();

}  // namespace rv
