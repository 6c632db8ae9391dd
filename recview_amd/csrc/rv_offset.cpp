// rv_offset.cpp -- offset-CT (360-degree, off-centre axis) entry points.
#include "recview_hip.h"
#include "rv_common.h"

extern "C" {

// Frame pairing of gazoDoc.cpp:2967-2978 and 3562-3572: output row i (frames
// below 180 deg, i < iRevIdx0) pairs with the first later sampled frame at
// >= 180 deg further; unsampled rows get -1.  Returns the number of output
// rows, or -21051 (no frame beyond 180 deg) / -21052 (partner search ran out).
int RecviewOffsetCTPairs(const float* fdeg, const char* binc, int nframes, int* pair)
{
    int iRevIdx = -1;
    for (int i = 0; i < nframes; i++) {
        if (fdeg[i] > 180.) { iRevIdx = i; break; }
    }
    if (iRevIdx < 0) return -21051;
    const int nout = iRevIdx;
    for (int i = 0; i < nout; i++) {
        pair[i] = -1;
        if (!(binc[i] & CGAZODOC_BINC_SAMPLE)) continue;
        const double th = fdeg[i];
        while ((fdeg[iRevIdx] - th < 180) | ((binc[iRevIdx] & CGAZODOC_BINC_SAMPLE) == 0)) {
            iRevIdx++;
            if (iRevIdx >= nframes) return -21052;
        }
        pair[i] = iRevIdx;
    }
    return nout;
}

int RecviewOffsetCTStitch(const void* d_sino, int is_float, int W, const int* d_pair, int nout,
                          double center, double deltac, int bin, int trim, int out_width,
                          void* d_out, rv_stream_t stream)
{
    if (W <= 0 || bin <= 0 || nout < 0 || out_width <= 0) {
        rv::set_message("invalid stitch geometry");
        return -1;
    }
    rv::StitchJob J{};
    J.sino = d_sino;
    J.W = W;
    J.bin = bin;
    J.trim = trim;
    J.center = center;
    J.deltac = deltac;
    J.pair = d_pair;
    J.nout = nout;
    J.out_width = out_width;
    J.out = d_out;
    J.is_float = is_float ? 1 : 0;
    rv::launch_stitch(J, (hipStream_t)stream);
    return rv::hip_ok(hipGetLastError(), "stitch launch") ? 0 : -1;
}

}  // extern "C"
