// recview_harness.cpp -- a C++ host that uses librecview_hip.so exactly the
// way RecView's reconstruction path does, through include/recview_hip.h only
// (no HIP or CUDA header): it fills RECONST_INFO like CGazoDoc::DeconvBackProj
// (gazoDoc.cpp:3929-3951) for one device, calls CudaReconstHost like
// DeconvBackProjThread (general.cpp:2172), applies the output stage on the
// device and writes the 16-bit slice as RecView's TIFF.
//
//   recview_harness [--repeat K] <sino.i16> <P> <W> <center_user> <fdeg.f32> <out.i32> [out.tif]
//
// sino.i16: P rows of W int16 samples (x3000); fdeg.f32: P angles (degrees).
// out.i32 receives the raw int32 image (W x W).  --repeat K times K more calls
// after the first (steady state: buffers already grown) and prints the mean
// wall time per call.  Exit code 0 on success.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "recview_hip.h"

static bool read_file(const char* path, void* dst, size_t bytes)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t n = std::fread(dst, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

int main(int argc, char** argv)
{
    int repeat = 0;
    if (argc > 2 && std::strcmp(argv[1], "--repeat") == 0) {
        repeat = std::atoi(argv[2]);
        argc -= 2;
        argv += 2;
    }
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s sino.i16 P W center_user fdeg.f32 out.i32 [out.tif]\n", argv[0]);
        return 2;
    }
    const int P = std::atoi(argv[2]), W = std::atoi(argv[3]);
    const double c_user = std::atof(argv[4]);
    if (P <= 0 || W <= 0) return 2;
    std::vector<short> sino((size_t)P * W);
    std::vector<float> fdeg(P);
    if (!read_file(argv[1], sino.data(), sino.size() * sizeof(short)) ||
        !read_file(argv[5], fdeg.data(), fdeg.size() * sizeof(float))) {
        std::fprintf(stderr, "cannot read inputs\n");
        return 2;
    }
    int ndim = 0, N = 0, zoom = 0;
    ndim = RecviewGeometry(W, 2, &N, &zoom);        // no binning / zoom
    std::vector<float> filt(ndim);
    RecviewSetFilter(ndim, 0, 0.5, 6.0, filt.data());  // Hann, RecView's defaults
    std::vector<short*> rows(P);
    for (int i = 0; i < P; i++) rows[i] = sino.data() + (size_t)i * W;
    std::vector<char> binc(P + 1, CGAZODOC_BINC_SAMPLE);
    int* image = nullptr;
    if (CUDA_MALLOC_HOST_INT(&image, sizeof(int) * (size_t)N * N) != 0) return 3;
    std::memset(image, 0, sizeof(int) * (size_t)N * N);
    int nsino = 0;

    RECONST_INFO ri;
    std::memset(&ri, 0, sizeof(ri));
    ri.iStatus = RECONST_INFO_BUSY;
    ri.ixdim = W;
    ri.iInterpolation = 2;
    ri.center = c_user + 0.5;                        // HALFPIXEL_OFFSET, no binning
    ri.iLenSinogr = P + 1;
    ri.iReconst = image;
    ri.iStartSino = 0;
    ri.iStepSino = 1;
    ri.nSinogr = &nsino;
    ri.bInc = binc.data();
    ri.fdeg = fdeg.data();
    ri.iMultiplex = 1;
    ri.iOffset = 0;
    ri.maxSinogrLen = P;
    ri.iSinogr = rows.data();
    ri.fFilter = filt.data();
    CudaReconstHost(&ri, 0, false, true);
    if (ri.iStatus == RECONST_INFO_ERROR) {
        std::fprintf(stderr, "CudaReconstHost failed: %s\n", RecviewLastMessage());
        return 4;
    }
    if (repeat > 0) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < repeat; k++) {
            ri.iStatus = RECONST_INFO_BUSY;
            nsino = 0;
            CudaReconstHost(&ri, 0, false, true);
            if (ri.iStatus == RECONST_INFO_ERROR) {
                std::fprintf(stderr, "CudaReconstHost failed: %s\n", RecviewLastMessage());
                return 4;
            }
        }
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / repeat;
        std::printf("ms_per_call %.3f\n", ms);
    }
    FILE* f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(image, sizeof(int), (size_t)N * N, f) != (size_t)N * N) return 5;
    std::fclose(f);
    if (argc > 7) {
        // 16-bit slice: the output stage on the GPU (the int32 image uploaded from
        // iReconst, which the engine wrote directly since it is pinned), then
        // RecView's TIFF layout.
        std::vector<unsigned short> u16((size_t)N * N);
        unsigned short* d_out = nullptr;
        int* d_img = nullptr;
        double lo = 0, hi = 0;
        if (RecviewDeviceAlloc((void**)&d_out, sizeof(unsigned short) * (size_t)N * N) != 0 ||
            RecviewDeviceAlloc((void**)&d_img, sizeof(int) * (size_t)N * N) != 0 ||
            RecviewCopyToDevice(d_img, image, sizeof(int) * (size_t)N * N) != 0 ||
            RecviewOutputStage(d_img, 0, N, c_user, 2, d_out, &lo, &hi, nullptr) != 0 ||
            RecviewCopyToHost(u16.data(), d_out, sizeof(unsigned short) * u16.size()) != 0)
            return 6;
        RecviewDeviceFree(d_out);
        RecviewDeviceFree(d_img);
        std::vector<int> as_int(u16.begin(), u16.end());
        char desc[256];
        RecviewImageDescription(desc, sizeof(desc), nsino, 0.5, 1.0, lo, c_user, 0);
        if (RecviewWriteTif16(argv[7], as_int.data(), N, N, desc, nullptr) != 0) return 7;
    }
    CudaReconstResourceFree(&ri, false);
    CUDA_FREE_HOST(image);
    std::printf("ok %d projections, %dx%d\n", nsino, N, N);
    return 0;
}
