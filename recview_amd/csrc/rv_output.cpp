// rv_output.cpp -- the output side of a reconstruction: the caller-side
// normalisation on the GPU (gazoDoc.cpp:3993-4021) and the 16-bit TIFF writer
// with RecView's ImageDescription metadata (general.cpp:1313-1470,
// gazoDoc.cpp:1986-1989), byte for byte the reference's file layout.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "recview_hip.h"
#include "rv_common.h"

namespace {

// TIFF tags and field types, constants.h:13-35.
enum : uint16_t {
    kTifAscii = 2, kTifShort = 3, kTifLong = 4,
    kImageWidth = 0x0100, kImageLength = 0x0101, kBitsPerSample = 0x0102, kCompression = 0x0103,
    kPhotometric = 0x0106, kImageDescription = 0x010e, kStripOffsets = 0x0111,
    kSamplesPerPixel = 0x0115, kRowsPerStrip = 0x0116, kStripByteCounts = 0x0117,
    kPlanarConfiguration = 0x011c, kArtist = 0x013b
};
constexpr int kCargSize = 200;   // READTIF_CARG_SIZE

void put16(std::vector<unsigned char>& b, uint16_t v) { b.push_back(v & 0xff); b.push_back(v >> 8); }
void put32(std::vector<unsigned char>& b, uint32_t v)
{
    for (int i = 0; i < 4; i++) b.push_back((v >> (8 * i)) & 0xff);
}
void entry(std::vector<unsigned char>& b, uint16_t tag, uint16_t type, uint32_t count, uint32_t value)
{
    put16(b, tag);
    put16(b, type);
    put32(b, count);
    put32(b, value);
}

// The reference pads strings shorter than 4 characters with 4 blanks and cuts
// them at READTIF_CARG_SIZE-1; the stored count includes the terminating NUL.
std::string tif_text(const char* s, const char* dflt)
{
    std::string t = s ? s : dflt;
    if (t.size() < 4) t += "    ";
    if (t.size() > (size_t)kCargSize - 1) t.resize(kCargSize - 1);
    return t;
}

}  // namespace

extern "C" {

int RecviewOutputStage(const void* d_img, int is_float, int N, double center, int iInterpolation,
                       unsigned short* d_out, double* vmin, double* vmax, rv_stream_t stream)
{
    if (!d_img || !d_out || N <= 0 || iInterpolation < 0 || iInterpolation > 5) {
        rv::set_message("RecviewOutputStage: bad arguments");
        return -1;
    }
    const int irad2 = rv::output_irad2(N, center, iInterpolation);     // gazoDoc.cpp:4008-4010
    hipStream_t s = (hipStream_t)stream;
    void* scratch = nullptr;
    if (!rv::hip_ok(hipMalloc(&scratch, 2050 * 4), "hipMalloc")) return -1;
    rv::launch_output_stage(d_img, is_float, N, irad2, d_out, scratch, s);
    bool good = rv::hip_ok(hipGetLastError(), "output stage");
    unsigned char range[8];
    good = good && rv::hip_ok(hipMemcpyAsync(range, (char*)scratch + 2048 * 4, 8, hipMemcpyDeviceToHost, s), "D2H") &&
           rv::hip_ok(hipStreamSynchronize(s), "sync");
    (void)hipFree(scratch);
    if (!good) return -1;
    if (is_float) {
        float f[2];
        std::memcpy(f, range, 8);
        if (vmin) *vmin = f[0];
        if (vmax) *vmax = f[1];
    } else {
        int v[2];
        std::memcpy(v, range, 8);
        if (vmin) *vmin = v[0];
        if (vmax) *vmax = v[1];
    }
    return 0;
}

int RecviewImageDescription(char* buf, int bufsize, int numSinogr, double pixelWidth, double pixelDiv,
                            double pixelBase, double center, int ifilter)
{
    // gazoDoc.cpp:1986-1988
    if (!buf || bufsize <= 0) return -1;
    return std::snprintf(buf, (size_t)bufsize, "%d\t%.6f\t%.6f\t%.6f\t%.6f\t%d\t%.6f\t%.6f ", numSinogr,
                         pixelWidth, pixelDiv, pixelBase, center, ifilter + 1, 55.111697, 0.461227);
}

int RecviewWriteTif16(const char* path, const int* buffer, int height, int width, const char* desc,
                      const char* artist)
{
    // general.cpp:1313-1470: little-endian header, rows of uint16 (negative
    // values written as 0), 12 directory entries, ImageDescription, one strip
    // per row (offsets and byte counts), Artist.
    if (!buffer || height <= 0 || width <= 0) return 18100;
    const std::string d = tif_text(desc, "tif image"), a = tif_text(artist, "SP-uCT, gazoView");
    const uint32_t nDesc = (uint32_t)d.size() + 1, nArt = (uint32_t)a.size() + 1;
    const uint32_t offset = 8;
    const uint32_t oDirEnt = offset + (uint32_t)height * width * 2;
    const uint16_t nEnt = 12;
    const uint32_t oDesc = oDirEnt + 2 + nEnt * 12 + 4;
    uint32_t oStrips = oDesc + nDesc;
    if (nDesc & 1) oStrips++;
    const uint32_t oCounts = oStrips + (uint32_t)height * 4;
    const uint32_t oArtist = oCounts + (uint32_t)height * 4;
    std::vector<unsigned char> b;
    b.reserve((size_t)oArtist + nArt + 2);
    b.push_back('I'); b.push_back('I');
    put16(b, 42);
    put32(b, oDirEnt);
    for (long long i = 0; i < (long long)height * width; i++)
        put16(b, buffer[i] > 0 ? (uint16_t)buffer[i] : 0);
    put16(b, nEnt);
    entry(b, kImageWidth, kTifShort, 1, (uint32_t)width);
    entry(b, kImageLength, kTifShort, 1, (uint32_t)height);
    entry(b, kBitsPerSample, kTifShort, 1, 16);
    entry(b, kCompression, kTifShort, 1, 1);
    entry(b, kPhotometric, kTifShort, 1, 1);
    entry(b, kImageDescription, kTifAscii, nDesc, oDesc);
    entry(b, kStripOffsets, kTifLong, (uint32_t)height, oStrips);
    entry(b, kSamplesPerPixel, kTifShort, 1, 1);
    entry(b, kRowsPerStrip, kTifShort, 1, 1);
    entry(b, kStripByteCounts, kTifLong, (uint32_t)height, oCounts);
    entry(b, kPlanarConfiguration, kTifShort, 1, 1);
    entry(b, kArtist, kTifAscii, nArt, oArtist);
    put32(b, 0);
    b.insert(b.end(), d.begin(), d.end());
    b.push_back(0);
    if (nDesc & 1) b.push_back(0);
    for (int i = 0; i < height; i++) put32(b, offset + (uint32_t)i * width * 2);
    for (int i = 0; i < height; i++) put32(b, (uint32_t)width * 2);
    b.insert(b.end(), a.begin(), a.end());
    b.push_back(0);
    if (nArt & 1) b.push_back(0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return 18100;
    const size_t w = std::fwrite(b.data(), 1, b.size(), f);
    const int c = std::fclose(f);
    return (w == b.size() && c == 0) ? 0 : 18102;
}

}  // extern "C"
