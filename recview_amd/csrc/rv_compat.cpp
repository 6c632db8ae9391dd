// rv_compat.cpp -- the cudaReconst.h entry points (Part 1 of recview_hip.h).
//
// CudaReconstHost keeps the control flow of cudaReconstHost.cpp:629-813
// (bDeconv / kept-sinogram re-projection with iCenterOffset, frame selection,
// progress/abort callback, grow-only buffers, error codes) but replaces the
// per-projection CPU filter + two tiny kernels per projection with:
//   one H2D of all selected int16 rows -> one LDS FFT filter launch over all
//   rows -> one tiled LUT16 back-projection launch over all projections ->
//   one D2H of the int32 image.
#include "recview_hip.h"
#include "rv_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

// cudaReconst.cu:19
int blocksize = 256;

// Imported from the application when it defines it (general.cpp:2352).
extern "C" bool DBProjDlgCtrl(RECONST_INFO* ri, int iProgStep, int iSino, int* pCurrStep)
    __attribute__((weak));

namespace {

rv_dlgctrl_fn g_dlgctrl = nullptr;
int g_lut_formula = 1;   // cudaReconst.cu:237-239
int g_index_rule = 1;    // trunc then range check, cudaReconst.cu:36-38 with tex1Dfetch

bool call_dlgctrl(RECONST_INFO* ri, int iProgStep, int iSino, int* pCurrStep)
{
    if (g_dlgctrl) return g_dlgctrl(ri, iProgStep, iSino, pCurrStep);
    if (DBProjDlgCtrl) return DBProjDlgCtrl(ri, iProgStep, iSino, pCurrStep);
    return false;
}

// Grow-only device buffer held in a RECONST_INFO field pair (cudaReconstHost.cpp:44-81).
template <typename T>
bool grow(T*& ptr, unsigned int& maxb, size_t bytes)
{
    if (bytes > 0xffffffffull) return false;      // the reference sizes are unsigned int
    if (ptr && bytes > maxb) {
        if (hipFree(ptr) != hipSuccess) return false;
        ptr = nullptr;
        maxb = 0;
    }
    if (!ptr) {
        if (hipMalloc((void**)&ptr, bytes ? bytes : 4) != hipSuccess) { ptr = nullptr; return false; }
        maxb = (unsigned int)bytes;
    }
    return true;
}

struct Selected {
    int i, slot, sidx;
};

size_t align64(size_t b) { return (b + 63) / 64 * 64; }

// Grow-only pinned host staging per RECONST_INFO (rows and launch metadata of
// one call; the reference keeps pinned host rows per worker too,
// cudaReconstHost.cpp:678).  Every call ends with a stream synchronisation, so
// the next call may overwrite it.  Released by CudaReconstMemFree.
std::mutex g_stage_mu;
std::map<const RECONST_INFO*, std::pair<void*, size_t>> g_stage;

void* host_stage(const RECONST_INFO* ri, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    auto& h = g_stage[ri];
    if (h.first && h.second >= bytes) return h.first;
    if (h.first) (void)hipHostFree(h.first);
    h = {nullptr, 0};
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, 0) != hipSuccess) return nullptr;
    h = {p, bytes};
    return p;
}

// Work rows of CudaDeconv's two-pass transforms (one strip per call), grow-only
// per device and held until the process ends or the device is reset.
std::mutex g_deconv_mu;
std::map<int, std::pair<void*, size_t>> g_deconv_work;

void* deconv_work(int dev, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_deconv_mu);
    auto& w = g_deconv_work[dev];
    if (w.first && w.second >= bytes) return w.first;
    if (w.first) (void)hipFree(w.first);
    w = {nullptr, 0};
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    w = {p, bytes};
    return p;
}

void release_stage(const RECONST_INFO* ri)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    auto it = g_stage.find(ri);
    if (it == g_stage.end()) return;
    if (it->second.first) (void)hipHostFree(it->second.first);
    g_stage.erase(it);
}

// Device address of pinned (page-locked, GPU-mapped) host memory, or nullptr
// for pageable memory (whose failed query leaves no error behind).
void* pinned_device_pointer(const void* p)
{
    hipPointerAttribute_t a{};
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return (a.type == hipMemoryTypeHost && a.devicePointer) ? a.devicePointer : nullptr;
}

// Frame selection of cudaReconstHost.cpp:712-730: the projections to use, in
// order.  The progress / abort callback of :715-717 falls at every loop
// position i with (i - iStartSino) / iStepSino a multiple of 20; each such
// point is returned with the number of projections selected before it (an
// abort there keeps exactly those).  The callbacks themselves are made while
// the back-projection runs (see reconst_body).
struct CallbackPoint {
    int i, nsel_before;
};

std::vector<Selected> select_frames(RECONST_INFO* ri, bool bReport, std::vector<CallbackPoint>& points)
{
    std::vector<Selected> sel;
    for (int i = ri->iStartSino; i < ri->iLenSinogr - 1; i += ri->iStepSino) {
        const int isino = (i - ri->iStartSino) / ri->iStepSino;
        if (bReport && (isino % 20 == 0)) points.push_back({i, (int)sel.size()});
        if (!(ri->bInc[i] & CGAZODOC_BINC_SAMPLE)) continue;
        const int sidx = i * ri->iMultiplex + ri->iOffset;
        if (sidx >= ri->maxSinogrLen) break;
        if (ri->dReconFlags & (RQFLAGS_USEONLYEVENFRAMES | RQFLAGS_USEONLYODDFRAMES)) {
            if (i & 1) {
                if (ri->dReconFlags & RQFLAGS_USEONLYEVENFRAMES) continue;
            } else {
                if (ri->dReconFlags & RQFLAGS_USEONLYODDFRAMES) continue;
            }
        }
        sel.push_back({i, isino, sidx});
    }
    return sel;
}

// Shared body of CudaReconstHost (cudafft = false) and CudaReconstHostFFT (true).
void reconst_body(RECONST_INFO* ri, int idev, bool bReport, bool bEnStream, bool cudafft)
{
    if (!rv::hip_ok(hipSetDevice(idev), "hipSetDevice")) {
        rv::log_code(28800);
        return;
    }
    const rv::Geom g = rv::geometry(ri->ixdim, ri->iInterpolation, ri->center, true);
    const int W = g.W, N = g.N, n = g.n;
    if (W < 4 || g.ndimp < 3 || g.ndimp > rv::kMaxFftLog2) {
        rv::set_message("unsupported detector width / zoom (FFT length must be 8..2^20)");
        rv::log_code(28808);
        ri->iStatus = RECONST_INFO_ERROR;
        return;
    }
    const size_t mem_ifp = sizeof(int) * (size_t)N * N;
    const int igpdimy = ((ri->iLenSinogr - 1) - ri->iStartSino + ri->iStepSino - 1) / ri->iStepSino;
    const bool bDeconv = ((ri->dReconFlags & RQFLAGS_SINOGRAMKEPT) == 0) || (ri->d_ifp == nullptr) ||
                         (ri->max_d_ifp < mem_ifp);
    if (bDeconv) {
        // CudaReconstMemAlloc (cudaReconstHost.cpp:19-141): image, rows, filter.
        const size_t rows = (size_t)(igpdimy > 0 ? igpdimy : 1);
        if (!grow(ri->d_ifp, ri->max_d_ifp, mem_ifp) ||
            !grow(ri->d_strip, ri->max_d_strip, sizeof(short) * rows * W) ||
            !grow(ri->d_px, ri->max_d_px, sizeof(float) * rows * N) ||
            !grow(ri->d_filt, ri->max_d_filt, sizeof(float) * (size_t)n)) {
            rv::set_message("Out of device memory");
            ri->iStatus = RECONST_INFO_ERROR;
            rv::log_code(28801);
            return;
        }
        if (!rv::hip_ok(hipMemcpy(ri->d_filt, ri->fFilter, sizeof(float) * n, hipMemcpyHostToDevice),
                        "filter upload")) {
            ri->iStatus = RECONST_INFO_ERROR;
            rv::log_code(28801);
            return;
        }
        ri->iSinoCenter = g.ic;                                       // :677
    }
    // Streams (cudaReconstHost.cpp:685-702): failure falls back to the null stream.
    hipStream_t st = nullptr;
    if (bEnStream && !cudafft) {
        hipError_t e1 = hipSuccess;
        if (ri->stream1 == nullptr) {
            hipStream_t s = nullptr;
            e1 = hipStreamCreate(&s);
            if (e1 == hipSuccess) ri->stream1 = (rv_stream_t)s;
        }
        if (e1 != hipSuccess) {
            ri->stream1 = nullptr;
            rv::log_code(28807);
        } else {
            st = (hipStream_t)ri->stream1;
        }
    }

    std::vector<CallbackPoint> points;
    std::vector<Selected> sel = select_frames(ri, bReport, points);
    const int nsel = (int)sel.size();
    const int iProgStep = ri->iLenSinogr / 20;                     // PROGRESS_BAR_UNIT
    int iCurrStep = 0;
    bool host_direct = false;
    if (nsel == 0) {
        for (const auto& pt : points)
            if (call_dlgctrl(ri, iProgStep, pt.i, &iCurrStep)) break;
        if (!rv::hip_ok(hipMemsetAsync(ri->d_ifp, 0, mem_ifp, st), "image clear")) rv::log_code(28806);
    } else {
        // Metadata of the launch -- angle table, slice descriptor, row slots -- staged
        // with the selected int16 rows in one pinned host buffer and uploaded
        // asynchronously into the d_p scratch field (unused by this path) and d_strip.
        const size_t tab_bytes = align64(sizeof(rv::AngleEntry) * (size_t)nsel);
        const size_t meta_bytes = tab_bytes + align64(sizeof(rv::BpSlice)) + align64(sizeof(int) * (size_t)nsel);
        const size_t row_bytes = bDeconv ? align64(sizeof(short) * (size_t)nsel * W) : 0;
        unsigned char* h = (unsigned char*)host_stage(ri, meta_bytes + row_bytes);
        if (!h || !grow(ri->d_p, ri->max_d_p, meta_bytes)) {
            ri->iStatus = RECONST_INFO_ERROR;
            rv::log_code(h ? 28801 : 28803);
            return;
        }
        unsigned char* d = (unsigned char*)ri->d_p;
        rv::AngleEntry* h_tab = (rv::AngleEntry*)h;
        rv::BpSlice* h_slice = (rv::BpSlice*)(h + tab_bytes);
        int* h_slots = (int*)(h + tab_bytes + align64(sizeof(rv::BpSlice)));
        rv::AngleEntry* d_tab = (rv::AngleEntry*)d;
        rv::BpSlice* d_slice = (rv::BpSlice*)(d + tab_bytes);
        int* d_slots = (int*)(d + tab_bytes + align64(sizeof(rv::BpSlice)));

        const int icoff = g.ic - ri->iSinoCenter;                          // :775
        for (int q = 0; q < nsel; q++) {
            rv::AngleEntry& e = h_tab[q];
            std::memset(&e, 0, sizeof(e));
            rv::lut_params(ri->fdeg[sel[q].i], ri->fTiltAngle, g, icoff, &e.fcos, &e.fsin, &e.foffset);
            e.row = sel[q].slot;
            h_slots[q] = sel[q].slot;
        }
        // A pinned image (RecView allocates it with CUDA_MALLOC_HOST_INT,
        // gazoDoc.cpp:3893) is mapped into the GPU's address space: the
        // back-projection then writes each finished tile straight into it, so the
        // image download overlaps the kernel instead of following it.  Progress
        // reporting back-projects in segments that add into the device image.
        void* host_img = bReport ? nullptr : pinned_device_pointer(ri->iReconst);
        host_direct = host_img != nullptr;
        // one item per tile there: 32x32 tiles balance the single launch better
        // (2048^2 x 1800 from the C++ host: 2.71 ms per slice with 64x64 tiles,
        // 2.59 ms with 32x32 and 48 angles per chunk)
        const int tile = host_img ? 2 : 0;
        const int tw = rv::bp_tile_side(tile), th = rv::bp_tile_h(tile);
        rv::BpSlice& S = *h_slice;
        std::memset(&S, 0, sizeof(S));
        S.rows = ri->d_px;
        S.scales = nullptr;
        S.row_stride = N;
        S.row_len = N;
        S.nang = nsel;
        S.ang = d_tab;
        S.out = host_img ? host_img : (void*)ri->d_ifp;
        S.out_n = N;
        S.tiles_x = (N + tw - 1) / tw;
        S.tiles_y = (N + th - 1) / th;
        S.tile_base = 0;
        S.accumulate = 0;
        S.lut_formula = cudafft ? 1 : g_lut_formula;
        S.store = 0;
        S.out_scale = 1.f;
        bool ok = rv::hip_ok(hipMemcpyAsync(d, h, meta_bytes, hipMemcpyHostToDevice, st), "metadata upload");

        if (ok && bDeconv) {
            // Gather the selected int16 rows, one H2D, one filter launch.
            short* h_rows = (short*)(h + meta_bytes);
            for (int q = 0; q < nsel; q++)
                std::memcpy(h_rows + (size_t)q * W, ri->iSinogr[sel[q].sidx], sizeof(short) * W);
            const float* twd = rv::device_twiddles(idev, g.ndimp);
            ok = twd && rv::hip_ok(hipMemcpyAsync(ri->d_strip, h_rows, sizeof(short) * (size_t)nsel * W,
                                                  hipMemcpyHostToDevice, st), "row upload");
            if (ok) {
                rv::FilterJob J{};
                J.sino = ri->d_strip;
                J.in_stride = W;
                J.row_in = nullptr;
                J.row_out = d_slots;
                J.nrows = nsel;
                J.in_kind = 0;
                J.W = W; J.z = g.z; J.ic = g.ic; J.n = n; J.np = g.ndimp; J.N = N;
                J.place = cudafft ? n / 2 : n / 2 - 1;
                J.pad_mode = cudafft ? 1 : 0;
                J.crop = cudafft ? n / 2 - g.h : g.ihoffset;
                J.filt = ri->d_filt;
                J.tw = twd;
                J.out = ri->d_px;
                J.out_stride = N;
                J.store = 0;
                if (g.ndimp <= rv::kMaxLdsFftLog2) {
                    ok = rv::hip_ok(rv::launch_filter(J, st), "filter launch");
                } else {
                    // two-pass transform: its work rows live in the d_igp field, which
                    // this path does not otherwise use (grow-only like the others)
                    const size_t wrows = rv::filter_large_rows(n, (size_t)1 << 30, nsel);
                    ok = grow(ri->d_igp, ri->max_d_igp, 2 * sizeof(float) * 2 * (size_t)n * wrows) &&
                         rv::hip_ok(rv::launch_filter_large(J, ri->d_igp, wrows, st), "filter launch");
                }
            }
            if (!ok) {
                (void)hipStreamSynchronize(st);      // the staging buffer is reused by the next call
                rv::log_code(28804);
                ri->iStatus = RECONST_INFO_ERROR;
                return;
            }
        }

        // Back-projection of the selected projections: one launch, or with progress
        // reporting one launch per segment between callback points.
        int used = nsel;
        if (ok) {
            rv::BpLaunch L{};
            L.slices = d_slice;
            L.nslices = 1;
            L.total_items = S.tiles_x * S.tiles_y;
            L.item_order = nullptr;
            L.counter = nullptr;
            L.mode = rv::BP_LUT_I32;
            L.rule = cudafft ? 1 : g_index_rule;
            L.tile = tile;
            L.chunk = host_img ? 48 : rv::bp_default_chunk(rv::BP_LUT_I32, tile);
            // Each tile's projections are split over four work items whose integer
            // sums meet in the zeroed image (one 2048^2 slice is 1024 tiles, one
            // round of resident workgroups that finish unevenly; 4096 smaller
            // items let the dispatcher balance them).  An image in host memory is
            // written once per pixel instead (no atomics across the host link).
            L.angle_parts = host_img ? 1 : 4;
            L.total_items *= L.angle_parts;
            if (!host_img) rv::launch_zero_u32((unsigned int*)ri->d_ifp, N * N, st);
            if (!bReport) {
                rv::launch_bp(L, L.total_items, st);
            } else {
                // DBProjDlgCtrl at every 20th loop position (cudaReconstHost.cpp:715-717)
                // while the GPU works: segments of at least max(40, 1/8 of the
                // projections) end at callback points, the GPU at most one segment
                // ahead of the callbacks (the reference synchronises every 40
                // projections, :796).  Returning true keeps the projections before
                // that point, as the reference's loop break does.
                const int seg_min = std::max(40, (nsel + 7) / 8);
                hipEvent_t ev[2] = {nullptr, nullptr};
                const bool have_ev =
                    hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
                    hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
                int launched = 0, nseg = 0;
                auto launch_segment = [&](int a1) {
                    L.ang_begin = launched;
                    L.ang_end = a1;
                    rv::launch_bp(L, L.total_items, st);
                    launched = a1;
                };
                for (const auto& pt : points) {
                    if (pt.nsel_before - launched >= seg_min) {
                        launch_segment(pt.nsel_before);
                        if (have_ev) {
                            (void)hipEventRecord(ev[nseg & 1], st);
                            if (nseg > 0) (void)hipEventSynchronize(ev[(nseg - 1) & 1]);
                        } else {
                            (void)hipStreamSynchronize(st);
                        }
                        nseg++;
                    }
                    if (call_dlgctrl(ri, iProgStep, pt.i, &iCurrStep)) {
                        used = pt.nsel_before;
                        break;
                    }
                }
                if (used > launched) launch_segment(used);
                (void)hipStreamSynchronize(st);
                if (ev[0]) (void)hipEventDestroy(ev[0]);
                if (ev[1]) (void)hipEventDestroy(ev[1]);
            }
            ok = rv::hip_ok(hipGetLastError(), "back-projection launch");
        }
        if (ri->nSinogr) *(ri->nSinogr) += used;
        if (!ok) {
            (void)hipStreamSynchronize(st);
            rv::log_code(28806);
            ri->iStatus = RECONST_INFO_ERROR;
            return;
        }
    }
    // D2H of the image (cudaReconstHost.cpp:810-811); the copy waits for the stream.
    // A pinned image has already been written by the back-projection.
    if (!rv::hip_ok(hipStreamSynchronize(st), "reconstruction") ||
        (!host_direct &&
         !rv::hip_ok(hipMemcpy(ri->iReconst, ri->d_ifp, mem_ifp, hipMemcpyDeviceToHost), "image download")))
        rv::log_code(28805);
}

}  // namespace

extern "C" {

void RecviewSetDlgCtrl(rv_dlgctrl_fn fn) { g_dlgctrl = fn; }

void RecviewSetCompatOptions(int lut_formula, int index_rule)
{
    g_lut_formula = lut_formula ? 1 : 0;
    g_index_rule = index_rule ? 1 : 0;
}

void CudaReconstHost(RECONST_INFO* ri, int idev, bool bReport, bool bEnStream)
{
    reconst_body(ri, idev, bReport, bEnStream, false);
}

void CudaReconstHostFFT(RECONST_INFO* ri, int idev, bool bReport)
{
    reconst_body(ri, idev, bReport, false, true);
}

void CudaReconstMemFree(RECONST_INFO* ri)
{
    (void)hipSetDevice(ri->iStartSino);                            // cudaReconstHost.cpp:143
    if (ri->d_ifp && hipFree(ri->d_ifp) == hipSuccess) { ri->d_ifp = nullptr; ri->max_d_ifp = 0; }
    if (ri->d_strip && hipFree(ri->d_strip) == hipSuccess) { ri->d_strip = nullptr; ri->max_d_strip = 0; }
    if (ri->d_px && hipFree(ri->d_px) == hipSuccess) { ri->d_px = nullptr; ri->max_d_px = 0; }
    if (ri->d_igp && hipFree(ri->d_igp) == hipSuccess) { ri->d_igp = nullptr; ri->max_d_igp = 0; }
    if (ri->d_p && hipFree(ri->d_p) == hipSuccess) { ri->d_p = nullptr; ri->max_d_p = 0; }
    if (ri->d_filt && hipFree(ri->d_filt) == hipSuccess) { ri->d_filt = nullptr; ri->max_d_filt = 0; }
    release_stage(ri);
    ri->fftplan = nullptr;
    ri->ifftdim = 0;
}

void CudaReconstResourceFree(RECONST_INFO* ri, bool bCudaDeviceReset)
{
    (void)hipSetDevice(ri->iStartSino);                            // cudaReconstHost.cpp:206-213
    if (ri->stream1) { (void)hipStreamDestroy((hipStream_t)ri->stream1); ri->stream1 = nullptr; }
    if (ri->stream2) { (void)hipStreamDestroy((hipStream_t)ri->stream2); ri->stream2 = nullptr; }
    CudaReconstMemFree(ri);
    if (bCudaDeviceReset) {
        (void)hipDeviceSynchronize();
        (void)hipDeviceReset();
        rv::drop_device_caches(ri->iStartSino);
        std::lock_guard<std::mutex> lk(g_deconv_mu);
        g_deconv_work.erase(ri->iStartSino);                       // memory gone with the reset
    }
}

rv_error_t CUDA_MALLOC_HOST_INT(int** ptr, size_t size)
{
    return (rv_error_t)hipHostMalloc((void**)ptr, size, 0);
}

rv_error_t CUDA_FREE_HOST(void* ptr) { return (rv_error_t)hipHostFree(ptr); }

rv_error_t RecviewDeviceAlloc(void** ptr, size_t bytes) { return (rv_error_t)hipMalloc(ptr, bytes); }
rv_error_t RecviewDeviceFree(void* ptr) { return (rv_error_t)hipFree(ptr); }
rv_error_t RecviewCopyToHost(void* dst, const void* src, size_t bytes)
{
    return (rv_error_t)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}
rv_error_t RecviewCopyToDevice(void* dst, const void* src, size_t bytes)
{
    return (rv_error_t)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

}  // extern "C"

namespace rv {

bool sinogram_jobs(const SinoWalk& w, std::vector<SinoJob>& jobs)
{
    const char* binc = w.binc;
    const int isino = w.nframes;
    int i0 = -1, i1 = -1;
    for (int i = 0; i < isino - 1; i++) {
        if (binc[i] & CGAZODOC_BINC_NOUSE) continue;
        if (!(binc[i] & CGAZODOC_BINC_SAMPLE)) { i0 = i; break; }
    }
    for (int i = 0; i < isino - 1; i++) {
        if ((binc[i] & CGAZODOC_BINC_NOUSE) && !(binc[i] & CGAZODOC_BINC_SAMPLE)) continue;
        if (!(binc[i] & CGAZODOC_BINC_SAMPLE)) { i0 = i; i1 = -1; continue; }
        if (i1 < 0) {
            for (int j = i; j < isino - 1; j++) {
                if (binc[j] & CGAZODOC_BINC_SAMPLE) continue;
                i1 = j;
                break;
            }
            if (i1 < 0) {
                if (i0 >= 0) i1 = i0;
                else return false;
            }
        }
        if (i0 < 0 || i1 < 0) return false;
        if (i % w.step != w.start) continue;
        float t0 = w.fexp[i0];
        const float t1 = w.fexp[i1], ti = w.fexp[i];
        if (t0 == t1) t0 = 0;
        else t0 = (ti - t0) / (t1 - t0);
        const double th = theta_rad(w.fdeg[i], 0.f), cth = std::cos(th), sth = std::sin(th);
        int idrift = 0, action = 1;
        if (w.drEnd > w.drStart && (w.flags & RQFLAGS_DRIFTPARAMS)) {
            if (w.drStart < i && i < w.drEnd && w.drOmit) action = 2;
            else if (i >= w.drEnd) idrift = (int)(w.drX * cth - w.drY * sth);
            else if (i > w.drStart)
                idrift = (int)((w.drX * cth - w.drY * sth) * (i - w.drStart) / (w.drEnd - w.drStart));
        }
        if (action == 1 && (binc[i] & CGAZODOC_BINC_SKIP)) action = 2;
        if (action == 1 && (w.flags & RQFLAGS_DRIFTLIST) && w.piDrift) idrift -= w.piDrift[i];
        for (int k = 0; k < w.mult; k++) {
            SinoJob J{};
            J.frame = i; J.k = k; J.i0 = i0; J.i1 = i1; J.t0 = t0;
            J.shift = action == 1 ? idrift / w.binning : 0;
            J.action = action;
            jobs.push_back(J);
        }
    }
    return true;
}

}  // namespace rv

extern "C" {

int RecviewSinogramDevice(const unsigned short* d_frames, int nframes, int mult, int W, const char* binc,
                          const float* fexp, const float* fdeg, unsigned flags, int drStart, int drEnd,
                          double drX, double drY, int drOmit, const int* piDrift, int is8bit, int start,
                          int step, int binning, int out_float, void* d_out, rv_stream_t stream)
{
    if (!d_frames || !d_out || !binc || !fexp || !fdeg || nframes < 2 || mult <= 0 || W <= 0 || step <= 0 ||
        binning <= 0) {
        rv::set_message("RecviewSinogramDevice: bad arguments");
        return -1;
    }
    rv::SinoWalk walk{};
    walk.nframes = nframes; walk.mult = mult; walk.binc = binc; walk.fexp = fexp; walk.fdeg = fdeg;
    walk.flags = flags; walk.drStart = drStart; walk.drEnd = drEnd; walk.drOmit = drOmit;
    walk.drX = drX; walk.drY = drY; walk.piDrift = piDrift;
    walk.start = start; walk.step = step; walk.binning = binning;
    std::vector<rv::SinoJob> jobs;
    const bool good_walk = rv::sinogram_jobs(walk, jobs);
    if (!jobs.empty()) {
        hipStream_t s = (hipStream_t)stream;
        rv::SinoJob* d_jobs = nullptr;
        bool good = rv::hip_ok(hipMalloc(&d_jobs, jobs.size() * sizeof(rv::SinoJob)), "hipMalloc") &&
                    rv::hip_ok(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(rv::SinoJob),
                                              hipMemcpyHostToDevice, s), "H2D");
        if (good) {
            rv::launch_sinogram(d_frames, nframes, mult, W, d_jobs, (int)jobs.size(), d_out, is8bit ? 3 : 20,
                                (flags & RQFLAGS_ZERNIKE) ? 1 : 0, out_float ? 1 : 0, 1, s);
            good = rv::hip_ok(hipGetLastError(), "sinogram kernel") && rv::hip_ok(hipStreamSynchronize(s), "sync");
        }
        if (d_jobs) (void)hipFree(d_jobs);
        if (!good) return -1;
    }
    if (!good_walk) {
        rv::set_message("RecviewSinogramDevice: no incident frame for a sample frame");
        return -2;
    }
    return (int)jobs.size();
}

void CudaSinogramHost(RECONST_INFO* ri, int idev, bool)
{
    // Production semantics of GenerateSinogramThread (general.cpp:2377-2532): the
    // frame walk (rv::sinogram_jobs) is sequential and runs on the host; the
    // per-sample work runs in one kernel launch over every converted row.  Rows
    // are converted in place.
    if (!ri || !ri->iSinogr || !ri->bInc || !ri->fexp || !ri->fdeg || ri->ixdim <= 0 ||
        ri->iMultiplex <= 0 || ri->maxSinogrLen < 2 || ri->iStepSino <= 0) {
        rv::set_message("CudaSinogramHost: incomplete RECONST_INFO");
        rv::log_code(28901);
        if (ri) ri->iStatus = RECONST_INFO_ERROR;
        return;
    }
    const int isino = ri->maxSinogrLen, W = ri->ixdim, M = ri->iMultiplex;
    rv::SinoWalk walk{};
    walk.nframes = isino;
    walk.mult = M;
    walk.binc = ri->bInc;
    walk.fexp = ri->fexp;
    walk.fdeg = ri->fdeg;
    walk.flags = ri->dReconFlags;
    walk.drStart = ri->drStart; walk.drEnd = ri->drEnd; walk.drOmit = ri->drOmit;
    walk.drX = ri->drX; walk.drY = ri->drY;
    walk.piDrift = ri->piDrift;
    walk.start = ri->iStartSino;
    walk.step = ri->iStepSino;
    walk.binning = ri->iInterpolation == 0 ? 4 : (ri->iInterpolation == 1 ? 2 : 1);
    std::vector<rv::SinoJob> jobs;
    const bool error = !rv::sinogram_jobs(walk, jobs);
    if (!jobs.empty()) {
        if (!rv::hip_ok(hipSetDevice(idev), "hipSetDevice")) { ri->iStatus = RECONST_INFO_ERROR; return; }
        const size_t nrow = (size_t)isino * M, frame_bytes = nrow * W * sizeof(uint16_t);
        const size_t out_bytes = jobs.size() * W * sizeof(int16_t);
        uint16_t *h_frames = nullptr, *d_frames = nullptr;
        int16_t *h_out = nullptr, *d_out = nullptr;
        rv::SinoJob* d_jobs = nullptr;
        hipStream_t st = nullptr;
        bool good = rv::hip_ok(hipHostMalloc((void**)&h_frames, frame_bytes, hipHostMallocDefault), "pinned") &&
                    rv::hip_ok(hipHostMalloc((void**)&h_out, out_bytes, hipHostMallocDefault), "pinned") &&
                    rv::hip_ok(hipMalloc(&d_frames, frame_bytes), "hipMalloc") &&
                    rv::hip_ok(hipMalloc(&d_out, out_bytes), "hipMalloc") &&
                    rv::hip_ok(hipMalloc(&d_jobs, jobs.size() * sizeof(rv::SinoJob)), "hipMalloc") &&
                    rv::hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
        if (good) {
            for (size_t r = 0; r < nrow; r++) std::memcpy(h_frames + r * W, ri->iSinogr[r], W * sizeof(uint16_t));
            const int pixmin = (ri->iFlag & READTIF8bit) ? 3 : 20;   // SINOGRAM_PIXEL_MIN8bit / 16bit
            good = rv::hip_ok(hipMemcpyAsync(d_frames, h_frames, frame_bytes, hipMemcpyHostToDevice, st), "H2D") &&
                   rv::hip_ok(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(rv::SinoJob),
                                             hipMemcpyHostToDevice, st), "H2D");
            if (good) {
                rv::launch_sinogram(d_frames, isino, M, W, d_jobs, (int)jobs.size(), d_out, pixmin,
                                    (ri->dReconFlags & RQFLAGS_ZERNIKE) ? 1 : 0, 0, 0, st);
                good = rv::hip_ok(hipGetLastError(), "sinogram kernel") &&
                       rv::hip_ok(hipMemcpyAsync(h_out, d_out, out_bytes, hipMemcpyDeviceToHost, st), "D2H") &&
                       rv::hip_ok(hipStreamSynchronize(st), "sync");
            }
            if (good)
                for (size_t q = 0; q < jobs.size(); q++)
                    std::memcpy(ri->iSinogr[(size_t)jobs[q].frame * M + jobs[q].k], h_out + q * W,
                                W * sizeof(int16_t));
        }
        if (st) (void)hipStreamDestroy(st);
        if (d_jobs) (void)hipFree(d_jobs);
        if (d_out) (void)hipFree(d_out);
        if (d_frames) (void)hipFree(d_frames);
        if (h_out) (void)hipHostFree(h_out);
        if (h_frames) (void)hipHostFree(h_frames);
        if (!good) {
            rv::log_code(28901);
            ri->iStatus = RECONST_INFO_ERROR;
            return;
        }
    }
    ri->iStatus = error ? RECONST_INFO_ERROR : RECONST_INFO_IDLE;
}

void CudaDeconv(int ixdim, int iIntpDim, int ndim, float center, float* d_filt, short* d_strip,
                int* d_igp, rv_float2* d_p, rv_fft_plan_t*)
{
    // cudaReconst.cu:158-192 for one strip: zero padding, float zoom, crop at
    // n/2 - N/2, LUT with the p2igp formula.  d_p is scratch for the row.
    int np = 0;
    while ((1 << np) < ndim) np++;
    if ((1 << np) != ndim || np < 3 || np > rv::kMaxFftLog2) {
        rv::log_code(28808);
        return;
    }
    const int dev = rv::current_device();
    const float* tw = rv::device_twiddles(dev, np);
    if (!tw) { rv::log_code(28804); return; }
    const int N = ixdim * iIntpDim;
    rv::FilterJob J{};
    J.sino = d_strip;
    J.in_stride = ixdim;
    J.nrows = 1;
    J.in_kind = 0;
    J.W = ixdim; J.z = iIntpDim; J.ic = (int)center; J.n = ndim; J.np = np; J.N = N;
    J.place = ndim / 2;
    J.pad_mode = 1;
    J.crop = ndim / 2 - N / 2;
    J.filt = d_filt;
    J.tw = tw;
    J.out = (float*)d_p;
    J.out_stride = N;
    hipError_t e;
    if (np <= rv::kMaxLdsFftLog2) {
        e = rv::launch_filter(J, nullptr);
    } else {
        void* work = deconv_work(dev, 2 * sizeof(float) * 2 * (size_t)ndim);
        e = work ? rv::launch_filter_large(J, work, 1, nullptr) : hipErrorOutOfMemory;
    }
    if (!rv::hip_ok(e, "filter launch")) {
        rv::log_code(28804);
        return;
    }
    rv::launch_px2igp((const float*)d_p, d_igp, N, 1, 100.f, blocksize, nullptr);
}

void CudaBackProjStream(int ixdimp, float center, int iCenterOffset, int iIntpDim, double theta,
                        int* d_ifp, int* d_igp, rv_stream_t stream)
{
    // cudaReconst.cu:199-222
    const int ixdimh = ixdimp / 2;
    const float fcos = (float)(std::cos(theta) * 16);
    const float fsin = (float)(-std::sin(theta) * 16);
    const float fcenter = (iIntpDim == 1) ? (ixdimh + center - (int)(center)) * 16
                                          : (ixdimh + (center - (int)(center)) * iIntpDim) * 16;
    const float foffset = fcenter - ixdimh * (fcos + fsin) + 16 * iCenterOffset;
    rv::launch_bp_lut_global(d_ifp, d_igp, ixdimp, fcos, fsin, foffset, g_index_rule, blocksize,
                             (hipStream_t)stream);
}

void CudaSinoPx2igpStream(int ixdimp, int* d_igp, float* d_px, rv_stream_t stream)
{
    rv::launch_px2igp(d_px, d_igp, ixdimp, 1, 100.f, blocksize, (hipStream_t)stream);
}

// ------------------------------------------------------------ device queries
// cudaReconst.cu:449-538 with hipGetDeviceProperties.

int GetCudaDeviceCount(int iMinComputeCapability)
{
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e == hipErrorNoDevice) return 0;
    if (e == hipErrorInsufficientDriver) return CUDA_ERROR_INSUFFICIENT_DRIVER;
    if (e != hipSuccess) return CUDA_ERROR_DEVICE_GETCOUNT;
    for (int i = 0; i < count; i++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) != hipSuccess) return i | CUDA_ERROR_DEVICE_GETPROPERTY;
        if (p.major * 100 + p.minor * 10 < iMinComputeCapability)
            return i | CUDA_ERROR_INSUFFICIENT_COMPUTE_CAPABILITY;
    }
    return count;
}

int GetCudaDeviceName(int iDevice, char* pcName, int iszcName)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    const int isz = (iszcName < 256) ? iszcName : 256;
    for (int i = 0; i < isz; i++) {
        pcName[i] = p.name[i];
        if (p.name[i] == 0) break;
    }
    return 0;
}

int GetCudaDeviceComputingCapability(int iDevice, int* piMajor, int* piMinor)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    if (piMajor) *piMajor = p.major;
    if (piMinor) *piMinor = p.minor;
    return 0;
}

int GetCudaMaxThreadsPerBlock(int iDevice)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    return p.maxThreadsPerBlock;
}

int GetCudaWarpSize(int iDevice)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    return p.warpSize;
}

int GetCudaNumberOfCores(int iDevice, int* piCores, int* piProcessors)
{
    // CDNA compute unit: 4 SIMD x 16 lanes = 64 FP32 lanes per CU.
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    if (piCores) *piCores = 64;
    if (piProcessors) *piProcessors = p.multiProcessorCount;
    return 0;
}

int GetCudaClockRate(int iDevice, int* piClockRate, int* piMemRate)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, iDevice) != hipSuccess) return CUDA_ERROR_DEVICE_GETPROPERTY;
    if (piClockRate) *piClockRate = p.clockRate;
    if (piMemRate) *piMemRate = p.memoryClockRate;
    return 0;
}

}  // extern "C"
