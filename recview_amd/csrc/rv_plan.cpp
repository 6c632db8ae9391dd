// rv_plan.cpp -- native batched API (Part 2 of recview_hip.h).
//
// A plan fixes everything that does not depend on the sinogram values: the
// geometry of every slice, its filter table, the FP64 angle table, the
// filtered-sinogram workspace in HBM, the offset-CT pairing and stitch buffers,
// and the cost-sorted tile work lists.  Execution is then: one stitch launch per
// offset-CT slice, one filter launch per slice (LDS FFT, all rows), and one
// back-projection launch per tile configuration (tiles of every slice of that
// configuration, in one grid or through a persistent work counter), optionally
// captured once into a hipGraph and replayed.  The reference has no
// counterpart: it reconstructs one slice at a time (gazoDoc.cpp:1965-1991).
#include "recview_hip.h"
#include "rv_common.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <map>
#include <numeric>
#include <string>
#include <vector>

namespace {

struct SliceState {
    RvSliceSpec spec;
    std::vector<float> fdeg;     // angles of the reconstructed rows
    rv::Geom g;                  // geometry of the reconstructed rows
    int out_n = 0;
    int nang = 0;
    size_t rows_off = 0;         // element offset of row 0 in the filtered workspace
    size_t ang_off = 0;          // first entry in the angle table
    size_t filt_off = 0;         // first entry in the filter tables
    int tiles_x = 0, tiles_y = 0, tile_base = 0;
    int group = 0;
    int tile = 0;
    int fclass = 0;              // first slice with the same filter job (batched launches)
    int rows_src = -1;           // slice whose filtered rows this one back-projects (-1: its own)
    size_t scale_off = 0;        // first per-row scale of the rows it reads (16-bit stores)
    // offset CT
    bool offset = false;
    int raw_W = 0;
    size_t stitch_off = 0;       // element offset in the stitch buffer
    size_t pair_off = 0;         // first entry in the pair table
};

struct Group {
    int tile = 0, chunk = 0;
    int mode = 0;                // BP mode of the group's launch
    bool tgeom = false;          // BP_LINEAR2: every entry has a tile geometry table
    std::vector<int> slices;     // first slice of each work-item entry
    std::vector<std::array<int, rv::BP_MAX_SLICES_PER_ITEM - 1>> others;   // BP_LINEAR2: other slices (-1: none)
    std::vector<rv::BpSlice> h;
    rv::BpSlice* d_slices = nullptr;
    int* d_order = nullptr;
    int total_items = 0;
    int grid = 0;
};

}  // namespace

struct rv_plan_opaque {
    int dev = 0;
    RvOptions opt{};
    int mode = rv::BP_LINEAR;
    int store = 0;
    std::vector<SliceState> sl;
    std::vector<int> filter_order;       // slices grouped by filter class (batched launches)
    std::vector<Group> groups;
    void* d_rows = nullptr;
    float* d_scales = nullptr;
    float* d_filts = nullptr;            // [2][filter entries]: one-sided, symmetrised
    size_t filt_total = 0;
    bool pair_filter = false;            // LINEAR modes without exact_filter
    rv::AngleEntry* d_ang = nullptr;
    unsigned int* d_counter = nullptr;   // kXcds counters per group
    void* d_stitch = nullptr;            // stitched offset-CT rows
    int* d_pairs = nullptr;
    unsigned char* d_packed = nullptr;   // multi-slice work items: packed rows
    float* d_gscale = nullptr;           // FP16-pair items: item scales
    unsigned char* d_tgeom = nullptr;    // FP32 pair groups: tile geometry tables
    rv::PackJob* d_pack_jobs = nullptr;
    int n_pack = 0;
    bool pack_half = false;
    struct PackRange { int first, count, kind, rows; };
    std::vector<PackRange> pack_ranges;  // one per group: jobs of one kind
    void* d_fft_work = nullptr;          // two-pass transforms (n > 16384): work rows
    size_t fft_work_rows = 0;
    std::vector<void*> cur_out;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void*> g_sino;
    std::vector<void*> g_out;
};

namespace {

size_t align_up(size_t b, size_t a) { return (b + a - 1) / a * a; }

void plan_free(rv_plan_t p)
{
    if (!p) return;
    (void)hipSetDevice(p->dev);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    for (auto& G : p->groups) {
        (void)hipFree(G.d_slices);
        (void)hipFree(G.d_order);
    }
    (void)hipFree(p->d_rows);
    (void)hipFree(p->d_scales);
    (void)hipFree(p->d_filts);
    (void)hipFree(p->d_ang);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_stitch);
    (void)hipFree(p->d_pairs);
    (void)hipFree(p->d_packed);
    (void)hipFree(p->d_gscale);
    (void)hipFree(p->d_tgeom);
    (void)hipFree(p->d_pack_jobs);
    (void)hipFree(p->d_fft_work);
    delete p;
}

int fail(rv_plan_t p, const std::string& m)
{
    rv::set_message(m);
    plan_free(p);
    return -1;
}

// Image pointers of the slice tables, patched by kernels on the execution's
// stream (no host copy that could overtake work still queued on that stream).
// all = true patches every entry (a graph capture: each replay then restores
// its own pointers, whatever eager executions wrote in between).
bool upload_slices(rv_plan_t p, void* const* d_out, hipStream_t st, bool all)
{
    rv::PtrPatch P{};
    hipError_t err = hipSuccess;
    auto add = [&](void** dst, void* val) {
        P.dst[P.n] = dst;
        P.val[P.n] = val;
        if (++P.n == rv::kPtrPatch) {
            rv::launch_ptr_patch(P, st);
            if (err == hipSuccess) err = hipGetLastError();
            P.n = 0;
        }
    };
    for (auto& G : p->groups) {
        for (size_t q = 0; q < G.slices.size(); q++) {
            const int s = G.slices[q];
            if (all || p->cur_out[s] != d_out[s]) {
                G.h[q].out = d_out[s];
                add(&G.d_slices[q].out, d_out[s]);
            }
            for (int e = 0; e + 1 < rv::BP_MAX_SLICES_PER_ITEM; e++) {
                const int s2 = G.others[q][e];
                if (s2 >= 0 && (all || p->cur_out[s2] != d_out[s2])) {
                    G.h[q].out_x[e] = d_out[s2];
                    add(&G.d_slices[q].out_x[e], d_out[s2]);
                }
            }
        }
    }
    if (P.n) {
        rv::launch_ptr_patch(P, st);
        if (err == hipSuccess) err = hipGetLastError();
    }
    if (!rv::hip_ok(err, "slice table patch")) return false;
    // a capture runs nothing: the caller records the pointers once the graph is launched
    if (!all)
        for (size_t s = 0; s < p->sl.size(); s++) p->cur_out[s] = d_out[s];
    return true;
}

rv::FilterJob filter_job(rv_plan_t p, size_t s)
{
    const SliceState& S = p->sl[s];
    const rv::Geom& g = S.g;
    const size_t esz = (p->store == 0) ? 4 : 2;
    rv::FilterJob J{};
    J.in_stride = g.W;
    J.nrows = S.nang;
    J.in_kind = (S.spec.sino_dtype != RV_SINO_I16) ? 1 : 0;
    J.W = g.W; J.z = g.z; J.ic = g.ic; J.n = g.n; J.np = g.ndimp; J.N = g.N;
    const bool zero = (S.spec.pad_mode == RV_PAD_ZERO);
    J.place = zero ? g.n / 2 : g.n / 2 - 1;
    J.pad_mode = zero ? 1 : 0;
    J.crop = zero ? g.n / 2 - g.h : g.ihoffset;
    J.filt = p->d_filts + S.filt_off;
    J.tw = rv::device_twiddles(p->dev, g.ndimp);
    if (p->pair_filter && g.ndimp <= rv::kMaxLdsFftLog2) {
        J.pair = 1;
        J.filt = p->d_filts + p->filt_total + S.filt_off;
        J.tw = rv::device_twiddles_full(p->dev, g.ndimp);
    } else if (p->pair_filter) {
        // two-pass transform (CFft's network) with exp(-2 pi i k / n) rounded from
        // double instead of the reference's float-angle twiddles: the same
        // operator, FP32-accurate at 2^15..2^20 points (exact_filter keeps the
        // reference's table and its rounding)
        J.tw = rv::device_twiddles_full(p->dev, g.ndimp);
    }
    J.out_stride = g.N;
    J.store = p->store;
    if (p->store == 0) {
        J.out = (float*)p->d_rows + S.rows_off;
    } else {
        J.out_lp = (unsigned char*)p->d_rows + S.rows_off * esz;
        J.out_scale = p->d_scales + S.scale_off;
    }
    return J;
}

// Slices of one filter class (same transform, filter table, row count and
// input kind) are filtered by one launch, up to kFilterBatch slices each; the
// slices are visited class by class (filter_order), whatever their order in
// the plan.
int launch_filters(rv_plan_t p, const void* const* d_sino, hipStream_t st)
{
    rv::FilterJob head{};
    rv::FilterBatch fb{};
    int count = 0, cls = -1;
    hipError_t err = hipSuccess;
    auto flush = [&]() {
        if (count && err == hipSuccess) err = rv::launch_filter_batch(head, fb, count, st);
        count = 0;
    };
    for (const int s : p->filter_order) {
        const SliceState& S = p->sl[s];
        const bool is_float = (S.spec.sino_dtype != RV_SINO_I16);
        const void* src = d_sino[s];
        if (S.offset) {
            rv::StitchJob J{};
            J.sino = d_sino[s];
            J.W = S.raw_W;
            J.bin = 1;
            J.trim = 0;
            J.center = S.spec.center;
            J.deltac = 0.0;
            J.pair = p->d_pairs + S.pair_off;
            J.nout = S.nang;
            J.out_width = S.g.W;
            J.out = (unsigned char*)p->d_stitch + S.stitch_off * (is_float ? 4 : 2);
            J.is_float = is_float ? 1 : 0;
            rv::launch_stitch(J, st);
            src = J.out;
        }
        rv::FilterJob J = filter_job(p, s);
        if (!J.tw) { rv::set_message("twiddle table allocation failed"); return -1; }
        if (S.g.ndimp > rv::kMaxLdsFftLog2) {
            // two-pass transform through the plan's work rows, one slice at a time
            flush();
            cls = -1;
            J.sino = src;
            if (err == hipSuccess) err = rv::launch_filter_large(J, p->d_fft_work, p->fft_work_rows, st);
            continue;
        }
        if (count == rv::kFilterBatch || S.fclass != cls) flush();
        if (count == 0) {
            head = J;
            cls = S.fclass;
        }
        fb.sino[count] = src;
        fb.out[count] = J.out;
        fb.out_lp[count] = J.out_lp;
        fb.out_scale[count] = J.out_scale;
        count++;
    }
    flush();
    return rv::hip_ok(err, "filter launch") && rv::hip_ok(hipGetLastError(), "filter launch") ? 0 : -1;
}

void launch_packing(rv_plan_t p, hipStream_t st)
{
    if (p->pack_half) rv::launch_pack_scales(p->d_pack_jobs, p->n_pack, st);
    for (const auto& R : p->pack_ranges)
        rv::launch_pack(p->d_pack_jobs + R.first, R.count, R.kind, p->store, R.rows, st);
}

int launch_backprojection(rv_plan_t p, hipStream_t st, bool pack = true)
{
    if (pack) launch_packing(p, st);
    for (size_t gi = 0; gi < p->groups.size(); gi++) {
        const Group& G = p->groups[gi];
        rv::BpLaunch L{};
        L.slices = G.d_slices;
        L.nslices = (int)G.slices.size();
        L.total_items = G.total_items;
        L.item_order = G.d_order;
        unsigned int* ctr = p->d_counter + gi * (size_t)(rv::kXcds * rv::kCounterStride);
        L.counter = p->opt.persistent ? ctr : nullptr;
        L.run = rv::work_run(G.total_items, rv::bp_deals_runs(G.mode));
        L.mode = G.mode;
        L.rule = p->opt.index_rule ? 1 : 0;
        L.chunk = G.chunk;
        L.tile = G.tile;
        L.tgeom = G.tgeom ? 1 : 0;
        if (L.counter) rv::launch_zero_u32(ctr, rv::kXcds * rv::kCounterStride, st);
        rv::launch_bp(L, G.grid, st);
    }
    return rv::hip_ok(hipGetLastError(), "back-projection launch") ? 0 : -1;
}

}  // namespace

extern "C" {

int rvPlanCreate(rv_plan_t* out, int device, const RvSliceSpec* specs, int nslices, const RvOptions* opt)
{
    *out = nullptr;
    if (nslices <= 0 || !specs || !opt) { rv::set_message("no slices"); return -1; }
    if (!rv::hip_ok(hipSetDevice(device), "hipSetDevice")) return -1;
    rv_plan_t p = new rv_plan_opaque();
    p->dev = device;
    p->opt = *opt;
    p->mode = opt->interp;
    if (p->mode < 0 || p->mode > 2) return fail(p, "unknown interpolation mode");
    p->store = (p->mode == rv::BP_LINEAR) ? opt->store : 0;
    if (p->store < 0 || p->store > 2) return fail(p, "unknown storage precision");
    // 16-bit filtered sinograms are expanded to FP32 pairs in LDS by default (the
    // kernel is LDS-bound at 8 B/update and VALU-bound at 4 B/update with the
    // extra fract); lds_half = 1 selects the FP16-pair variant.
    if (p->mode == rv::BP_LINEAR && p->store != 0 && opt->lds_half) p->mode = rv::BP_LINEAR_H;
    const bool linear = (p->mode == rv::BP_LINEAR || p->mode == rv::BP_LINEAR_H);
    p->pair_filter = linear && !opt->exact_filter;
    const int plan_tile = (opt->tile >= 0 && opt->tile <= 4) ? opt->tile : 0;

    size_t rows_elems = 0, nang_total = 0, filt_total = 0, stitch_elems = 0, pair_total = 0;
    std::vector<int> all_pairs;
    p->sl.resize(nslices);
    for (int s = 0; s < nslices; s++) {
        SliceState& S = p->sl[s];
        S.spec = specs[s];
        if (!specs[s].fdeg || specs[s].P <= 0) return fail(p, "slice without angles");
        S.spec.fdeg = nullptr;
        S.spec.binc = nullptr;
        double center = S.spec.center;
        int W = S.spec.W;
        if (S.spec.offset_ct) {
            // pair the 360-degree frames, stitch to 2W (gazoDoc.cpp:3536-3665)
            if (S.spec.iInterpolation != 2) return fail(p, "offset CT needs iInterpolation = 2");
            std::vector<int> pr(S.spec.P);
            std::vector<char> binc(S.spec.P, (char)CGAZODOC_BINC_SAMPLE);
            if (specs[s].binc) binc.assign(specs[s].binc, specs[s].binc + S.spec.P);
            const int nout = RecviewOffsetCTPairs(specs[s].fdeg, binc.data(), S.spec.P, pr.data());
            if (nout <= 0) return fail(p, "offset-CT pairing failed (no partner frames)");
            S.offset = true;
            S.raw_W = W;
            S.fdeg.assign(specs[s].fdeg, specs[s].fdeg + nout);
            S.pair_off = pair_total;
            all_pairs.insert(all_pairs.end(), pr.begin(), pr.begin() + nout);
            pair_total += nout;
            // DeconvBackProj's centre remap on the user centre (gazoDoc.cpp:3832-3837)
            if (center - 0.5 <= W / 2.0) center += W;
            W = 2 * W;
            S.stitch_off = stitch_elems;
            stitch_elems += (size_t)nout * W;
        } else {
            S.fdeg.assign(specs[s].fdeg, specs[s].fdeg + specs[s].P);
        }
        S.g = rv::geometry(W, S.spec.iInterpolation, center, true);
        if (W < 4 || S.g.ndimp < 3 || S.g.ndimp > rv::kMaxFftLog2)
            return fail(p, "detector width / zoom outside the supported FFT range (8 <= n <= 2^20)");
        if (S.spec.out_n > 0) {
            if (!linear) return fail(p, "custom output grids need the LINEAR mode");
            if (S.g.z != 1) return fail(p, "custom output grids need iInterpolation = 2");
            if (!(S.spec.pix_zoom >= 1.0)) return fail(p, "pix_zoom must be >= 1");
            S.out_n = S.spec.out_n;
        } else {
            S.out_n = S.g.N;
        }
        S.nang = (int)S.fdeg.size();
        S.ang_off = nang_total;
        S.filt_off = filt_total;
        nang_total += (size_t)S.nang;
        if (S.spec.rows_of > 0) {
            // the filtered rows of an earlier slice, shifted by the integer centre offset
            if (S.spec.rows_of > s) return fail(p, "rows_of must name an earlier slice");
            int r = S.spec.rows_of - 1;
            if (p->sl[r].rows_src >= 0) r = p->sl[r].rows_src;
            const SliceState& R = p->sl[r];
            if (S.offset || R.offset) return fail(p, "offset-CT slices cannot share filtered rows");
            if (S.g.W != R.g.W || S.g.z != R.g.z || S.g.n != R.g.n || S.nang != R.nang ||
                S.spec.filter != R.spec.filter || S.spec.cutoff != R.spec.cutoff || S.spec.order != R.spec.order ||
                S.spec.pad_mode != R.spec.pad_mode)
                return fail(p, "rows_of needs the same detector, zoom, projections and filter");
            S.rows_src = r;
            S.rows_off = R.rows_off;
            S.scale_off = R.scale_off;
        } else {
            S.rows_off = rows_elems;
            S.scale_off = S.ang_off;
            rows_elems += (size_t)S.nang * S.g.N;
        }
        filt_total += (size_t)S.g.n;
        S.tile = (S.spec.tile >= 0 && S.spec.tile <= 4) ? S.spec.tile : plan_tile;
        S.tiles_x = (S.out_n + rv::bp_tile_side(S.tile) - 1) / rv::bp_tile_side(S.tile);
        S.tiles_y = (S.out_n + rv::bp_tile_h(S.tile) - 1) / rv::bp_tile_h(S.tile);
    }

    p->filt_total = filt_total;
    // filter tables (one-sided, then the symmetrised copies used when two rows
    // share a transform) and angle tables
    std::vector<float> hf(2 * filt_total);
    std::vector<rv::AngleEntry> ha(nang_total);
    for (auto& S : p->sl) {
        float* f = hf.data() + S.filt_off;
        rv::set_filter(S.g.n, S.spec.filter, S.spec.cutoff, S.spec.order, f);
        float* fs = hf.data() + filt_total + S.filt_off;
        for (int k = 0; k < S.g.n; k++) fs[k] = 0.5f * (f[k] + f[(S.g.n - k) % S.g.n]);
        // slices on another slice's rows: that slice's rows place sample ic_src where
        // this slice's own would place ic, so the coordinate moves by the difference
        const int icoff = S.rows_src >= 0 ? (S.g.ic - p->sl[S.rows_src].g.ic) * S.g.z : 0;
        const double fc = (double)S.g.fcenter16 / 16.0 + icoff;
        const double h = S.g.h;
        for (int a = 0; a < S.nang; a++) {
            rv::AngleEntry& e = ha[S.ang_off + a];
            std::memset(&e, 0, sizeof(e));
            const double th = rv::theta_rad(S.fdeg[a], S.spec.tilt);
            const double c = std::cos(th), sn = -std::sin(th);
            if (S.spec.out_n > 0) {
                const double zf = S.spec.pix_zoom;
                e.u0 = (S.spec.x0 - h) * c + (S.spec.y0 - h) * sn + fc;
                e.dux = c / zf;
                e.duy = sn / zf;
            } else {
                e.u0 = -h * c - h * sn + fc;
                e.dux = c;
                e.duy = sn;
            }
            rv::lut_params(S.fdeg[a], S.spec.tilt, S.g, icoff, &e.fcos, &e.fsin, &e.foffset);
            e.row = a;
        }
    }

    // filter classes: equal transform geometry, row count, input kind and filter table
    for (int s = 0; s < nslices; s++) {
        SliceState& S = p->sl[s];
        S.fclass = s;
        if (S.rows_src >= 0) continue;
        for (int t = 0; t < s; t++) {
            const SliceState& T = p->sl[t];
            if (T.fclass != t || T.rows_src >= 0) continue;
            if (T.g.W == S.g.W && T.g.z == S.g.z && T.g.ic == S.g.ic && T.g.n == S.g.n &&
                T.g.ndimp == S.g.ndimp && T.g.N == S.g.N && T.g.h == S.g.h && T.g.ihoffset == S.g.ihoffset &&
                T.nang == S.nang && T.spec.pad_mode == S.spec.pad_mode &&
                (T.spec.sino_dtype != RV_SINO_I16) == (S.spec.sino_dtype != RV_SINO_I16) &&
                std::memcmp(hf.data() + T.filt_off, hf.data() + S.filt_off, sizeof(float) * S.g.n) == 0) {
                S.fclass = t;
                break;
            }
        }
    }
    // filter launch order: class by class, each in slice order (slices that
    // back-project another slice's rows are not filtered)
    {
        std::vector<char> placed(nslices, 0);
        for (int s = 0; s < nslices; s++) {
            if (placed[s] || p->sl[s].rows_src >= 0) continue;
            for (int t = s; t < nslices; t++)
                if (!placed[t] && p->sl[t].rows_src < 0 && p->sl[t].fclass == p->sl[s].fclass) {
                    p->filter_order.push_back(t);
                    placed[t] = 1;
                }
        }
    }

    // LINEAR modes: slices whose back-projection geometry is identical (angle
    // table, sizes, tile) share work items -- up to four with FP32 values in LDS,
    // up to 32 with FP16 pairs (classes of equal geometry, in slice order).
    std::vector<std::vector<int>> classes;
    if (opt->slices_per_item != 0 && opt->slices_per_item != 1 && opt->slices_per_item != 2 &&
        opt->slices_per_item != 4 && opt->slices_per_item != 8 && opt->slices_per_item != 16 &&
        opt->slices_per_item != 32)
        return fail(p, "slices_per_item must be 0, 1, 2, 4, 8, 16 or 32");
    // default: four FP32 slices or sixteen FP16-pair slices per item, all on the
    // vector ALU (packed rows; sixteen form on the 32x64 and 32x32 tiles, eight on
    // 64x64).  Measured per 2048^2 x 1800 slice (64-slice stack): FP32 0.57 ms with
    // four per item, 0.58 ms with two; FP16 pairs 0.367 / 0.381 / 0.492 ms with
    // sixteen (32x64) / eight (64x64) / four per item.  A class's leftovers run as
    // a pair and a single.  32 per item (opt-in) runs the matrix-core kernel: 0.343 ms.
    const int max_per_item = (p->mode == rv::BP_LINEAR_H) ? 32 : 4;
    const int per_item = std::min(opt->slices_per_item == 0 ? (p->mode == rv::BP_LINEAR_H ? 16 : 4)
                                                            : opt->slices_per_item, max_per_item);
    if ((p->mode == rv::BP_LINEAR || p->mode == rv::BP_LINEAR_H) && per_item > 1) {
        auto same = [&](int a, int b) {
            const SliceState& A = p->sl[a];
            const SliceState& B = p->sl[b];
            return A.nang == B.nang && A.g.N == B.g.N && A.out_n == B.out_n && A.tile == B.tile &&
                   std::memcmp(ha.data() + A.ang_off, ha.data() + B.ang_off, sizeof(rv::AngleEntry) * A.nang) == 0;
        };
        std::map<std::vector<long long>, std::vector<int>> by_key;   // key -> class indices
        for (int s = 0; s < nslices; s++) {
            const SliceState& S = p->sl[s];
            // cheap key first (sizes and the first and last angle entries), full comparison on a match
            const rv::AngleEntry& e0 = ha[S.ang_off];
            const rv::AngleEntry& e1 = ha[S.ang_off + S.nang - 1];
            std::vector<long long> key{S.nang, S.g.N, S.out_n, S.tile};
            for (const double d : {e0.u0, e0.dux, e0.duy, e1.u0, e1.dux, e1.duy}) {
                long long bits;
                std::memcpy(&bits, &d, sizeof(bits));
                key.push_back(bits);
            }
            auto& lst = by_key[key];
            bool found = false;
            for (const int c : lst) {
                if (same(classes[c][0], s)) {
                    classes[c].push_back(s);
                    found = true;
                    break;
                }
            }
            if (!found) {
                lst.push_back((int)classes.size());
                classes.push_back({s});
            }
        }
    } else {
        for (int s = 0; s < nslices; s++) classes.push_back({s});
    }

    // groups by tile configuration and slices per item; work items cost-sorted within each group
    std::map<std::pair<int, int>, int> gidx;
    for (const auto& cls : classes) {
        size_t i = 0;
        while (i < cls.size()) {
            const size_t left = cls.size() - i;
            const int tile = p->sl[cls[i]].tile;
            const bool tile8 = tile == 0 || tile == 2 || tile == 4;    // eight per item: 64^2, 32^2, 32x64
            const bool tile16 = tile == 2 || tile == 4;                // sixteen: 32^2 and 32x64
            const int k = (per_item >= 32 && left >= 32) ? 32                  // matrix cores, 32x32 tiles
                        : (per_item >= 16 && left >= 16 && tile16) ? 16
                        : (per_item >= 8 && left >= 8 && tile8) ? 8
                        : (per_item >= 4 && left >= 4) ? 4 : (per_item >= 2 && left >= 2) ? 2 : 1;
            const bool half = (p->mode == rv::BP_LINEAR_H);
            // FP16 pairs in LDS pay from four slices per item on; the pairs and singles
            // left over by wider items run the FP32-pair kernels on the 16-bit rows
            // (2048^2 x 1800: 37.6 vs 45.9 ms for two per item, 52.7 vs 76.0 ms for
            // one), unless the plan asked for at most two per item
            const bool narrow_fp32 = half && per_item > 2;
            if (k == 32)
                for (int e = 0; e < k; e++) {
                    SliceState& T = p->sl[cls[i + e]];
                    T.tile = 2;
                    T.tiles_x = (T.out_n + rv::bp_tile_side(2) - 1) / rv::bp_tile_side(2);
                    T.tiles_y = (T.out_n + rv::bp_tile_h(2) - 1) / rv::bp_tile_h(2);
                }
            const int mode = k == 32 ? (int)rv::BP_LINEAR_M32
                           : k == 16 ? (int)rv::BP_LINEAR_H16
                           : k == 8 ? (int)rv::BP_LINEAR_H8
                           : k == 4 ? (half ? (int)rv::BP_LINEAR_H4 : (int)rv::BP_LINEAR4)
                           : k == 2 ? (half && !narrow_fp32 ? (int)rv::BP_LINEAR_H2 : (int)rv::BP_LINEAR2)
                                    : (narrow_fp32 ? (int)rv::BP_LINEAR : p->mode);
            const int s = cls[i];
            const auto gk = std::make_pair(p->sl[s].tile, mode);
            if (gidx.find(gk) == gidx.end()) {
                gidx[gk] = (int)p->groups.size();
                Group G;
                G.tile = p->sl[s].tile;
                G.mode = mode;
                G.chunk = p->sl[s].spec.chunk > 0 ? p->sl[s].spec.chunk
                        : opt->angles_per_chunk > 0 ? opt->angles_per_chunk : rv::bp_default_chunk(mode, G.tile);
                if (G.chunk > rv::bp_max_chunk(mode)) G.chunk = rv::bp_max_chunk(mode);
                while (G.chunk > 1 && rv::bp_lds_bytes(mode, G.tile, G.chunk) > 160 * 1024) G.chunk /= 2;
                p->groups.push_back(G);
            }
            Group& G = p->groups[gidx[gk]];
            G.slices.push_back(s);
            std::array<int, rv::BP_MAX_SLICES_PER_ITEM - 1> others;
            others.fill(-1);
            for (int e = 1; e < k; e++) others[e - 1] = cls[i + e];
            G.others.push_back(others);
            for (int e = 0; e < k; e++) p->sl[cls[i + e]].group = gidx[gk];
            i += k;
        }
    }

    const size_t esz = (p->store == 0) ? 4 : 2;
    const size_t in_esz = 4;   // stitch buffer sized for float rows (int16 uses half)
    if (!rv::hip_ok(hipMalloc(&p->d_rows, esz * std::max<size_t>(rows_elems, 1)), "filtered-sinogram workspace") ||
        !rv::hip_ok(hipMalloc(&p->d_scales, sizeof(float) * nang_total), "row scales") ||
        !rv::hip_ok(hipMalloc(&p->d_filts, sizeof(float) * 2 * filt_total), "filters") ||
        !rv::hip_ok(hipMalloc(&p->d_ang, sizeof(rv::AngleEntry) * nang_total), "angle table") ||
        !rv::hip_ok(hipMalloc(&p->d_counter, sizeof(unsigned int) * rv::kXcds * rv::kCounterStride *
                                     (p->groups.size() + 1)), "work counters"))
        return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
    if (stitch_elems) {
        if (!rv::hip_ok(hipMalloc(&p->d_stitch, in_esz * stitch_elems), "stitch buffer") ||
            !rv::hip_ok(hipMalloc(&p->d_pairs, sizeof(int) * pair_total), "pair table") ||
            !rv::hip_ok(hipMemcpy(p->d_pairs, all_pairs.data(), sizeof(int) * pair_total, hipMemcpyHostToDevice),
                        "pairs"))
            return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
    }
    if (!rv::hip_ok(hipMemcpy(p->d_filts, hf.data(), sizeof(float) * 2 * filt_total, hipMemcpyHostToDevice), "filters") ||
        !rv::hip_ok(hipMemcpy(p->d_ang, ha.data(), sizeof(rv::AngleEntry) * nang_total, hipMemcpyHostToDevice), "angles"))
        return fail(p, std::string("plan upload failed: ") + RecviewLastMessage());

    // packed rows of the multi-slice work items: FP32 pairs 16 B, FP16 pairs 8 / 16 /
    // 32 / 64 B per sample (two / four / eight / sixteen slices); 32 FP16 slices in
    // chunks of eight samples
    auto packed_size = [](int mode) -> size_t {
        return mode == rv::BP_LINEAR_M32 ? 64 : mode == rv::BP_LINEAR_H16 ? 64 : mode == rv::BP_LINEAR_H8 ? 32
             : mode == rv::BP_LINEAR4 ? 32
             : mode == rv::BP_LINEAR2 || mode == rv::BP_LINEAR_H4 ? 16 : mode == rv::BP_LINEAR_H2 ? 8 : 0;
    };
    auto packed_row = [&](int mode, int N) -> size_t {     // bytes per angle row
        return mode == rv::BP_LINEAR_M32 ? (size_t)rv::bp_m32_chunks(N) * 8 * packed_size(mode)
                                         : (size_t)N * packed_size(mode);
    };
    size_t packed_bytes = 0, n_half = 0;
    for (const auto& G : p->groups)
        if (packed_size(G.mode))
            for (int s : G.slices) {
                packed_bytes += align_up((size_t)p->sl[s].nang * packed_row(G.mode, p->sl[s].g.N), 256);
                n_half += G.mode != rv::BP_LINEAR2 && G.mode != rv::BP_LINEAR4;
            }
    std::vector<rv::PackJob> pack_jobs;
    if (packed_bytes &&
        !rv::hip_ok(hipMalloc(&p->d_packed, packed_bytes), "packed multi-slice rows"))
        return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
    if (n_half && !rv::hip_ok(hipMalloc(&p->d_gscale, sizeof(float) * n_half), "item scales"))
        return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
    size_t packed_off = 0, half_idx = 0;

    // Tile geometry tables of the FP32 pair groups, one per class of identical
    // geometry (its slices share every angle entry): [tiles][angles] of 16 B,
    // formed once here; at most 8 GiB (the 256-slice batch of uneven shapes up to
    // 8192^2 x 3600 needs 3.8 GiB), groups beyond that form theirs per chunk.
    // A table depends on the geometry and the tile configuration: the matrix-core
    // grouping moves its slices to 32x32 tiles after the classes are formed, so a
    // class can run on two tile configurations.
    std::vector<int> cls_of(nslices, -1);
    for (size_t c = 0; c < classes.size(); c++)
        for (int s : classes[c]) cls_of[s] = (int)c;
    std::map<std::pair<int, int>, std::pair<size_t, int>> tg_tab;   // (class, tile) -> (offset, slice)
    size_t tg_total = 0;
    // RECVIEW_GEOMETRY_TABLES=0 (diagnostics, tests): every group forms its geometry
    // per chunk, as when the tables would exceed the budget
    const char* tg_env = std::getenv("RECVIEW_GEOMETRY_TABLES");
    const bool tg_allowed = !(tg_env && tg_env[0] == '0');
    for (auto& G : p->groups) {
        if (!tg_allowed || !rv::bp_has_tgeom(G.mode)) continue;
        G.tgeom = true;
        for (int s : G.slices) {
            const int c = cls_of[s];
            if (c < 0) { G.tgeom = false; break; }
            const SliceState& S = p->sl[s];
            const auto key = std::make_pair(c, S.tile);
            if (tg_tab.count(key)) continue;
            // + 64 B: a chunk's prefetch may read one entry past the tile's last
            const size_t bytes = align_up((size_t)S.tiles_x * S.tiles_y * S.nang * 16 + 64, 256);
            if (tg_total + bytes > ((size_t)8 << 30)) { G.tgeom = false; break; }
            tg_tab[key] = std::make_pair(tg_total, s);
            tg_total += bytes;
        }
    }
    auto tg_of = [&](int s) -> const float4* {
        return (const float4*)(p->d_tgeom + tg_tab.at(std::make_pair(cls_of[s], p->sl[s].tile)).first);
    };
    if (tg_total) {
        if (!rv::hip_ok(hipMalloc(&p->d_tgeom, tg_total), "tile geometry tables"))
            return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
        for (const auto& kv : tg_tab) {
            const SliceState& S = p->sl[kv.second.second];
            rv::launch_tile_geometry(p->d_ang + S.ang_off, S.nang, S.tiles_x, S.tiles_x * S.tiles_y,
                                     rv::bp_tile_side(S.tile), rv::bp_tile_h(S.tile),
                                     (float4*)(p->d_tgeom + kv.second.first), nullptr);
        }
        if (!rv::hip_ok(hipGetLastError(), "tile geometry") || !rv::hip_ok(hipDeviceSynchronize(), "tile geometry"))
            return fail(p, std::string("plan set-up failed: ") + RecviewLastMessage());
    }

    for (auto& G : p->groups) {
        // slice order inside the group: longest projection lists first
        std::vector<int> ord = G.slices;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return p->sl[a].nang > p->sl[b].nang; });
        int base = 0;
        for (int s : G.slices) {
            p->sl[s].tile_base = base;
            base += p->sl[s].tiles_x * p->sl[s].tiles_y;
        }
        G.total_items = base;
        std::vector<int> order;
        order.reserve(base);
        for (int s : ord) {
            const SliceState& S = p->sl[s];
            for (int t = 0; t < S.tiles_x * S.tiles_y; t++) order.push_back(S.tile_base + t);
        }
        G.h.resize(G.slices.size());
        for (size_t q = 0; q < G.slices.size(); q++) {
            const SliceState& S = p->sl[G.slices[q]];
            rv::BpSlice& B = G.h[q];
            std::memset(&B, 0, sizeof(B));
            B.rows = (const unsigned char*)p->d_rows + S.rows_off * esz;
            B.scales = (p->store == 0) ? nullptr : p->d_scales + S.scale_off;
            B.row_stride = S.g.N;
            B.row_len = S.g.N;
            B.nang = S.nang;
            B.ang = p->d_ang + S.ang_off;
            B.out = nullptr;
            B.out_n = S.out_n;
            B.tiles_x = S.tiles_x;
            B.tiles_y = S.tiles_y;
            B.tile_base = S.tile_base;
            B.accumulate = 0;
            B.lut_formula = opt->lut_formula ? 1 : 0;
            B.store = p->store;
            B.out_scale = opt->normalize ? (float)(6.283185307179586 / S.nang) : 1.f;
            B.tgeom = G.tgeom ? tg_of(G.slices[q]) : nullptr;
            for (int e = 0; e + 1 < rv::BP_MAX_SLICES_PER_ITEM; e++) {
                const int s2 = G.others[q][e];
                if (s2 < 0) continue;
                const SliceState& S2 = p->sl[s2];
                B.rows_x[e] = (const unsigned char*)p->d_rows + S2.rows_off * esz;
                B.scales_x[e] = (p->store == 0) ? nullptr : p->d_scales + S2.scale_off;
            }
            if (packed_size(G.mode)) {
                const int ns = rv::bp_slices_per_item(G.mode);
                rv::PackJob J{};
                J.rows[0] = B.rows;
                J.scales[0] = B.scales;
                for (int e = 1; e < ns; e++) {
                    J.rows[e] = B.rows_x[e - 1];
                    J.scales[e] = B.scales_x[e - 1];
                }
                J.ang = B.ang;
                J.nang = S.nang;
                J.ns = ns;
                J.half = G.mode != rv::BP_LINEAR2 && G.mode != rv::BP_LINEAR4;
                J.store = p->store;
                J.nrows = S.nang;
                J.N = S.g.N;
                J.out = p->d_packed + packed_off;
                J.gscale = J.half ? p->d_gscale + half_idx++ : nullptr;
                if (q == 0)
                    p->pack_ranges.push_back({(int)pack_jobs.size(), 0, J.half ? ns : ns == 4 ? 3 : 0, 0});
                auto& R = p->pack_ranges.back();
                R.count++;
                R.rows = std::max(R.rows, S.nang);
                pack_jobs.push_back(J);
                p->pack_half = p->pack_half || J.half;
                B.rows = J.out;                    // the kernel reads the packed rows
                B.row_stride = G.mode == rv::BP_LINEAR_M32 ? rv::bp_m32_chunks(S.g.N) : S.g.N;
                B.gscale = J.gscale;
                packed_off += align_up((size_t)S.nang * packed_row(G.mode, S.g.N), 256);
            }
        }
        if (!rv::hip_ok(hipMalloc(&G.d_slices, sizeof(rv::BpSlice) * G.h.size()), "slice table") ||
            !rv::hip_ok(hipMalloc(&G.d_order, sizeof(int) * std::max<size_t>(order.size(), 1)), "work list") ||
            !rv::hip_ok(hipMemcpy(G.d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice), "order") ||
            !rv::hip_ok(hipMemcpy(G.d_slices, G.h.data(), sizeof(rv::BpSlice) * G.h.size(), hipMemcpyHostToDevice),
                        "slices"))
            return fail(p, std::string("plan upload failed: ") + RecviewLastMessage());
        (void)rv::bp_blocks_per_cu_tg(G.mode, G.tile, G.chunk, G.tgeom);     // sets LDS attributes outside capture
        if (opt->persistent) {
            const int per_cu = rv::bp_blocks_per_cu_tg(G.mode, G.tile, G.chunk, G.tgeom);
            G.grid = std::max(1, std::min(per_cu * rv::device_cu_count(device), G.total_items));
        } else {
            G.grid = G.total_items;
        }
    }
    p->cur_out.assign(nslices, nullptr);
    if (!pack_jobs.empty()) {
        p->n_pack = (int)pack_jobs.size();
        if (!rv::hip_ok(hipMalloc(&p->d_pack_jobs, sizeof(rv::PackJob) * pack_jobs.size()), "pack jobs") ||
            !rv::hip_ok(hipMemcpy(p->d_pack_jobs, pack_jobs.data(), sizeof(rv::PackJob) * pack_jobs.size(),
                                  hipMemcpyHostToDevice), "pack jobs"))
            return fail(p, std::string("plan upload failed: ") + RecviewLastMessage());
    }

    // work rows of the two-pass transforms (1 GiB at most; batches of rows beyond)
    int big_n = 0, big_rows = 0;
    for (const auto& S : p->sl)
        if (S.g.ndimp > rv::kMaxLdsFftLog2) {
            big_n = std::max(big_n, S.g.n);
            big_rows = std::max(big_rows, S.nang);
        }
    if (big_n) {
        p->fft_work_rows = rv::filter_large_rows(big_n, (size_t)1 << 30, big_rows);
        if (!rv::hip_ok(hipMalloc(&p->d_fft_work, 2 * sizeof(float) * 2 * (size_t)big_n * p->fft_work_rows),
                        "two-pass transform work rows"))
            return fail(p, std::string("plan allocation failed: ") + RecviewLastMessage());
    }

    // Everything that allocates or sets kernel attributes happens here, so that
    // execution (and hipGraph capture of it) only launches kernels.
    int max_n = 0;
    for (const auto& S : p->sl) {
        if (!rv::device_twiddles(device, S.g.ndimp) ||
            (p->pair_filter && !rv::device_twiddles_full(device, S.g.ndimp)))
            return fail(p, "twiddle table allocation failed");
        if (S.g.ndimp <= rv::kMaxLdsFftLog2) max_n = std::max(max_n, S.g.n);
    }
    rv::prepare_filter(max_n);
    *out = p;
    return 0;
}

// Every slice needs a non-null device pointer (the arrays hold one per slice).
static bool all_set(rv_plan_t p, const void* const* ptrs, const char* what)
{
    if (!ptrs) { rv::set_message(std::string(what) + " array is null"); return false; }
    for (size_t s = 0; s < p->sl.size(); s++)
        if (!ptrs[s]) { rv::set_message(std::string(what) + " pointer of slice " + std::to_string(s) + " is null"); return false; }
    return true;
}

int rvPlanFilter(rv_plan_t p, const void* const* d_sino, rv_stream_t stream)
{
    if (!p || !all_set(p, d_sino, "sinogram")) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    return launch_filters(p, d_sino, (hipStream_t)stream);
}

int rvPlanBackproject(rv_plan_t p, void* const* d_out, rv_stream_t stream)
{
    if (!p || !all_set(p, (const void* const*)d_out, "image")) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    if (!upload_slices(p, d_out, (hipStream_t)stream, false)) return -1;
    return launch_backprojection(p, (hipStream_t)stream);
}

int rvPlanPack(rv_plan_t p, rv_stream_t stream)
{
    if (!p) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    launch_packing(p, (hipStream_t)stream);
    return rv::hip_ok(hipGetLastError(), "pack launch") ? 0 : -1;
}

int rvPlanBackprojectPacked(rv_plan_t p, void* const* d_out, rv_stream_t stream)
{
    if (!p || !all_set(p, (const void* const*)d_out, "image")) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    if (!upload_slices(p, d_out, (hipStream_t)stream, false)) return -1;
    return launch_backprojection(p, (hipStream_t)stream, false);
}

int rvPlanExecute(rv_plan_t p, const void* const* d_sino, void* const* d_out, rv_stream_t stream)
{
    if (!p || !all_set(p, d_sino, "sinogram") || !all_set(p, (const void* const*)d_out, "image")) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    if (!upload_slices(p, d_out, (hipStream_t)stream, false)) return -1;
    if (launch_filters(p, d_sino, (hipStream_t)stream)) return -1;
    return launch_backprojection(p, (hipStream_t)stream);
}

int rvPlanExecuteGraph(rv_plan_t p, const void* const* d_sino, void* const* d_out, rv_stream_t stream)
{
    if (!p || !all_set(p, d_sino, "sinogram") || !all_set(p, (const void* const*)d_out, "image")) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    const size_t ns = p->sl.size();
    bool same = (p->gexec != nullptr) && p->g_sino.size() == ns;
    for (size_t s = 0; same && s < ns; s++) same = (p->g_sino[s] == d_sino[s]) && (p->g_out[s] == d_out[s]);
    if (!same) {
        if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        if (!p->cap_stream && !rv::hip_ok(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking), "capture stream"))
            return -1;
        // The graph holds kernel nodes only (no memset/memcpy nodes); its first
        // nodes write its image pointers into the slice tables.
        if (!rv::hip_ok(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal), "begin capture"))
            return -1;
        int rc = upload_slices(p, d_out, p->cap_stream, true) ? 0 : -1;
        if (!rc) rc = launch_filters(p, d_sino, p->cap_stream);
        if (!rc) rc = launch_backprojection(p, p->cap_stream);
        hipGraph_t g = nullptr;
        const bool ended = rv::hip_ok(hipStreamEndCapture(p->cap_stream, &g), "end capture");
        if (rc || !ended) {
            if (g) (void)hipGraphDestroy(g);
            return -1;
        }
        p->graph = g;
        if (!rv::hip_ok(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0), "graph instantiate")) return -1;
        p->g_sino.assign(d_sino, d_sino + ns);
        p->g_out.assign(d_out, d_out + ns);
    }
    if (!rv::hip_ok(hipGraphLaunch(p->gexec, (hipStream_t)stream), "graph launch")) return -1;
    for (size_t s = 0; s < ns; s++) p->cur_out[s] = d_out[s];   // the replay wrote them
    return 0;
}

int rvPlanFilteredRows(rv_plan_t p, int s, void** rows, float** scales, int* row_len)
{
    if (!p || s < 0 || s >= (int)p->sl.size()) return -1;
    const SliceState& S = p->sl[s];
    const size_t esz = (p->store == 0) ? 4 : 2;
    if (rows) *rows = (unsigned char*)p->d_rows + S.rows_off * esz;
    if (scales) *scales = (p->store == 0) ? nullptr : p->d_scales + S.scale_off;
    if (row_len) *row_len = S.g.N;
    return 0;
}

int rvPlanSetFilteredRows(rv_plan_t p, int s, const float* d_rows, rv_stream_t stream)
{
    if (!p || s < 0 || s >= (int)p->sl.size() || p->store != 0) return -1;
    if (!rv::hip_ok(hipSetDevice(p->dev), "hipSetDevice")) return -1;
    const SliceState& S = p->sl[s];
    return rv::hip_ok(hipMemcpyAsync((float*)p->d_rows + S.rows_off, d_rows,
                                     sizeof(float) * (size_t)S.nang * S.g.N, hipMemcpyDeviceToDevice,
                                     (hipStream_t)stream), "rows copy") ? 0 : -1;
}

int rvPlanInfo(rv_plan_t p, int s, int* out_n, int* n_fft, int* tiles, int* tile_w, int* tile_h)
{
    if (!p || s < 0 || s >= (int)p->sl.size()) return -1;
    const SliceState& S = p->sl[s];
    if (out_n) *out_n = S.out_n;
    if (n_fft) *n_fft = S.g.n;
    if (tiles) *tiles = S.tiles_x * S.tiles_y;
    if (tile_w) *tile_w = rv::bp_tile_side(S.tile);
    if (tile_h) *tile_h = rv::bp_tile_h(S.tile);
    return 0;
}

int rvPlanRows(rv_plan_t p, int s, int* nrows, int* width)
{
    if (!p || s < 0 || s >= (int)p->sl.size()) return -1;
    if (nrows) *nrows = p->sl[s].nang;
    if (width) *width = p->sl[s].g.W;
    return 0;
}

void rvPlanDestroy(rv_plan_t p) { plan_free(p); }

}  // extern "C"
