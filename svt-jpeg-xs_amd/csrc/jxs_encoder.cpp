/*
 * jxs_encoder.cpp -- encoder runtime of the MI355X library: launch plans
 * (job lists and frame descriptors resident in HBM), the plain
 * send_picture / get_packet path, and the C-ABI entry points of
 * SvtJpegxsEnc.h and svt_jpegxs_amd.h.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "SvtJpegxsEnc.h"
#include "jxs_enc.h"
#include "svt_jpegxs_amd.h"

extern "C" {
hipError_t jxs_launch_dwt(const JxsDwtJob*, const uint32_t*, int, uint32_t, uint32_t, uint32_t*, uint32_t, hipStream_t);
hipError_t jxs_launch_rate(const JxsFrameDesc*, int, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);
hipError_t jxs_launch_pack(const JxsUnit*, uint32_t, uint32_t, hipStream_t);
hipError_t jxs_launch_join(const JxsUnit*, uint32_t, hipStream_t);
hipError_t jxs_launch_enc_slice(const JxsFrameDesc*, int, uint32_t, hipStream_t);
}

namespace jxs {

#define CHECK_HIP(x)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "svt-jpeg-xs-amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return SvtJxsErrorInsufficientResources;                                        \
        }                                                                                   \
    } while (0)

/* Shadow words of a frame: one per codestream dword (the packer's boundary
 * words), or, for a frame k_enc_slice codes, the staging words of its
 * slices, each slice starting a word of its own. */
static uint64_t shadow_bytes(const JxsGeom& g) { return ((uint64_t)g.lcod + 3) / 4 * 4 + 4ull * g.slices + 16; }

/* Frames whose coding tools k_enc_slice handles. */
static bool geom_ordered(const JxsGeom& g) { return g.vpred != 0 || g.signs != 0; }

/* Work bytes of one frame of a geometry (inputs and outputs excluded). */
static uint64_t frame_work_bytes(const JxsGeom& g) {
    uint64_t b = 0;
    b += align256(g.n_coeff * 2);
    b += align256(g.n_gcli);
    b += align256(g.n_smax);
    b += align256(g.n_i32 * 4);
    b += align256((uint64_t)g.rows * g.n_bl * JXS_HREC * 2);
    b += align256((uint64_t)g.rows * sizeof(JxsPrec));
    b += align256((uint64_t)g.rows * sizeof(JxsRateRec));
    b += align256(shadow_bytes(g));
    return b;
}

void DevGeom::release() {
    if (d_geom)
        (void)hipFree(d_geom);
    if (d_slice_off)
        (void)hipFree(d_slice_off);
    if (d_slice_bud)
        (void)hipFree(d_slice_bud);
    d_geom = nullptr;
    d_slice_off = d_slice_bud = nullptr;
}

SvtJxsErrorType_t DevGeom::upload() {
    CHECK_HIP(hipMalloc(&d_geom, sizeof(JxsGeom)));
    CHECK_HIP(hipMemcpy(d_geom, &host.g, sizeof(JxsGeom), hipMemcpyHostToDevice));
    const size_t n = host.g.slices * sizeof(uint32_t);
    CHECK_HIP(hipMalloc(&d_slice_off, n));
    CHECK_HIP(hipMalloc(&d_slice_bud, n));
    CHECK_HIP(hipMemcpy(d_slice_off, host.slice_off.data(), n, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_slice_bud, host.slice_bud.data(), n, hipMemcpyHostToDevice));
    work = frame_work_bytes(host.g);
    /* how many transform jobs one frame has per stage does not depend on
     * where its buffers lie: count them on placeholder addresses (nothing is
     * dereferenced) */
    static const uint8_t placeholder[256] = {};
    const void* planes[4] = {placeholder, placeholder, placeholder, placeholder};
    const uint32_t strides[4] = {host.g.w, host.g.w, host.g.w, host.g.w};
    FwdJobs jobs;
    fwd_jobs(host.g, planes, strides, nullptr, nullptr, nullptr, nullptr, nullptr, jobs);
    for (int s = 0; s < 2; s++)
        njobs[s] = jobs.n[s];
    return SvtJxsErrorNone;
}

uint64_t BatchSizer::upload_bytes() const {
    uint64_t b = align256((uint64_t)n * sizeof(JxsFrameDesc));
    for (int s = 0; s < 2; s++)
        b += 3 * (align256((uint64_t)njobs[s] * sizeof(JxsDwtJob)) + align256((uint64_t)njobs[s] * 4));
    return b;
}

uint64_t BatchSizer::bytes() const {
    return upload_bytes() + align256((uint64_t)n * 4) + align256((uint64_t)units * sizeof(JxsUnit)) + work;
}

struct EncoderPriv;

}  // namespace jxs

/* A launch plan: one device block holding a batch (descriptors, job lists,
 * packing units and work buffers of every frame), optionally captured in a
 * HIP graph. */
struct svt_jpeg_xs_amd_plan {
    uint32_t nfr = 0;
    std::vector<const jxs::DevGeom*> geoms;
    void* d_block = nullptr;
    jxs::BatchLaunch L;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace jxs {

struct EncoderPriv {
    DevGeom geom;
    uint32_t verbose = 0;
    /* plain path */
    struct Slot {
        void* d_in = nullptr;
        uint8_t* d_out = nullptr;
        svt_jpeg_xs_amd_plan_t* plan = nullptr;
        hipEvent_t done = nullptr;
        uint32_t* h_status = nullptr; /* pinned */
        svt_jpeg_xs_frame_t frame;
        bool busy = false;
        bool finished = false; /* set by the finisher once the frame's work is done */
        bool failed = false;
    };
    std::vector<Slot> slots;
    std::deque<int> pending; /* in send order, until get_packet takes them */
    std::deque<int> watch;   /* in send order, until the finisher has seen them complete */
    std::thread finisher;
    bool stop = false;
    svt_jpeg_xs_encoder_api_t* api = nullptr;
    /* slice packetization (FinalStageProcess.c:126-213): the frame whose
     * packets are being handed out, the next packet (0: header, 1 + s: slice
     * s) and its byte offset in the codestream */
    bool slice_mode = false;
    int rel_slot = -1;
    uint32_t rel_pkt = 0, rel_off = 0;
    bool rel_failed = false;
    std::mutex mu;
    std::condition_variable cv_free, cv_pending, cv_watch;
    hipStream_t stream = nullptr;
    uint64_t in_bytes = 0;
    size_t plane_off[4] = {0, 0, 0, 0};
    uint32_t plane_stride[4] = {0, 0, 0, 0}; /* samples */
    uint32_t plane_rows[4] = {0, 0, 0, 0};
    uint32_t plane_row_bytes[4] = {0, 0, 0, 0};
    uint32_t nplanes = 0;
};

const DevGeom* encoder_geom(const svt_jpeg_xs_encoder_api_t* api) {
    return api && api->private_ptr ? &((const EncoderPriv*)api->private_ptr)->geom : nullptr;
}

const DevGeom* spec_geom(const svt_jpeg_xs_amd_frame_spec_t& s) {
    /* the packing kernel writes the codestream as aligned dwords */
    if (!s.d_out || ((uintptr_t)s.d_out & 3) != 0)
        return nullptr;
    const DevGeom* dg = encoder_geom(s.encoder);
    if (!dg)
        return nullptr;
    const JxsGeom& g = dg->host.g;
    const uint32_t np = g.packed ? 1 : g.nc;
    for (uint32_t c = 0; c < np; c++)
        if (!s.d_planes[c] || s.strides[c] < (g.packed ? 3 * g.w : g.c[c].w))
            return nullptr;
    return dg;
}

bool batch_place(const svt_jpeg_xs_amd_frame_spec_t* specs, uint32_t count, uint8_t* base, BatchLaunch& L,
                 std::vector<uint8_t>& img) {
    BatchSizer sz;
    for (uint32_t i = 0; i < count; i++) {
        const DevGeom* dg = spec_geom(specs[i]);
        if (!dg)
            return false;
        sz.add(*dg);
    }
    img.assign(sz.upload_bytes(), 0);
    return batch_place_into(specs, count, base, L, img.data(), img.size());
}

bool batch_place_into(const svt_jpeg_xs_amd_frame_spec_t* specs, uint32_t count, uint8_t* base, BatchLaunch& L,
                      uint8_t* img, uint64_t img_bytes) {
    BatchSizer sz;
    for (uint32_t i = 0; i < count; i++) {
        const DevGeom* dg = spec_geom(specs[i]);
        if (!dg)
            return false;
        sz.add(*dg);
    }
    if (img_bytes != sz.upload_bytes())
        return false;
    memset(img, 0, img_bytes);
    L = BatchLaunch();
    L.nfr = count;
    /* block: frame descriptors | job list and block bases per stage and
     * chunk class | (uploaded part ends) error words | packing units | the
     * frames' work buffers */
    uint64_t off = 0;
    L.frames = (JxsFrameDesc*)(base + off);
    const uint64_t frames_off = off;
    off += align256((uint64_t)count * sizeof(JxsFrameDesc));
    uint64_t job_off[2][3], base_off[2][3];
    for (int s = 0; s < 2; s++)
        for (int c = 0; c < 3; c++) {
            job_off[s][c] = off;
            L.dwt[s][c] = (JxsDwtJob*)(base + off);
            off += align256((uint64_t)sz.njobs[s] * sizeof(JxsDwtJob));
            base_off[s][c] = off;
            L.dbase[s][c] = (uint32_t*)(base + off);
            off += align256((uint64_t)sz.njobs[s] * 4);
        }
    L.status = (uint32_t*)(base + off);
    off += align256((uint64_t)count * 4);
    L.units = (JxsUnit*)(base + off);
    off += align256((uint64_t)sz.units * sizeof(JxsUnit));
    uint8_t* cur = base + off;
    uint32_t prec_base = 0, slice_base = 0, unit_base = 0, most_rows = 0;
    uint32_t fill[2][3] = {}, blk[2][3] = {};
    for (uint32_t i = 0; i < count; i++) {
        const DevGeom* dg = spec_geom(specs[i]);
        const JxsGeom& g = dg->host.g;
        JxsFrameDesc F;
        memset(&F, 0, sizeof(F));
        F.g = dg->d_geom;
        F.slice_off = dg->d_slice_off;
        F.slice_bud = dg->d_slice_bud;
        F.coeff = (uint16_t*)cur;
        cur += align256(g.n_coeff * 2);
        F.gcli = cur;
        cur += align256(g.n_gcli);
        F.smax = cur;
        cur += align256(g.n_smax);
        int32_t* scratch = (int32_t*)cur;
        cur += align256(g.n_i32 * 4);
        F.hist = (uint16_t*)cur;
        cur += align256((uint64_t)g.rows * g.n_bl * JXS_HREC * 2);
        F.prec = (JxsPrec*)cur;
        cur += align256((uint64_t)g.rows * sizeof(JxsPrec));
        F.rrec = (JxsRateRec*)cur;
        cur += align256((uint64_t)g.rows * sizeof(JxsRateRec));
        F.shadow = (uint32_t*)cur;
        cur += align256(shadow_bytes(g));
        F.units = L.units + unit_base;
        F.out = specs[i].d_out;
        F.status = L.status + i;
        F.rows = g.rows;
        F.npk = g.npk;
        F.prec_base = prec_base;
        F.slice_base = slice_base;
        F.unit_base = unit_base;
        F.n_units = g.n_units;
        F.pps = g.pps;
        F.slices = g.slices;
        F.n_bl = g.n_bl;
        F.ordered = geom_ordered(g) ? 1u : 0u;
        F.slice_rc = g.rc_mode >= 2 ? 1u : 0u;
        L.ordered |= F.ordered != 0;
        if (!F.ordered)
            L.rate_kinds |= 1u << F.slice_rc;
        if (!F.ordered && !F.slice_rc && g.rows > most_rows) {
            most_rows = g.rows;
            L.rate_waves = rate_waves_get(g);
        }
        memcpy(img + frames_off + (uint64_t)i * sizeof(JxsFrameDesc), &F, sizeof(F));
        prec_base += g.rows;
        slice_base += g.slices;
        unit_base += F.ordered ? 0u : g.rows * g.n_units; /* k_enc_slice packs without units */
        FwdJobs jobs;
        fwd_jobs(g, specs[i].d_planes, specs[i].strides, F.coeff, F.gcli, F.smax, F.hist, scratch, jobs);
        for (int s = 0; s < 2; s++) {
            /* the components of a packed plane read the same rows: their
             * blocks are interleaved when the jobs are alike */
            bool il = g.packed && s == 0 && jobs.n[s] > 1;
            for (uint32_t ji = 1; il && ji < jobs.n[s]; ji++) {
                const JxsDwtJob &a = jobs.dwt[s][0], &b = jobs.dwt[s][ji];
                il = b.chunk == a.chunk && b.nblk == a.nblk && b.rows == a.rows && b.rpb == a.rpb;
            }
            for (uint32_t ji = 0; ji < jobs.n[s]; ji++) {
                JxsDwtJob j = jobs.dwt[s][ji];
                const int c = j.chunk == 8 ? 0 : j.chunk == 16 ? 1 : j.chunk == 32 ? 2 : -1;
                /* outside what the transform kernel's layout covers */
                if (c < 0 || j.lev_l < 1 || (1u << j.lev_l) > j.chunk || j.tpr > (j.chunk == 8 ? 512u : 256u) ||
                    j.rpb == 0 || fill[s][c] >= sz.njobs[s])
                    return false;
                j.base = blk[s][c];
                j.il = il ? jobs.n[s] : 0u;
                j.il_first = il ? fill[s][c] - ji : 0u;
                memcpy(img + job_off[s][c] + (uint64_t)fill[s][c] * sizeof(JxsDwtJob), &j, sizeof(j));
                memcpy(img + base_off[s][c] + (uint64_t)fill[s][c] * 4, &blk[s][c], 4);
                fill[s][c]++;
                blk[s][c] += j.nblk;
            }
        }
    }
    for (int s = 0; s < 2; s++)
        for (int c = 0; c < 3; c++) {
            L.ndwt[s][c] = (int)fill[s][c];
            L.dblocks[s][c] = blk[s][c];
        }
    L.total_slices = slice_base;
    L.total_prec = prec_base;
    L.total_units = unit_base;
    return (uint64_t)(cur - base) == sz.bytes();
}

hipError_t batch_launch(const BatchLaunch& L, hipStream_t st, hipEvent_t* ev) {
    hipError_t e;
    /* every non-empty chunk class of a stage is one launch; the first launch
     * of the batch also clears the frames' error words */
    bool cleared = false;
    for (int s = 0; s < 2; s++) {
        if (ev)
            (void)hipEventRecord(ev[s], st);
        for (int c = 0; c < 3; c++) {
            if (!L.ndwt[s][c])
                continue;
            e = jxs_launch_dwt(L.dwt[s][c], L.dbase[s][c], L.ndwt[s][c], L.dblocks[s][c], 8u << c,
                               cleared ? nullptr : L.status, cleared ? 0 : L.nfr, st);
            if (e != hipSuccess)
                return e;
            cleared = true;
        }
    }
    if (ev)
        (void)hipEventRecord(ev[2], st);
    if ((e = jxs_launch_rate(L.frames, (int)L.nfr, L.total_slices, L.total_prec, L.rate_kinds, L.rate_waves, st)) !=
        hipSuccess)
        return e;
    if (ev)
        (void)hipEventRecord(ev[3], st);
    /* launches of more than about one 4K frame's packing units take the
     * shadow-word path of the packer and its join kernel */
    const bool shadow = L.total_units > JXS_SHADOW_UNITS;
    if ((e = jxs_launch_pack(L.units, L.total_units, shadow ? 1u : 0u, st)) != hipSuccess)
        return e;
    if (shadow && (e = jxs_launch_join(L.units, L.total_units, st)) != hipSuccess)
        return e;
    if (L.ordered && (e = jxs_launch_enc_slice(L.frames, (int)L.nfr, L.total_slices, st)) != hipSuccess)
        return e;
    if (ev)
        (void)hipEventRecord(ev[SVT_JPEG_XS_AMD_STAGES], st);
    return hipSuccess;
}

static void plan_free(svt_jpeg_xs_amd_plan_t* p) {
    if (!p)
        return;
    if (p->graph_exec)
        (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->d_block);
    delete p;
}

static svt_jpeg_xs_amd_plan_t* plan_build(const svt_jpeg_xs_amd_frame_spec_t* specs, uint32_t count) {
    if (!specs || count == 0)
        return nullptr;
    BatchSizer sz;
    std::vector<const DevGeom*> geoms(count);
    for (uint32_t i = 0; i < count; i++) {
        geoms[i] = spec_geom(specs[i]);
        if (!geoms[i])
            return nullptr;
        sz.add(*geoms[i]);
    }
    auto* p = new (std::nothrow) svt_jpeg_xs_amd_plan;
    if (!p)
        return nullptr;
    p->nfr = count;
    p->geoms = geoms;
    std::vector<uint8_t> img;
    if (hipMalloc(&p->d_block, sz.bytes()) != hipSuccess ||
        !batch_place(specs, count, (uint8_t*)p->d_block, p->L, img) ||
        hipMemcpy(p->d_block, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
        plan_free(p);
        return nullptr;
    }
    return p;
}

static EncoderPriv* priv_of(svt_jpeg_xs_encoder_api_t* api) {
    return api ? (EncoderPriv*)api->private_ptr : nullptr;
}

static bool version_ok(uint64_t maj, uint64_t min) {
    return !(maj > SVT_JPEGXS_API_VER_MAJOR || (maj == SVT_JPEGXS_API_VER_MAJOR && min > SVT_JPEGXS_API_VER_MINOR));
}

static void encoder_destroy(EncoderPriv* e) {
    if (!e)
        return;
    if (e->finisher.joinable()) {
        {
            std::lock_guard<std::mutex> lk(e->mu);
            e->stop = true;
        }
        e->cv_watch.notify_all();
        e->finisher.join();
    }
    if (e->stream)
        (void)hipStreamSynchronize(e->stream);
    for (auto& s : e->slots) {
        plan_free(s.plan);
        (void)hipFree(s.d_in);
        (void)hipFree(s.d_out);
        if (s.done)
            (void)hipEventDestroy(s.done);
        if (s.h_status)
            (void)hipHostFree(s.h_status);
    }
    if (e->stream)
        (void)hipStreamDestroy(e->stream);
    e->geom.release();
    delete e;
}

static const int kSlots = 3;

/* The batch spec of one plain-path slot: its device input planes and
 * codestream buffer. */
static svt_jpeg_xs_amd_frame_spec_t slot_spec(const EncoderPriv* e, svt_jpeg_xs_encoder_api_t* api,
                                              const EncoderPriv::Slot& s) {
    svt_jpeg_xs_amd_frame_spec_t spec;
    memset(&spec, 0, sizeof(spec));
    spec.encoder = api;
    for (uint32_t c = 0; c < e->nplanes; c++) {
        spec.d_planes[c] = (const uint8_t*)s.d_in + e->plane_off[c];
        spec.strides[c] = e->plane_stride[c];
    }
    spec.d_out = s.d_out;
    return spec;
}

/* The finisher thread: takes the frames in send order, waits for each one's
 * GPU work, marks it finished and announces it through
 * callback_get_data_available, once per packet get_packet will hand out for
 * it (the whole frame, or the header and one packet per slice), as the
 * reference's final stage does when it posts output (FinalStageProcess.c:
 * 173-175, 208-209, 236-237). */
static void finisher_loop(EncoderPriv* e) {
    const uint32_t packets = e->slice_mode ? 1 + e->geom.host.g.slices : 1;
    for (;;) {
        int slot;
        {
            std::unique_lock<std::mutex> lk(e->mu);
            e->cv_watch.wait(lk, [&] { return e->stop || !e->watch.empty(); });
            if (e->watch.empty())
                return;
            slot = e->watch.front();
            e->watch.pop_front();
        }
        EncoderPriv::Slot& s = e->slots[slot];
        const hipError_t he = hipEventSynchronize(s.done);
        {
            std::lock_guard<std::mutex> lk(e->mu);
            s.failed = he != hipSuccess || *s.h_status != 0;
            s.finished = true;
        }
        e->cv_pending.notify_all();
        svt_jpeg_xs_encoder_api_t* api = e->api;
        if (api->callback_get_data_available)
            for (uint32_t k = 0; k < packets; k++)
                api->callback_get_data_available(api, api->callback_get_data_available_context);
    }
}

static SvtJxsErrorType_t encoder_create(svt_jpeg_xs_encoder_api_t* api) {
    EncParams p;
    p.width = api->source_width;
    p.height = api->source_height;
    p.depth = api->input_bit_depth;
    p.format = (uint32_t)api->colour_format;
    p.bpp_num = api->bpp_numerator;
    p.bpp_den = api->bpp_denominator;
    p.decomp_v = api->ndecomp_v;
    p.decomp_h = api->ndecomp_h;
    p.quantization = api->quantization;
    p.slice_height = api->slice_height;
    p.significance = api->coding_significance;
    p.rc_mode = api->rate_control_mode;
    p.vpred = api->coding_vertical_prediction_mode;
    p.signs = api->coding_signs_handling;
    /* EncHandle.c:242-278.  Full sign coding with the per-slice rate control
     * modes is refused: there the reference's per-band cache of two
     * truncations carries sizes from the slice-wide search (made with fast
     * sign coding) into the precincts' own searches, and that is not
     * restated. */
    if (api->rate_control_mode > 3 || api->coding_vertical_prediction_mode > 2 || api->coding_signs_handling > 2 ||
        (api->coding_signs_handling == 2 && api->rate_control_mode >= 2)) {
        if (api->verbose >= VERBOSE_ERRORS)
            fprintf(stderr, "svt-jpeg-xs-amd: unsupported rate control / vertical prediction / sign coding mode\n");
        return SvtJxsErrorBadParameter;
    }
    if (api->slice_packetization_mode > 1 || api->cpu_profile > 1)
        return SvtJxsErrorBadParameter;
    auto* e = new (std::nothrow) EncoderPriv;
    if (!e)
        return SvtJxsErrorInsufficientResources;
    e->verbose = api->verbose;
    e->slice_mode = api->slice_packetization_mode == 1;
    int32_t err = build_encoder_geometry(p, e->geom.host, api->verbose >= VERBOSE_ERRORS);
    if (err) {
        delete e;
        return (SvtJxsErrorType_t)err;
    }
    if (api->slice_height >= api->source_height)
        api->slice_height = api->source_height;
    SvtJxsErrorType_t r = e->geom.upload();
    if (r) {
        encoder_destroy(e);
        return r;
    }
    const JxsGeom& g = e->geom.host.g;
    /* device input layout: tightly packed planes */
    const uint32_t elem = g.depth <= 8 ? 1 : 2;
    e->nplanes = g.packed ? 1 : g.nc;
    uint64_t off = 0;
    for (uint32_t c = 0; c < e->nplanes; c++) {
        const uint32_t w = g.packed ? 3 * g.w : g.c[c].w, h = g.packed ? g.h : g.c[c].h;
        e->plane_off[c] = off;
        e->plane_stride[c] = w;
        e->plane_rows[c] = h;
        e->plane_row_bytes[c] = w * elem;
        off = align256(off + (uint64_t)w * h * elem);
    }
    e->in_bytes = off;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        encoder_destroy(e);
        return SvtJxsErrorInsufficientResources;
    }
    api->private_ptr = e;
    e->slots.resize(kSlots);
    for (auto& s : e->slots) {
        if (hipMalloc(&s.d_in, e->in_bytes) != hipSuccess || hipMalloc(&s.d_out, ((size_t)g.lcod + 3) & ~(size_t)3) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&s.h_status, sizeof(uint32_t), 0) != hipSuccess) {
            api->private_ptr = nullptr;
            encoder_destroy(e);
            return SvtJxsErrorInsufficientResources;
        }
        const svt_jpeg_xs_amd_frame_spec_t spec = slot_spec(e, api, s);
        s.plan = plan_build(&spec, 1);
        if (!s.plan) {
            api->private_ptr = nullptr;
            encoder_destroy(e);
            return SvtJxsErrorInsufficientResources;
        }
    }
    e->api = api;
    try {
        e->finisher = std::thread(finisher_loop, e);
    } catch (...) {
        api->private_ptr = nullptr;
        encoder_destroy(e);
        return SvtJxsErrorInsufficientResources;
    }
    return SvtJxsErrorNone;
}

}  // namespace jxs

using namespace jxs;

extern "C" {

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_load_default_parameters(uint64_t version_api_major,
                                                                         uint64_t version_api_minor,
                                                                         svt_jpeg_xs_encoder_api_t* enc_api) {
    if (!version_ok(version_api_major, version_api_minor))
        return SvtJxsErrorInvalidApiVersion;
    if (!enc_api)
        return SvtJxsErrorBadParameter;
    memset(enc_api, 0, sizeof(*enc_api));
    enc_api->colour_format = COLOUR_FORMAT_INVALID;
    enc_api->bpp_denominator = 1;
    enc_api->ndecomp_v = 2;
    enc_api->ndecomp_h = 5;
    enc_api->quantization = 0;
    enc_api->slice_height = 16;
    enc_api->use_cpu_flags = CPU_FLAGS_ALL;
    enc_api->verbose = VERBOSE_SYSTEM_INFO;
    enc_api->rate_control_mode = 1;
    enc_api->coding_significance = 1;
    return SvtJxsErrorNone;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_get_image_config(uint64_t version_api_major, uint64_t version_api_minor,
                                                                  svt_jpeg_xs_encoder_api_t* enc_api,
                                                                  svt_jpeg_xs_image_config_t* cfg, uint32_t* out_bytes) {
    if (!version_ok(version_api_major, version_api_minor))
        return SvtJxsErrorInvalidApiVersion;
    if (!enc_api || !cfg)
        return SvtJxsErrorBadParameter;
    cfg->width = enc_api->source_width;
    cfg->height = enc_api->source_height;
    cfg->bit_depth = enc_api->input_bit_depth;
    cfg->format = enc_api->colour_format;
    if (cfg->bit_depth < 8 || cfg->bit_depth > 14)
        return SvtJxsErrorBadParameter;
    uint32_t nc = 0, sx[4] = {1, 1, 1, 1}, sy[4] = {1, 1, 1, 1};
    switch (cfg->format) {
    case COLOUR_FORMAT_PLANAR_YUV422: nc = 3; sx[1] = sx[2] = 2; break;
    case COLOUR_FORMAT_PLANAR_YUV420: nc = 3; sx[1] = sx[2] = sy[1] = sy[2] = 2; break;
    case COLOUR_FORMAT_PLANAR_YUV444_OR_RGB:
    case COLOUR_FORMAT_PACKED_YUV444_OR_RGB: nc = 3; break;
    case COLOUR_FORMAT_PLANAR_4_COMPONENTS: nc = 4; break;
    case COLOUR_FORMAT_GRAY: nc = 1; break;
    default: return SvtJxsErrorBadParameter;
    }
    const uint32_t pix = cfg->bit_depth <= 8 ? 1 : 2;
    if (cfg->format == COLOUR_FORMAT_PACKED_YUV444_OR_RGB) {
        const uint64_t b = (uint64_t)enc_api->source_width * enc_api->source_height * 3 * pix;
        if (b > UINT32_MAX)
            return SvtJxsErrorBadParameter;
        cfg->components_num = 1;
        cfg->components[0].width = enc_api->source_width;
        cfg->components[0].height = enc_api->source_height;
        cfg->components[0].byte_size = (uint32_t)b;
    } else {
        cfg->components_num = (uint8_t)nc;
        for (uint32_t c = 0; c < nc; c++) {
            cfg->components[c].width = cfg->width >> (sx[c] - 1);
            cfg->components[c].height = cfg->height >> (sy[c] - 1);
            const uint64_t b = (uint64_t)cfg->components[c].width * cfg->components[c].height * pix;
            if (b > UINT32_MAX)
                return SvtJxsErrorBadParameter;
            cfg->components[c].byte_size = (uint32_t)b;
        }
    }
    if (out_bytes) {
        if (enc_api->bpp_denominator == 0)
            return SvtJxsErrorBadParameter;
        *out_bytes = (uint32_t)(((uint64_t)enc_api->source_width * enc_api->source_height * enc_api->bpp_numerator /
                                     enc_api->bpp_denominator +
                                 7) /
                                8);
    }
    return SvtJxsErrorNone;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_init(uint64_t version_api_major, uint64_t version_api_minor,
                                                      svt_jpeg_xs_encoder_api_t* enc_api) {
    if (!version_ok(version_api_major, version_api_minor))
        return SvtJxsErrorInvalidApiVersion;
    if (!enc_api)
        return SvtJxsErrorBadParameter;
    enc_api->private_ptr = nullptr;
    SvtJxsErrorType_t r = encoder_create(enc_api);
    if (r)
        return r;
    EncoderPriv* e = priv_of(enc_api);
    if (enc_api->callback_send_data_available)
        enc_api->callback_send_data_available(enc_api, enc_api->callback_send_data_available_context);
    (void)e;
    return SvtJxsErrorNone;
}

PREFIX_API void svt_jpeg_xs_encoder_close(svt_jpeg_xs_encoder_api_t* enc_api) {
    if (!enc_api || !enc_api->private_ptr)
        return;
    encoder_destroy(priv_of(enc_api));
    enc_api->private_ptr = nullptr;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_send_picture(svt_jpeg_xs_encoder_api_t* enc_api,
                                                              svt_jpeg_xs_frame_t* in, uint8_t blocking_flag) {
    EncoderPriv* e = priv_of(enc_api);
    if (!e || !in)
        return SvtJxsErrorBadParameter;
    const JxsGeom& g = e->geom.host.g;
    if (in->bitstream.allocation_size < g.lcod || !in->bitstream.buffer)
        return SvtJxsErrorBadParameter;
    const uint32_t pix = g.depth <= 8 ? 1 : 2;
    /* plane checks as EncHandle.c:969-990 (all components are validated,
     * also for the packed format) */
    for (uint32_t c = 0; c < g.nc; c++) {
        if (!in->image.data_yuv[c])
            return SvtJxsErrorBadParameter;
        const uint64_t min_size =
            (uint64_t)in->image.stride[c] * pix * (g.c[c].h - 1) + (uint64_t)g.c[c].w * pix;
        if (in->image.alloc_size[c] < min_size)
            return SvtJxsErrorBadParameter;
    }
    int slot = -1;
    {
        std::unique_lock<std::mutex> lk(e->mu);
        for (;;) {
            for (int i = 0; i < (int)e->slots.size(); i++)
                if (!e->slots[i].busy) {
                    slot = i;
                    break;
                }
            if (slot >= 0 || !blocking_flag)
                break;
            e->cv_free.wait(lk);
        }
        if (slot < 0)
            return SvtJxsErrorNoErrorEmptyQueue;
        e->slots[slot].busy = true;
    }
    EncoderPriv::Slot& s = e->slots[slot];
    s.frame = *in;
    hipError_t he = hipSuccess;
    for (uint32_t c = 0; c < e->nplanes && he == hipSuccess; c++) {
        const uint32_t src_pitch = in->image.stride[c] * pix;
        he = hipMemcpy2DAsync((uint8_t*)s.d_in + e->plane_off[c], e->plane_row_bytes[c], in->image.data_yuv[c], src_pitch,
                              e->plane_row_bytes[c], e->plane_rows[c], hipMemcpyHostToDevice, e->stream);
    }
    if (he == hipSuccess)
        he = batch_launch(s.plan->L, e->stream, nullptr);
    if (he == hipSuccess)
        he = hipMemcpyAsync(in->bitstream.buffer, s.d_out, g.lcod, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.h_status, s.plan->L.status, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess)
        he = hipEventRecord(s.done, e->stream);
    std::lock_guard<std::mutex> lk(e->mu);
    if (he != hipSuccess) {
        s.busy = false;
        e->cv_free.notify_one();
        return SvtJxsErrorUndefined;
    }
    s.finished = false;
    e->pending.push_back(slot);
    e->watch.push_back(slot);
    e->cv_watch.notify_one();
    return SvtJxsErrorNone;
}

/* Takes the oldest frame off the pending queue once the finisher has seen
 * its GPU work complete (waiting for that when blocking); -1 when none is
 * ready. */
static int next_finished(EncoderPriv* e, bool blocking, bool& failed) {
    std::unique_lock<std::mutex> lk(e->mu);
    auto ready = [&] { return !e->pending.empty() && e->slots[e->pending.front()].finished; };
    if (!ready()) {
        if (!blocking)
            return -1;
        e->cv_pending.wait(lk, ready);
    }
    const int slot = e->pending.front();
    e->pending.pop_front();
    failed = e->slots[slot].failed;
    return slot;
}

/* Returns a slot to the free pool. */
static void release_slot(EncoderPriv* e, svt_jpeg_xs_encoder_api_t* enc_api, int slot) {
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->slots[slot].busy = false;
    }
    e->cv_free.notify_one();
    if (enc_api->callback_send_data_available)
        enc_api->callback_send_data_available(enc_api, enc_api->callback_send_data_available_context);
}

/* Whole frames (EncHandle.c:1030-1060, FinalStageProcess.c:216-239), or with
 * slice packetization the frame header, then one packet per slice into
 * the same buffer, the last one marked last and releasable
 * (FinalStageProcess.c:126-213).  Callers hand packets out in order from one
 * thread (as the reference's single output queue does). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_get_packet(svt_jpeg_xs_encoder_api_t* enc_api,
                                                            svt_jpeg_xs_frame_t* out, uint8_t blocking_flag) {
    EncoderPriv* e = priv_of(enc_api);
    if (!e || !out)
        return SvtJxsErrorBadParameter;
    const JxsGeom& g = e->geom.host.g;
    if (!e->slice_mode) {
        bool failed = false;
        const int slot = next_finished(e, blocking_flag != 0, failed);
        if (slot < 0) {
            out->user_prv_ctx_ptr = nullptr;
            return SvtJxsErrorNoErrorEmptyQueue;
        }
        *out = e->slots[slot].frame;
        out->bitstream.used_size = g.lcod;
        out->bitstream.ready_to_release = 1;
        out->bitstream.last_packet_in_frame = 1;
        out->image.ready_to_release = 1;
        release_slot(e, enc_api, slot);
        return failed ? SvtJxsErrorEncodeFrameError : SvtJxsErrorNone;
    }
    if (e->rel_slot < 0) {
        bool failed = false;
        const int slot = next_finished(e, blocking_flag != 0, failed);
        if (slot < 0) {
            out->user_prv_ctx_ptr = nullptr;
            return SvtJxsErrorNoErrorEmptyQueue;
        }
        e->rel_slot = slot;
        e->rel_pkt = 0;
        e->rel_off = 0;
        e->rel_failed = failed;
    }
    const int slot = e->rel_slot;
    *out = e->slots[slot].frame;
    out->bitstream.last_packet_in_frame = 0;
    out->bitstream.ready_to_release = 0;
    out->image.ready_to_release = 0;
    bool last = false;
    if (e->rel_pkt == 0) {
        out->bitstream.used_size = g.hdr_len;
        e->rel_off = g.hdr_len;
    } else {
        const uint32_t k = e->rel_pkt - 1;
        out->bitstream.buffer += e->rel_off;
        out->bitstream.used_size = e->geom.host.slice_size[k];
        e->rel_off += out->bitstream.used_size;
        last = k + 1 == g.slices;
    }
    e->rel_pkt++;
    const bool failed = e->rel_failed;
    if (last) {
        out->bitstream.last_packet_in_frame = 1;
        out->bitstream.ready_to_release = 1;
        out->image.ready_to_release = 1;
        e->rel_slot = -1;
        release_slot(e, enc_api, slot);
    }
    return failed ? SvtJxsErrorEncodeFrameError : SvtJxsErrorNone;
}

/* ---------------- extension entry points ---------------- */

PREFIX_API svt_jpeg_xs_amd_plan_t* svt_jpeg_xs_amd_plan_create(const svt_jpeg_xs_amd_frame_spec_t* frames, uint32_t count) {
    return plan_build(frames, count);
}

PREFIX_API uint64_t svt_jpeg_xs_amd_plan_footprint(const svt_jpeg_xs_amd_frame_spec_t* frames, uint32_t count) {
    if (!frames)
        return 0;
    BatchSizer sz;
    for (uint32_t i = 0; i < count; i++) {
        const DevGeom* dg = encoder_geom(frames[i].encoder);
        if (dg)
            sz.add(*dg);
    }
    return sz.n ? sz.bytes() : 0;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_run(svt_jpeg_xs_amd_plan_t* plan, void* stream, uint32_t use_graph) {
    if (!plan)
        return SvtJxsErrorBadParameter;
    hipStream_t st = (hipStream_t)stream;
    if (use_graph && st) {
        if (plan->graph_exec && plan->graph_stream != st) {
            (void)hipGraphExecDestroy(plan->graph_exec);
            (void)hipGraphDestroy(plan->graph);
            plan->graph_exec = nullptr;
            plan->graph = nullptr;
        }
        if (!plan->graph_exec) {
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return SvtJxsErrorUndefined;
            hipError_t le = batch_launch(plan->L, st, nullptr);
            hipError_t ee = hipStreamEndCapture(st, &plan->graph);
            if (le != hipSuccess || ee != hipSuccess)
                return SvtJxsErrorUndefined;
            if (hipGraphInstantiate(&plan->graph_exec, plan->graph, nullptr, nullptr, 0) != hipSuccess)
                return SvtJxsErrorUndefined;
            plan->graph_stream = st;
        }
        return hipGraphLaunch(plan->graph_exec, st) == hipSuccess ? SvtJxsErrorNone : SvtJxsErrorUndefined;
    }
    return batch_launch(plan->L, st, nullptr) == hipSuccess ? SvtJxsErrorNone : SvtJxsErrorUndefined;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_run_timed(svt_jpeg_xs_amd_plan_t* plan, void* stream, float* stage_ms) {
    if (!plan || !stage_ms)
        return SvtJxsErrorBadParameter;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t ev[SVT_JPEG_XS_AMD_STAGES + 1];
    for (int i = 0; i <= SVT_JPEG_XS_AMD_STAGES; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess)
            return SvtJxsErrorInsufficientResources;
    hipError_t e = batch_launch(plan->L, st, ev);
    if (e == hipSuccess)
        e = hipEventSynchronize(ev[SVT_JPEG_XS_AMD_STAGES]);
    for (int i = 0; i < SVT_JPEG_XS_AMD_STAGES && e == hipSuccess; i++)
        (void)hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
    for (int i = 0; i <= SVT_JPEG_XS_AMD_STAGES; i++)
        (void)hipEventDestroy(ev[i]);
    return e == hipSuccess ? SvtJxsErrorNone : SvtJxsErrorUndefined;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_status(svt_jpeg_xs_amd_plan_t* plan, uint32_t* frame_errors) {
    if (!plan)
        return SvtJxsErrorBadParameter;
    std::vector<uint32_t> st(plan->nfr);
    if (hipMemcpy(st.data(), plan->L.status, plan->nfr * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return SvtJxsErrorUndefined;
    bool any = false;
    for (uint32_t i = 0; i < plan->nfr; i++) {
        if (frame_errors)
            frame_errors[i] = st[i];
        any |= st[i] != 0;
    }
    return any ? SvtJxsErrorEncodeFrameError : SvtJxsErrorNone;
}

PREFIX_API void svt_jpeg_xs_amd_plan_destroy(svt_jpeg_xs_amd_plan_t* plan) { plan_free(plan); }

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_plan_precinct_offsets(svt_jpeg_xs_amd_plan_t* plan, uint32_t frame,
                                                                   const uint32_t** d_offsets, uint32_t* stride) {
    if (!plan || frame >= plan->nfr || !d_offsets || !stride)
        return SvtJxsErrorBadParameter;
    JxsFrameDesc F;
    if (hipMemcpy(&F, plan->L.frames + frame, sizeof(F), hipMemcpyDeviceToHost) != hipSuccess)
        return SvtJxsErrorUndefined;
    *d_offsets = &F.prec->offset;
    *stride = sizeof(JxsPrec) / sizeof(uint32_t);
    return SvtJxsErrorNone;
}

PREFIX_API int32_t svt_jpeg_xs_amd_autotune(svt_jpeg_xs_encoder_api_t* enc_api, void* stream,
                                            svt_jpeg_xs_amd_tile_t* out, uint32_t max_out) {
    EncoderPriv* e = priv_of(enc_api);
    if (!e)
        return -1;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        for (const auto& s : e->slots)
            if (s.busy)
                return -1; /* frames in flight */
    }
    hipStream_t st = (hipStream_t)stream;
    /* the job shapes of one frame (placeholder addresses: only shapes matter) */
    static const uint8_t placeholder[256] = {};
    const void* planes[4] = {placeholder, placeholder, placeholder, placeholder};
    const uint32_t strides[4] = {e->geom.host.g.w, e->geom.host.g.w, e->geom.host.g.w, e->geom.host.g.w};
    FwdJobs jobs;
    fwd_jobs(e->geom.host.g, planes, strides, nullptr, nullptr, nullptr, nullptr, nullptr, jobs);
    std::vector<JxsDwtJob> shapes;
    for (int sg = 0; sg < 2; sg++)
        for (uint32_t ji = 0; ji < jobs.n[sg]; ji++) {
            const JxsDwtJob& j = jobs.dwt[sg][ji];
            bool seen = false;
            for (const auto& k : shapes)
                seen |= k.kind == j.kind && k.step == j.step && k.vert == j.vert && k.lev_l == j.lev_l && k.w == j.w;
            if (!seen)
                shapes.push_back(j);
        }
    EncoderPriv::Slot& s0 = e->slots[0];
    {
        /* a pseudo-random frame within the bit depth (xorshift32): every
         * candidate must give the built-in widths' codestream for it */
        const JxsGeom& g = e->geom.host.g;
        const uint32_t mask = (1u << g.depth) - 1, elem = g.depth <= 8 ? 1 : 2;
        std::vector<uint8_t> host(e->in_bytes);
        uint32_t x = 0x9E3779B9u;
        for (uint64_t i = 0; i + elem <= host.size(); i += elem) {
            x ^= x << 13;
            x ^= x >> 17;
            x ^= x << 5;
            const uint32_t v = (x >> 7) & mask;
            host[i] = (uint8_t)v;
            if (elem == 2)
                host[i + 1] = (uint8_t)(v >> 8);
        }
        if (hipMemcpy(s0.d_in, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
            return -1;
    }
    const uint32_t lcod = e->geom.host.g.lcod;
    std::vector<uint8_t> ref_cs, cs(lcod);
    const svt_jpeg_xs_amd_frame_spec_t spec = slot_spec(e, enc_api, s0);
    hipEvent_t ev[SVT_JPEG_XS_AMD_STAGES + 1];
    for (int i = 0; i <= SVT_JPEG_XS_AMD_STAGES; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess)
            return -1;
    /* the transform's time (both vertical levels) of one frame, best of a few
     * runs, with the tile widths now in the table */
    auto time_transform = [&]() -> float {
        svt_jpeg_xs_amd_plan_t* p = plan_build(&spec, 1);
        if (!p)
            return -1.0f;
        float best = -1.0f;
        for (int rep = 0; rep < 6; rep++) {
            float ms = 0.0f;
            if (batch_launch(p->L, st, ev) != hipSuccess || hipEventSynchronize(ev[SVT_JPEG_XS_AMD_STAGES]) != hipSuccess ||
                hipEventElapsedTime(&ms, ev[0], ev[2]) != hipSuccess) {
                best = -1.0f;
                break;
            }
            if (rep > 0 && (best < 0.0f || ms < best))
                best = ms;
        }
        /* the codestream must not depend on the tile width */
        if (best >= 0.0f && hipMemcpy(cs.data(), s0.d_out, lcod, hipMemcpyDeviceToHost) != hipSuccess)
            best = -1.0f;
        if (best >= 0.0f) {
            if (ref_cs.empty())
                ref_cs = cs;
            else if (cs != ref_cs)
                best = -1.0f;
        }
        plan_free(p);
        return best;
    };
    /* the built-in rule's codestream first */
    std::vector<std::pair<TileKey, uint32_t>> saved;
    for (const auto& j : shapes) {
        saved.push_back({tile_key(j), tile_get(tile_key(j))});
        tile_set(tile_key(j), 0);
    }
    bool ok = time_transform() >= 0.0f;
    for (const auto& kv : saved)
        tile_set(kv.first, kv.second);
    int32_t tuned = 0;
    for (const auto& j : shapes) {
        if (!ok)
            break;
        const TileKey key = tile_key(j);
        const uint32_t before = tile_get(key);
        uint32_t best_c = 0;
        float best_t = 0.0f;
        for (uint32_t c : {8u, 16u, 32u}) {
            if (!tile_legal(j.lev_l, j.w, c))
                continue;
            tile_set(key, c);
            const float t = time_transform();
            if (t < 0.0f) {
                ok = false;
                break;
            }
            if (!best_c || t < best_t) {
                best_c = c;
                best_t = t;
            }
        }
        tile_set(key, ok ? best_c : before);
        if (!ok)
            break;
        if (out && (uint32_t)tuned < max_out) {
            out[tuned].width = j.w;
            out[tuned].levels = j.lev_l;
            out[tuned].vertical = j.vert;
            out[tuned].chunk = best_c;
            out[tuned].frame_us = best_t * 1000.0f;
            out[tuned].stage = 0;
        }
        tuned++;
    }
    /* waves per precinct of the rate table kernel (rate control modes 0 and
     * 1): the rate stage's time of one frame for 4, 2 and 1 waves */
    const JxsGeom& rg = e->geom.host.g;
    if (ok && !geom_ordered(rg) && rg.rc_mode < 2) {
        const uint32_t before = rate_waves_get(rg);
        uint32_t best_w = 0;
        float best_t = 0.0f;
        for (uint32_t nw : {4u, 2u, 1u}) {
            rate_waves_set(rg, nw);
            svt_jpeg_xs_amd_plan_t* p = plan_build(&spec, 1);
            float best = -1.0f;
            for (int rep = 0; p && rep < 6; rep++) {
                float ms = 0.0f;
                if (batch_launch(p->L, st, ev) != hipSuccess ||
                    hipEventSynchronize(ev[SVT_JPEG_XS_AMD_STAGES]) != hipSuccess ||
                    hipEventElapsedTime(&ms, ev[2], ev[3]) != hipSuccess) {
                    best = -1.0f;
                    break;
                }
                if (rep > 0 && (best < 0.0f || ms < best))
                    best = ms;
            }
            if (p)
                plan_free(p);
            /* the codestream must not depend on the choice */
            if (best >= 0.0f && (hipMemcpy(cs.data(), s0.d_out, lcod, hipMemcpyDeviceToHost) != hipSuccess ||
                                 (!ref_cs.empty() && cs != ref_cs)))
                best = -1.0f;
            if (best < 0.0f) {
                ok = false;
                break;
            }
            if (!best_w || best < best_t) {
                best_w = nw;
                best_t = best;
            }
        }
        rate_waves_set(rg, ok ? best_w : before);
        if (ok && out && (uint32_t)tuned < max_out) {
            out[tuned].width = rg.w;
            out[tuned].levels = 0;
            out[tuned].vertical = 0;
            out[tuned].chunk = best_w;
            out[tuned].frame_us = best_t * 1000.0f;
            out[tuned].stage = 1;
        }
        if (ok)
            tuned++;
    }
    for (int i = 0; i <= SVT_JPEG_XS_AMD_STAGES; i++)
        (void)hipEventDestroy(ev[i]);
    /* the encoder's own plans take the chosen widths */
    for (auto& s : e->slots) {
        const svt_jpeg_xs_amd_frame_spec_t sp = slot_spec(e, enc_api, s);
        svt_jpeg_xs_amd_plan_t* p = plan_build(&sp, 1);
        if (!p) {
            ok = false;
            continue;
        }
        plan_free(s.plan);
        s.plan = p;
    }
    return ok ? tuned : -1;
}

/* ---------------- encode without allocation ---------------- */

}  // extern "C"

namespace jxs {

/* The work block of svt_jpeg_xs_amd_encode_noalloc: one frame's descriptors,
 * job lists, packing units and work buffers (the block a one-frame plan
 * allocates), in a file-scope device array that every GPU the library runs
 * on holds from module load until the process ends.  64 MiB holds a 1080p
 * 4:4:4 12-bit frame (46 MiB) or a 1080p 4:2:2 frame (32 MiB); 2160p frames
 * need more and are refused.  Host code never uses the array's name as an
 * address: it uses the per-device address from hipGetSymbolAddress. */
#define JXS_NOALLOC_BYTES (64ull << 20)
static __device__ __align__(256) uint8_t scratch[JXS_NOALLOC_BYTES];

/* Host staging of one frame's descriptor image (on the stack of a call). */
static const uint32_t kNoallocImg = 32768;
static_assert(256 + 6 * (((JXS_MAXC * sizeof(JxsDwtJob) + 255) & ~(size_t)255) + 256) <= kNoallocImg,
              "one frame's descriptor image must fit the staging buffer");

/* Kernel-argument upload: the descriptor image travels in the arguments of
 * small copy launches, which the runtime captures at launch, so the host
 * staging may be reused as soon as the launch is enqueued. */
#define JXS_UPLOAD_WORDS 128 /* 16-byte words per launch */
struct JxsUploadChunk {
    uint4 w[JXS_UPLOAD_WORDS];
};

__global__ void __launch_bounds__(JXS_UPLOAD_WORDS) k_upload(uint4* dst, uint32_t nw, JxsUploadChunk c) {
    const uint32_t i = threadIdx.x;
    if (i < nw)
        dst[i] = c.w[i];
}

static hipError_t upload_by_args(uint8_t* dst, const uint8_t* src, uint32_t bytes, hipStream_t st) {
    const uint32_t words = (bytes + 15) / 16;
    for (uint32_t w0 = 0; w0 < words; w0 += JXS_UPLOAD_WORDS) {
        JxsUploadChunk c;
        const uint32_t n = words - w0 < JXS_UPLOAD_WORDS ? words - w0 : JXS_UPLOAD_WORDS;
        memset(&c, 0, sizeof(c));
        memcpy(&c, src + (size_t)w0 * 16, (size_t)n * 16);
        hipLaunchKernelGGL(k_upload, dim3(1), dim3(JXS_UPLOAD_WORDS), 0, st, (uint4*)(dst + (size_t)w0 * 16), n, c);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

#define JXS_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[JXS_MAX_DEVICES];

/* The scratch array's device address on the current device, looked up once
 * per device (with that device current); null on failure. */
static uint8_t* scratch_address(int dev) {
    if (dev < 0 || dev >= JXS_MAX_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        uint8_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fprintf(stderr, "svt-jpeg-xs-amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace jxs

extern "C" {

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_encode_noalloc(const svt_jpeg_xs_amd_frame_spec_t* frame, void* stream,
                                                            uint32_t* d_status) {
    if (!frame || !d_status || ((uintptr_t)d_status & 3) != 0)
        return SvtJxsErrorBadParameter;
    const DevGeom* dg = spec_geom(*frame);
    if (!dg)
        return SvtJxsErrorBadParameter;
    BatchSizer sz;
    sz.add(*dg);
    if (sz.bytes() > JXS_NOALLOC_BYTES || sz.upload_bytes() > kNoallocImg) {
        fprintf(stderr,
                "svt-jpeg-xs-amd: encode_noalloc: the frame needs %llu work bytes, more than the %llu-byte scratch\n",
                (unsigned long long)sz.bytes(), (unsigned long long)JXS_NOALLOC_BYTES);
        return SvtJxsErrorBadParameter;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return SvtJxsErrorUndefined;
    uint8_t* base = scratch_address(dev);
    if (!base)
        return SvtJxsErrorInsufficientResources;
    alignas(16) uint8_t img[kNoallocImg];
    BatchLaunch L;
    if (!batch_place_into(frame, 1, base, L, img, sz.upload_bytes()))
        return SvtJxsErrorBadParameter;
    /* the frame's error word is the caller's (cleared by the first launch) */
    JxsFrameDesc F;
    memcpy(&F, img, sizeof(F));
    F.status = d_status;
    memcpy(img, &F, sizeof(F));
    L.status = d_status;
    hipStream_t st = (hipStream_t)stream;
    if (upload_by_args(base, img, (uint32_t)sz.upload_bytes(), st) != hipSuccess ||
        batch_launch(L, st, nullptr) != hipSuccess)
        return SvtJxsErrorUndefined;
    return SvtJxsErrorNone;
}

PREFIX_API const char* svt_jpeg_xs_amd_version(void) { return "svt-jpeg-xs-amd 0.1 (gfx950, API 0.10)"; }

}  // extern "C"
