/*
 * jxs_decoder.cpp -- decoder runtime: header parsing and frame sizing on the
 * host, decode plans (unpack -> inverse transform -> output scaling on the
 * GPU), and the C-ABI entry points of SvtJpegxsDec.h and the device-resident
 * decode entry points of svt_jpegxs_amd.h.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <vector>

#include "SvtJpegxsDec.h"
#include "svt_jpegxs_amd.h"
#include "jxs_dec.h"
#include "jxs_host.h"

extern "C" {
hipError_t jxs_launch_dec_scan(const JxsDecFrame*, int, hipStream_t);
hipError_t jxs_launch_dec_packets(const JxsDecFrame*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_dec_unpack(const JxsDecFrame*, int, uint32_t, uint32_t, hipStream_t);
hipError_t jxs_launch_inv_rows(const JxsInvRowJob*, const uint32_t*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_inv_vert(const JxsInvVertJob*, const uint32_t*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_zero(uint32_t*, uint32_t, hipStream_t);
}

namespace jxs {

static inline uint64_t al256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

static bool dec_version_ok(uint64_t maj, uint64_t min) {
    return !(maj > SVT_JPEGXS_API_VER_MAJOR || (maj == SVT_JPEGXS_API_VER_MAJOR && min > SVT_JPEGXS_API_VER_MINOR));
}

/* Codestream size of the frame starting at cs: header, every slice and the
 * EOC marker, following the precinct lengths (DecThreadInit.c:35-131). */
static int32_t frame_size(const uint8_t* cs, size_t len, HostGeom& hg, uint32_t* out_size,
                          std::vector<uint32_t>* prec_off) {
    uint32_t hdr_end = 0;
    int32_t e = parse_codestream_header(cs, len, hg, &hdr_end);
    if (e)
        return e;
    const JxsGeom& g = hg.g;
    uint64_t pos = hdr_end;
    if (prec_off)
        prec_off->assign(g.rows, 0);
    for (uint32_t s = 0; s < g.slices; s++) {
        if (pos + 6 > len)
            return SvtJxsErrorDecoderBitstreamTooShort;
        if (cs[pos] != 0xFF || cs[pos + 1] != 0x20 || ((cs[pos + 4] << 8) | cs[pos + 5]) != s)
            return SvtJxsErrorDecoderInvalidBitstream;
        pos += 6;
        const uint32_t first = s * g.pps, n = (s == g.slices - 1) ? g.rows - first : g.pps;
        for (uint32_t i = 0; i < n; i++) {
            if (pos + g.prec_hdr_bytes > len)
                return SvtJxsErrorDecoderBitstreamTooShort;
            if (prec_off)
                (*prec_off)[first + i] = (uint32_t)pos;
            const uint32_t lprc = ((uint32_t)cs[pos] << 16) | ((uint32_t)cs[pos + 1] << 8) | cs[pos + 2];
            pos += g.prec_hdr_bytes + (uint64_t)lprc;
        }
    }
    if (pos + 2 > len)
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (cs[pos] != 0xFF || cs[pos + 1] != 0x11)
        return SvtJxsErrorDecoderInvalidBitstream;
    *out_size = (uint32_t)(pos + 2);
    return SvtJxsErrorNone;
}

/* Inverse-transform jobs of one frame (stage 0: rows of the second vertical
 * level, 1: its columns, 2: rows of the first level or of components without
 * vertical levels, 3: columns of the first level). */
struct InvJobs {
    std::vector<JxsInvRowJob> rows[2];
    std::vector<JxsInvVertJob> vert[2];
};

static void inv_jobs(const JxsGeom& g, int16_t* coef, int32_t* scratch, void* const planes[4], const uint32_t strides[4],
                     InvJobs& J) {
    int32_t* scr = scratch;
    for (uint32_t c = 0; c < g.nc; c++) {
        const JxsComp& cp = g.c[c];
        auto band = [&](uint32_t b) { return (const int16_t*)(coef + cp.b[b].coeff_off); };
        auto rowjob = [&](uint32_t w, uint32_t rows) {
            JxsInvRowJob j;
            memset(&j, 0, sizeof(j));
            j.w = w;
            j.rows = rows;
            j.depth = g.depth;
            return j;
        };
        const uint32_t nb = cp.nb;
        if (cp.dv == 0) {
            JxsInvRowJob j = rowjob(cp.w, cp.h);
            j.lf16 = band(0);
            j.lf_stride = cp.b[0].cs;
            j.levels = cp.dh;
            for (uint32_t k = 0; k < cp.dh; k++) {
                j.hf[k] = band(nb - 1 - k);
                j.hf_stride[k] = cp.b[nb - 1 - k].cs;
            }
            j.out = planes[c];
            j.out_stride = strides[c];
            J.rows[1].push_back(j);
            continue;
        }
        const uint32_t w = cp.w, h = cp.h, w1 = (w + 1) / 2, h1 = (h + 1) / 2, h2 = h / 2;
        int32_t* A = scr;
        int32_t* B = scr + (size_t)w * h;
        scr += (size_t)w * h + (size_t)w1 * h1;
        JxsInvRowJob lo = rowjob(w, h1);
        lo.hf[0] = band(nb - 3);
        lo.hf_stride[0] = cp.b[nb - 3].cs;
        if (cp.dv == 1) {
            lo.lf16 = band(0);
            lo.lf_stride = cp.b[0].cs;
            lo.levels = cp.dh;
            for (uint32_t k = 1; k < cp.dh; k++) {
                lo.hf[k] = band(cp.dh - k);
                lo.hf_stride[k] = cp.b[cp.dh - k].cs;
            }
        } else {
            const uint32_t g1 = (h1 + 1) / 2, g2 = h1 / 2;
            JxsInvRowJob l2 = rowjob(w1, g1);
            l2.lf16 = band(0);
            l2.lf_stride = cp.b[0].cs;
            l2.levels = 1 + (cp.dh - 2);
            l2.hf[0] = band(nb - 6);
            l2.hf_stride[0] = cp.b[nb - 6].cs;
            for (uint32_t k = 1; k < l2.levels; k++) {
                l2.hf[k] = band(cp.dh - 2 - (k - 1));
                l2.hf_stride[k] = cp.b[cp.dh - 2 - (k - 1)].cs;
            }
            l2.out32 = A;
            l2.out_stride = w1;
            J.rows[0].push_back(l2);
            if (g2 > 0) {
                JxsInvRowJob h2j = rowjob(w1, g2);
                h2j.lf16 = band(nb - 5);
                h2j.lf_stride = cp.b[nb - 5].cs;
                h2j.levels = 1;
                h2j.hf[0] = band(nb - 4);
                h2j.hf_stride[0] = cp.b[nb - 4].cs;
                h2j.out32 = A + (size_t)g1 * w1;
                h2j.out_stride = w1;
                J.rows[0].push_back(h2j);
            }
            JxsInvVertJob v2;
            memset(&v2, 0, sizeof(v2));
            v2.src = A;
            v2.out32 = B;
            v2.w = w1;
            v2.h = h1;
            v2.out_stride = w1;
            v2.depth = g.depth;
            J.vert[0].push_back(v2);
            lo.lf32 = B;
            lo.lf_stride = w1;
            lo.levels = 1;
        }
        lo.out32 = A;
        lo.out_stride = w;
        J.rows[1].push_back(lo);
        if (h2 > 0) {
            JxsInvRowJob hi = rowjob(w, h2);
            hi.lf16 = band(nb - 2);
            hi.lf_stride = cp.b[nb - 2].cs;
            hi.levels = 1;
            hi.hf[0] = band(nb - 1);
            hi.hf_stride[0] = cp.b[nb - 1].cs;
            hi.out32 = A + (size_t)h1 * w;
            hi.out_stride = w;
            J.rows[1].push_back(hi);
        }
        JxsInvVertJob v1;
        memset(&v1, 0, sizeof(v1));
        v1.src = A;
        v1.out = planes[c];
        v1.w = w;
        v1.h = h;
        v1.out_stride = strides[c];
        v1.depth = g.depth;
        J.vert[1].push_back(v1);
    }
}

struct DecGeom {
    HostGeom host;   /* the codestream's geometry (packets are walked with it) */
    JxsGeom out;     /* the geometry of the output: host.g, or with levels dropped for a proxy decode */
    uint32_t npk_dec = 0; /* packets of each precinct decoded */
    uint32_t hdr_end = 0;
    JxsGeom* d_geom = nullptr;
};

static uint32_t dec_pkt_stride(const JxsGeom& g) { return JXS_DPKT_HDR + JXS_DPKT_WORDS * g.npk; }

static uint64_t dec_work_bytes(const JxsGeom& g) {
    return al256(g.n_coeff * 2) + al256(g.n_i32 * 4) + al256((uint64_t)g.rows * 4) +
        al256((uint64_t)g.rows * dec_pkt_stride(g) * 4);
}

}  // namespace jxs

struct svt_jpeg_xs_amd_decode_plan {
    uint32_t nfr = 0;
    void* d_work = nullptr;
    uint32_t* d_status = nullptr;
    JxsDecFrame* d_frames = nullptr;
    JxsInvRowJob* d_rows[2] = {nullptr, nullptr};
    uint32_t* d_rbase[2] = {nullptr, nullptr};
    JxsInvVertJob* d_vert[2] = {nullptr, nullptr};
    uint32_t* d_vbase[2] = {nullptr, nullptr};
    int nrows[2] = {0, 0}, nvert[2] = {0, 0};
    uint32_t rblocks[2] = {0, 0}, vblocks[2] = {0, 0};
    uint32_t total_prec = 0, max_npk = 0;
};

namespace jxs {

struct DecoderPriv {
    DecGeom geom;
    uint32_t verbose = 0;
    uint8_t* d_cs = nullptr;
    uint32_t cs_cap = 0;
    uint32_t* d_prec = nullptr;
    void* d_planes = nullptr;
    size_t plane_off[4] = {0, 0, 0, 0};
    svt_jpeg_xs_amd_decode_plan_t* plan = nullptr;
    hipStream_t stream = nullptr;
    uint32_t* h_status = nullptr;
    std::vector<uint32_t> prec_off;
    /* queue of decoded frames (decode completes inside send) */
    struct Done {
        svt_jpeg_xs_frame_t frame;
        SvtJxsErrorType_t err;
    };
    std::deque<Done> done;
    std::mutex mu;
    std::condition_variable cv;
    bool eoc = false;
    /* packet mode accumulation */
    std::vector<uint8_t> pending_bytes;
};

template <typename T>
static hipError_t up(const std::vector<T>& v, T** dst) {
    *dst = nullptr;
    if (v.empty())
        return hipSuccess;
    hipError_t e = hipMalloc(dst, v.size() * sizeof(T));
    if (e != hipSuccess)
        return e;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

static void dplan_free(svt_jpeg_xs_amd_decode_plan_t* p) {
    if (!p)
        return;
    (void)hipFree(p->d_work);
    (void)hipFree(p->d_status);
    (void)hipFree(p->d_frames);
    for (int s = 0; s < 2; s++) {
        (void)hipFree(p->d_rows[s]);
        (void)hipFree(p->d_rbase[s]);
        (void)hipFree(p->d_vert[s]);
        (void)hipFree(p->d_vbase[s]);
    }
    delete p;
}

struct DecSpec {
    const DecGeom* geom;
    const uint8_t* d_cs;
    uint32_t cs_len;
    void* planes[4];
    uint32_t strides[4];
    const uint32_t* prec_in;
    uint32_t prec_stride;
};

static svt_jpeg_xs_amd_decode_plan_t* dplan_build(const std::vector<DecSpec>& specs) {
    auto* p = new (std::nothrow) svt_jpeg_xs_amd_decode_plan;
    if (!p || specs.empty()) {
        delete p;
        return nullptr;
    }
    p->nfr = (uint32_t)specs.size();
    uint64_t work = 0;
    for (auto& s : specs)
        work += dec_work_bytes(s.geom->host.g);
    if (hipMalloc(&p->d_work, work) != hipSuccess || hipMalloc(&p->d_status, p->nfr * 4) != hipSuccess) {
        dplan_free(p);
        return nullptr;
    }
    std::vector<JxsDecFrame> frames(p->nfr);
    InvJobs J;
    uint8_t* cur = (uint8_t*)p->d_work;
    uint32_t base = 0;
    for (uint32_t i = 0; i < p->nfr; i++) {
        const DecSpec& s = specs[i];
        const JxsGeom& g = s.geom->host.g;
        JxsDecFrame& F = frames[i];
        memset(&F, 0, sizeof(F));
        F.g = s.geom->d_geom;
        F.cs = s.d_cs;
        F.cs_len = s.cs_len;
        F.hdr_end = s.geom->hdr_end;
        F.coef = (int16_t*)cur;
        cur += al256(g.n_coeff * 2);
        int32_t* scratch = (int32_t*)cur;
        cur += al256(g.n_i32 * 4);
        F.prec_off = (uint32_t*)cur;
        cur += al256((uint64_t)g.rows * 4);
        F.pkt_stride = dec_pkt_stride(g);
        F.pkt = (uint32_t*)cur;
        cur += al256((uint64_t)g.rows * F.pkt_stride * 4);
        F.status = p->d_status + i;
        F.prec_base = base;
        F.npk_dec = s.geom->npk_dec;
        F.prec_in = s.prec_in;
        F.prec_in_stride = s.prec_stride ? s.prec_stride : 1;
        base += g.rows;
        p->max_npk = s.geom->npk_dec > p->max_npk ? s.geom->npk_dec : p->max_npk;
        InvJobs one;
        inv_jobs(s.geom->out, F.coef, scratch, s.planes, s.strides, one);
        for (int k = 0; k < 2; k++) {
            J.rows[k].insert(J.rows[k].end(), one.rows[k].begin(), one.rows[k].end());
            J.vert[k].insert(J.vert[k].end(), one.vert[k].begin(), one.vert[k].end());
        }
    }
    p->total_prec = base;
    bool ok = up(frames, &p->d_frames) == hipSuccess;
    for (int k = 0; k < 2 && ok; k++) {
        std::vector<uint32_t> rb, vb;
        uint32_t blk = 0;
        for (auto& r : J.rows[k]) {
            r.base = blk;
            rb.push_back(blk);
            blk += r.rows;
        }
        p->rblocks[k] = blk;
        blk = 0;
        for (auto& v : J.vert[k]) {
            const uint32_t nl = (v.h + 1) / 2;
            v.tiles_x = (v.w + 255) / 256;
            v.tiles = v.tiles_x * ((nl + 15) / 16);
            v.base = blk;
            vb.push_back(blk);
            blk += v.tiles;
        }
        p->vblocks[k] = blk;
        p->nrows[k] = (int)J.rows[k].size();
        p->nvert[k] = (int)J.vert[k].size();
        ok = up(J.rows[k], &p->d_rows[k]) == hipSuccess && up(rb, &p->d_rbase[k]) == hipSuccess &&
            up(J.vert[k], &p->d_vert[k]) == hipSuccess && up(vb, &p->d_vbase[k]) == hipSuccess;
    }
    if (!ok) {
        dplan_free(p);
        return nullptr;
    }
    return p;
}

static hipError_t dplan_launch(svt_jpeg_xs_amd_decode_plan_t* p, hipStream_t st) {
    hipError_t e;
    if ((e = jxs_launch_zero(p->d_status, p->nfr, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_scan(p->d_frames, (int)p->nfr, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_packets(p->d_frames, (int)p->nfr, p->total_prec, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_unpack(p->d_frames, (int)p->nfr, p->max_npk, p->total_prec, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_rows(p->d_rows[0], p->d_rbase[0], p->nrows[0], p->rblocks[0], st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_vert(p->d_vert[0], p->d_vbase[0], p->nvert[0], p->vblocks[0], st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_rows(p->d_rows[1], p->d_rbase[1], p->nrows[1], p->rblocks[1], st)) != hipSuccess)
        return e;
    return jxs_launch_inv_vert(p->d_vert[1], p->d_vbase[1], p->nvert[1], p->vblocks[1], st);
}

/* Proxy decoding (pi_update_proxy_mode, Pi.c:567-626): half or quarter
 * resolution drops one or two vertical and horizontal levels from the
 * picture and every component, whose first 2V' + H' + 1 bands stay as they
 * are; only the packets carrying them are decoded (one per precinct, or four
 * at half resolution with two vertical levels).  A codestream with fewer
 * levels than the mode drops is refused. */
static int32_t apply_proxy(DecGeom& dg, uint32_t proxy) {
    const JxsGeom& g = dg.host.g;
    dg.out = g;
    dg.npk_dec = g.npk;
    if (proxy == proxy_mode_full)
        return SvtJxsErrorNone;
    if (proxy >= proxy_mode_max || proxy > g.dv || proxy > g.dh)
        return SvtJxsErrorBadParameter;
    dg.npk_dec = proxy == proxy_mode_half && g.dv != 1 ? 4u : 1u;
    JxsGeom& o = dg.out;
    o.w = (g.w + (1u << proxy) - 1) >> proxy;
    o.h = (g.h + (1u << proxy) - 1) >> proxy;
    for (uint32_t c = 0; c < o.nc; c++) {
        JxsComp& cp = o.c[c];
        if (proxy > cp.dv || proxy > cp.dh)
            return SvtJxsErrorBadParameter;
        cp.w = (cp.w + (1u << proxy) - 1) >> proxy;
        cp.h = (cp.h + (1u << proxy) - 1) >> proxy;
        cp.dv -= proxy;
        cp.dh -= proxy;
        cp.nb = 2 * cp.dv + cp.dh + 1;
    }
    return SvtJxsErrorNone;
}

static int32_t geom_from_codestream(const uint8_t* cs, size_t len, DecGeom& dg, uint32_t proxy) {
    int32_t e = parse_codestream_header(cs, len, dg.host, &dg.hdr_end);
    if (e)
        return e;
    for (uint32_t c = 0; c < dg.host.g.nc; c++)
        if (dg.host.g.c[c].w > JXS_MAX_ROW)
            return SvtJxsErrorDecoderInvalidBitstream;
    if (dg.host.g.npk > JXS_NPK || dg.host.g.n_bl > JXS_NBL)
        return SvtJxsErrorDecoderInvalidBitstream;
    return apply_proxy(dg, proxy);
}

static void fill_image_config(const JxsGeom& g, svt_jpeg_xs_image_config_t* cfg) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->width = g.w;
    cfg->height = g.h;
    cfg->bit_depth = (uint8_t)g.depth;
    cfg->components_num = (uint8_t)g.nc;
    const uint32_t pix = g.depth <= 8 ? 1 : 2;
    bool s420 = g.nc == 3 && g.c[1].sy == 2, s422 = g.nc == 3 && g.c[1].sx == 2 && g.c[1].sy == 1;
    cfg->format = g.nc == 1 ? COLOUR_FORMAT_GRAY
        : g.nc == 4        ? COLOUR_FORMAT_PLANAR_4_COMPONENTS
        : s420             ? COLOUR_FORMAT_PLANAR_YUV420
        : s422             ? COLOUR_FORMAT_PLANAR_YUV422
                           : COLOUR_FORMAT_PLANAR_YUV444_OR_RGB;
    for (uint32_t c = 0; c < g.nc; c++) {
        cfg->components[c].width = g.c[c].w;
        cfg->components[c].height = g.c[c].h;
        cfg->components[c].byte_size = g.c[c].w * g.c[c].h * pix;
    }
}

static void decoder_destroy(DecoderPriv* d) {
    if (!d)
        return;
    if (d->stream)
        (void)hipStreamSynchronize(d->stream);
    dplan_free(d->plan);
    (void)hipFree(d->d_cs);
    (void)hipFree(d->d_prec);
    (void)hipFree(d->d_planes);
    (void)hipFree(d->geom.d_geom);
    if (d->h_status)
        (void)hipHostFree(d->h_status);
    if (d->stream)
        (void)hipStreamDestroy(d->stream);
    delete d;
}

/* Decodes one whole codestream (host memory) into the frame's image buffer. */
static SvtJxsErrorType_t decode_one(DecoderPriv* d, const uint8_t* cs, size_t len, svt_jpeg_xs_frame_t* frame) {
    HostGeom hg;
    uint32_t fsize = 0;
    int32_t e = frame_size(cs, len, hg, &fsize, &d->prec_off);
    if (e)
        return (SvtJxsErrorType_t)e;
    const JxsGeom& g = d->geom.host.g;
    const JxsGeom& og = d->geom.out; /* output planes */
    /* a different configuration needs a new decoder (DecHandle.c) */
    if (hg.g.w != g.w || hg.g.h != g.h || hg.g.nc != g.nc || hg.g.depth != g.depth || hg.g.dv != g.dv ||
        hg.g.dh != g.dh || hg.g.pps != g.pps || memcmp(hg.g.c, g.c, sizeof(g.c)) != 0)
        return SvtJxsErrorDecoderConfigChange;
    if (fsize > d->cs_cap)
        return SvtJxsErrorDecoderInvalidBitstream;
    const uint32_t pix = g.depth <= 8 ? 1 : 2;
    for (uint32_t c = 0; c < g.nc; c++) {
        if (!frame->image.data_yuv[c] || frame->image.stride[c] < og.c[c].w)
            return SvtJxsErrorDecoderInvalidPointer;
        const uint64_t need = (uint64_t)frame->image.stride[c] * pix * (og.c[c].h - 1) + (uint64_t)og.c[c].w * pix;
        if (frame->image.alloc_size[c] < need)
            return SvtJxsErrorBadParameter;
    }
    hipError_t he = hipMemcpyAsync(d->d_cs, cs, fsize, hipMemcpyHostToDevice, d->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(d->d_prec, d->prec_off.data(), g.rows * 4, hipMemcpyHostToDevice, d->stream);
    if (he == hipSuccess)
        he = dplan_launch(d->plan, d->stream);
    for (uint32_t c = 0; c < g.nc && he == hipSuccess; c++)
        he = hipMemcpy2DAsync(frame->image.data_yuv[c], (size_t)frame->image.stride[c] * pix,
                              (uint8_t*)d->d_planes + d->plane_off[c], (size_t)og.c[c].w * pix, (size_t)og.c[c].w * pix,
                              og.c[c].h, hipMemcpyDeviceToHost, d->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(d->h_status, d->plan->d_status, 4, hipMemcpyDeviceToHost, d->stream);
    if (he == hipSuccess)
        he = hipStreamSynchronize(d->stream);
    if (he != hipSuccess)
        return SvtJxsErrorDecoderInternal;
    return *d->h_status ? SvtJxsErrorDecoderInvalidBitstream : SvtJxsErrorNone;
}

}  // namespace jxs

using namespace jxs;

extern "C" {

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(
    const uint8_t* bitstream_buf, size_t bitstream_buf_size, svt_jpeg_xs_image_config_t* out_image_config,
    uint32_t* frame_size_out, uint32_t fast_search, proxy_mode_t proxy_mode) {
    if (!bitstream_buf || !frame_size_out)
        return SvtJxsErrorDecoderInvalidPointer;
    DecGeom dg;
    HostGeom& hg = dg.host;
    uint32_t hdr_end = 0;
    int32_t e = parse_codestream_header(bitstream_buf, bitstream_buf_size, hg, &hdr_end);
    if (e)
        return (SvtJxsErrorType_t)e;
    e = apply_proxy(dg, (uint32_t)proxy_mode);
    if (e)
        return (SvtJxsErrorType_t)e;
    if (fast_search && hg.g.lcod > 0) {
        /* constant-bit-rate codestreams carry their size in Lcod (PIH) */
        *frame_size_out = hg.g.lcod;
    } else {
        e = frame_size(bitstream_buf, bitstream_buf_size, hg, frame_size_out, nullptr);
        if (e)
            return (SvtJxsErrorType_t)e;
    }
    if (out_image_config)
        fill_image_config(dg.out, out_image_config);
    return SvtJxsErrorNone;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_single_frame_size(const uint8_t* bitstream_buf,
                                                                       size_t bitstream_buf_size,
                                                                       svt_jpeg_xs_image_config_t* out_image_config,
                                                                       uint32_t* frame_size_out, uint32_t fast_search) {
    return svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(bitstream_buf, bitstream_buf_size, out_image_config,
                                                                frame_size_out, fast_search, proxy_mode_full);
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_init(uint64_t version_api_major, uint64_t version_api_minor,
                                                      svt_jpeg_xs_decoder_api_t* dec_api, const uint8_t* bitstream_buf,
                                                      size_t codestream_size, svt_jpeg_xs_image_config_t* out_image_config) {
    if (!dec_version_ok(version_api_major, version_api_minor))
        return SvtJxsErrorInvalidApiVersion;
    if (!dec_api || !bitstream_buf)
        return SvtJxsErrorDecoderInvalidPointer;
    dec_api->private_ptr = nullptr;
    if ((uint32_t)dec_api->proxy_mode >= proxy_mode_max)
        return SvtJxsErrorBadParameter;
    auto* d = new (std::nothrow) DecoderPriv;
    if (!d)
        return SvtJxsErrorInsufficientResources;
    d->verbose = dec_api->verbose;
    int32_t e = geom_from_codestream(bitstream_buf, codestream_size, d->geom, (uint32_t)dec_api->proxy_mode);
    if (e) {
        delete d;
        return (SvtJxsErrorType_t)e;
    }
    const JxsGeom& g = d->geom.host.g;
    const JxsGeom& og = d->geom.out;
    const uint32_t pix = g.depth <= 8 ? 1 : 2;
    uint64_t planes = 0;
    for (uint32_t c = 0; c < g.nc; c++) {
        d->plane_off[c] = planes;
        planes = al256(planes + (uint64_t)og.c[c].w * og.c[c].h * pix);
    }
    /* the largest codestream of this geometry: every precinct at the 20-bit
     * Lprc limit is not realistic; a constant-rate stream carries Lcod */
    d->cs_cap = g.lcod ? g.lcod : (uint32_t)codestream_size;
    if (hipMalloc(&d->geom.d_geom, sizeof(JxsGeom)) != hipSuccess ||
        hipMemcpy(d->geom.d_geom, &g, sizeof(JxsGeom), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&d->d_cs, d->cs_cap) != hipSuccess || hipMalloc(&d->d_prec, g.rows * 4) != hipSuccess ||
        hipMalloc(&d->d_planes, planes) != hipSuccess ||
        hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&d->h_status, 4, 0) != hipSuccess) {
        decoder_destroy(d);
        return SvtJxsErrorInsufficientResources;
    }
    DecSpec s;
    memset(&s, 0, sizeof(s));
    s.geom = &d->geom;
    s.d_cs = d->d_cs;
    s.cs_len = d->cs_cap;
    for (uint32_t c = 0; c < g.nc; c++) {
        s.planes[c] = (uint8_t*)d->d_planes + d->plane_off[c];
        s.strides[c] = og.c[c].w;
    }
    s.prec_in = d->d_prec;
    s.prec_stride = 1;
    d->plan = dplan_build({s});
    if (!d->plan) {
        decoder_destroy(d);
        return SvtJxsErrorInsufficientResources;
    }
    if (out_image_config)
        fill_image_config(og, out_image_config);
    dec_api->private_ptr = d;
    if (dec_api->callback_send_data_available)
        dec_api->callback_send_data_available(dec_api, dec_api->callback_send_data_available_context);
    return SvtJxsErrorNone;
}

PREFIX_API void svt_jpeg_xs_decoder_close(svt_jpeg_xs_decoder_api_t* dec_api) {
    if (!dec_api || !dec_api->private_ptr)
        return;
    decoder_destroy((DecoderPriv*)dec_api->private_ptr);
    dec_api->private_ptr = nullptr;
}

static SvtJxsErrorType_t finish_frame(svt_jpeg_xs_decoder_api_t* api, DecoderPriv* d, svt_jpeg_xs_frame_t* in,
                                      SvtJxsErrorType_t err) {
    in->image.ready_to_release = 1;
    in->bitstream.ready_to_release = 1;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->done.push_back({*in, err});
    }
    d->cv.notify_one();
    if (api->callback_get_data_available)
        api->callback_get_data_available(api, api->callback_get_data_available_context);
    return SvtJxsErrorNone;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                            svt_jpeg_xs_frame_t* dec_input, uint8_t blocking_flag) {
    (void)blocking_flag;
    if (!dec_api || !dec_api->private_ptr || !dec_input || !dec_input->bitstream.buffer)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    const size_t len = dec_input->bitstream.used_size ? dec_input->bitstream.used_size
                                                      : dec_input->bitstream.allocation_size;
    SvtJxsErrorType_t e = decode_one(d, dec_input->bitstream.buffer, len, dec_input);
    if (e == SvtJxsErrorDecoderConfigChange || e == SvtJxsErrorDecoderInvalidPointer || e == SvtJxsErrorBadParameter)
        return e;
    return finish_frame(dec_api, d, dec_input, e);
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_packet(svt_jpeg_xs_decoder_api_t* dec_api,
                                                             svt_jpeg_xs_frame_t* dec_input, uint32_t* bytes_used) {
    if (!dec_api || !dec_api->private_ptr || !dec_input || !dec_input->bitstream.buffer || !bytes_used)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    const uint8_t* b = dec_input->bitstream.buffer;
    const size_t n = dec_input->bitstream.used_size;
    d->pending_bytes.insert(d->pending_bytes.end(), b, b + n);
    *bytes_used = (uint32_t)n;
    HostGeom hg;
    uint32_t fsize = 0;
    int32_t e = frame_size(d->pending_bytes.data(), d->pending_bytes.size(), hg, &fsize, nullptr);
    if (e == SvtJxsErrorDecoderBitstreamTooShort)
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (e) {
        d->pending_bytes.clear();
        return (SvtJxsErrorType_t)e;
    }
    std::vector<uint8_t> frame_bytes(d->pending_bytes.begin(), d->pending_bytes.begin() + fsize);
    d->pending_bytes.erase(d->pending_bytes.begin(), d->pending_bytes.begin() + fsize);
    SvtJxsErrorType_t r = decode_one(d, frame_bytes.data(), frame_bytes.size(), dec_input);
    if (r == SvtJxsErrorDecoderConfigChange || r == SvtJxsErrorDecoderInvalidPointer || r == SvtJxsErrorBadParameter)
        return r;
    return finish_frame(dec_api, d, dec_input, r);
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                           svt_jpeg_xs_frame_t* dec_output, uint8_t blocking_flag) {
    if (!dec_api || !dec_api->private_ptr || !dec_output)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    std::unique_lock<std::mutex> lk(d->mu);
    if (d->done.empty()) {
        if (d->eoc)
            return SvtJxsDecoderEndOfCodestream;
        if (!blocking_flag) {
            dec_output->user_prv_ctx_ptr = nullptr;
            return SvtJxsErrorNoErrorEmptyQueue;
        }
        d->cv.wait(lk, [&] { return !d->done.empty() || d->eoc; });
        if (d->done.empty())
            return SvtJxsDecoderEndOfCodestream;
    }
    DecoderPriv::Done r = d->done.front();
    d->done.pop_front();
    *dec_output = r.frame;
    return r.err;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_eoc(svt_jpeg_xs_decoder_api_t* dec_api) {
    if (!dec_api || !dec_api->private_ptr)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->eoc = true;
    }
    d->cv.notify_all();
    return SvtJxsErrorNone;
}

/* ---------------- device-resident decode ---------------- */

PREFIX_API svt_jpeg_xs_amd_decode_plan_t* svt_jpeg_xs_amd_decode_plan_create(const svt_jpeg_xs_amd_decode_spec_t* frames,
                                                                           uint32_t count) {
    if (!frames || !count)
        return nullptr;
    std::vector<DecSpec> specs(count);
    for (uint32_t i = 0; i < count; i++) {
        const svt_jpeg_xs_amd_decode_spec_t& f = frames[i];
        if (!f.decoder || !f.decoder->private_ptr || !f.d_codestream)
            return nullptr;
        DecSpec& s = specs[i];
        memset(&s, 0, sizeof(s));
        s.geom = &((DecoderPriv*)f.decoder->private_ptr)->geom;
        s.d_cs = f.d_codestream;
        s.cs_len = f.codestream_size;
        for (uint32_t c = 0; c < 4; c++) {
            s.planes[c] = f.d_planes[c];
            s.strides[c] = f.strides[c];
        }
        s.prec_in = f.d_precinct_offsets;
        s.prec_stride = f.precinct_offsets_stride;
    }
    return dplan_build(specs);
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_decode_plan_run(svt_jpeg_xs_amd_decode_plan_t* plan, void* stream) {
    if (!plan)
        return SvtJxsErrorBadParameter;
    return dplan_launch(plan, (hipStream_t)stream) == hipSuccess ? SvtJxsErrorNone : SvtJxsErrorDecoderInternal;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_decode_plan_status(svt_jpeg_xs_amd_decode_plan_t* plan, uint32_t* errors) {
    if (!plan)
        return SvtJxsErrorBadParameter;
    std::vector<uint32_t> st(plan->nfr);
    if (hipMemcpy(st.data(), plan->d_status, plan->nfr * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return SvtJxsErrorDecoderInternal;
    bool any = false;
    for (uint32_t i = 0; i < plan->nfr; i++) {
        if (errors)
            errors[i] = st[i];
        any |= st[i] != 0;
    }
    return any ? SvtJxsErrorDecoderInvalidBitstream : SvtJxsErrorNone;
}

PREFIX_API void svt_jpeg_xs_amd_decode_plan_destroy(svt_jpeg_xs_amd_decode_plan_t* plan) { dplan_free(plan); }

}  // extern "C"
