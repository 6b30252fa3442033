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
#include <algorithm>
#include <new>
#include <thread>
#include <vector>
#include <cstddef>

#include "SvtJpegxsDec.h"
#include "svt_jpegxs_amd.h"
#include "jxs_dec.h"
#include "jxs_host.h"

extern "C" {
hipError_t jxs_launch_dec_scan(const JxsDecFrame*, int, hipStream_t);
hipError_t jxs_launch_dec_packets(const JxsDecFrame*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_dec_unpack(const JxsDecFrame*, int, uint32_t, uint32_t, hipStream_t);
hipError_t jxs_launch_dec_slice(const JxsDecFrame*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_inv_rows(const JxsInvRowJob*, const uint32_t*, int, uint32_t, uint32_t, hipStream_t);
hipError_t jxs_launch_inv_vert(const JxsInvVertJob*, const uint32_t*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_inv_last(const JxsInvLastJob*, const uint32_t*, int, uint32_t, hipStream_t);
hipError_t jxs_launch_zero(uint32_t*, uint32_t, hipStream_t);
}

namespace jxs {

static inline uint64_t al256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

static bool dec_version_ok(uint64_t maj, uint64_t min) {
    return !(maj > SVT_JPEGXS_API_VER_MAJOR || (maj == SVT_JPEGXS_API_VER_MAJOR && min > SVT_JPEGXS_API_VER_MINOR));
}

/* Inverse-transform jobs of one frame: rows[0] / vert[0] the second
 * vertical level (rows, then columns, into an int32 plane); rows[1] the
 * components without vertical levels (into the output plane) and the
 * coarser horizontal levels of the first vertical level's low rows (into an
 * int32 plane at half width); last the first vertical level with the finest
 * horizontal level of its low and high rows. */
#define INV_SEG 64 /* row pairs per block of the last stage */
struct InvJobs {
    std::vector<JxsInvRowJob> rows[2];
    std::vector<JxsInvVertJob> vert[2];
    std::vector<JxsInvLastJob> last;
};

static void inv_jobs(const JxsGeom& g, uint32_t bw, uint32_t fq, int16_t* coef, int32_t* scratch, void* const planes[4],
                     const uint32_t strides[4], InvJobs& J) {
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
            j.fq = fq;
            j.bw = bw;
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
        JxsInvLastJob f;
        memset(&f, 0, sizeof(f));
        if (cp.dv == 1) {
            if (cp.dh >= 2) {
                /* the low rows' coarser horizontal levels, into A at half width */
                JxsInvRowJob lo = rowjob(w1, h1);
                lo.lf16 = band(0);
                lo.lf_stride = cp.b[0].cs;
                lo.levels = cp.dh - 1;
                for (uint32_t k = 0; k + 1 < cp.dh; k++) {
                    lo.hf[k] = band(cp.dh - 1 - k);
                    lo.hf_stride[k] = cp.b[cp.dh - 1 - k].cs;
                }
                lo.out32 = A;
                lo.out_stride = w1;
                J.rows[1].push_back(lo);
                f.low32 = A;
                f.low_stride = w1;
            } else {
                f.low16 = band(0);
                f.low_stride = cp.b[0].cs;
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
            v2.bw = bw;
            J.vert[0].push_back(v2);
            f.low32 = B;
            f.low_stride = w1;
        }
        f.hl = band(nb - 3);
        f.hl_stride = cp.b[nb - 3].cs;
        f.lh = band(nb - 2);
        f.lh_stride = cp.b[nb - 2].cs;
        f.hh = band(nb - 1);
        f.hh_stride = cp.b[nb - 1].cs;
        f.out = planes[c];
        f.out_stride = strides[c];
        f.depth = g.depth;
        f.bw = bw;
        f.fq = fq;
        f.w = w;
        f.h = h;
        f.seg = INV_SEG;
        f.tiles_x = (w1 + 255) / 256;
        f.nblk = f.tiles_x * ((h1 + INV_SEG - 1) / INV_SEG);
        J.last.push_back(f);
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
        al256((uint64_t)g.rows * dec_pkt_stride(g) * 4) + al256(g.n_gcli);
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
    uint32_t rmaxw[2] = {0, 0}; /* widest row of each row launch */
    JxsInvLastJob* d_last = nullptr;
    uint32_t* d_lbase = nullptr;
    int nlast = 0;
    uint32_t lblocks = 0;
    uint32_t total_prec = 0, max_npk = 0, total_slices = 0;
    bool need_scan = false; /* a frame without the caller's precinct table */
    bool own_status = true; /* false in a part: the status words belong to the whole plan */
    /* A plan of several frames runs as parts on streams of its own, forked
     * from and joined back to the caller's stream, so that one part's
     * unpacking (instruction-bound) overlaps another's inverse transform
     * (memory-bound). */
    std::vector<svt_jpeg_xs_amd_decode_plan*> parts;
    std::vector<hipStream_t> part_st;
    std::vector<hipEvent_t> part_done;
    hipEvent_t fork = nullptr;
};

namespace jxs {

#define DPLAN_PARTS 4 /* parts a plan of several frames runs as, at most */

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
    for (auto* q : p->parts)
        dplan_free(q);
    for (hipStream_t s : p->part_st)
        (void)hipStreamDestroy(s);
    for (hipEvent_t e : p->part_done)
        (void)hipEventDestroy(e);
    if (p->fork)
        (void)hipEventDestroy(p->fork);
    (void)hipFree(p->d_work);
    if (p->own_status)
        (void)hipFree(p->d_status);
    (void)hipFree(p->d_frames);
    for (int s = 0; s < 2; s++) {
        (void)hipFree(p->d_rows[s]);
        (void)hipFree(p->d_rbase[s]);
        (void)hipFree(p->d_vert[s]);
        (void)hipFree(p->d_vbase[s]);
    }
    (void)hipFree(p->d_last);
    (void)hipFree(p->d_lbase);
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

/* status: the frames' status words (a part's slice of its plan's), or NULL
 * for words of its own */
static svt_jpeg_xs_amd_decode_plan_t* dplan_build(const std::vector<DecSpec>& specs, uint32_t* status) {
    auto* p = new (std::nothrow) svt_jpeg_xs_amd_decode_plan;
    if (!p || specs.empty()) {
        delete p;
        return nullptr;
    }
    p->nfr = (uint32_t)specs.size();
    uint64_t work = 0;
    for (auto& s : specs)
        work += dec_work_bytes(s.geom->host.g);
    if (status) {
        p->d_status = status;
        p->own_status = false;
    }
    if (hipMalloc(&p->d_work, work) != hipSuccess || (!status && hipMalloc(&p->d_status, p->nfr * 4) != hipSuccess)) {
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
        F.gcli = cur;
        cur += al256(g.n_gcli);
        F.slice_base = p->total_slices;
        p->total_slices += g.slices;
        F.status = p->d_status + i;
        F.prec_base = base;
        F.npk_dec = s.geom->npk_dec;
        F.prec_in = s.prec_in;
        F.prec_in_stride = s.prec_stride ? s.prec_stride : 1;
        p->need_scan |= s.prec_in == nullptr;
        F.rows_valid = g.rows;
        base += g.rows;
        p->max_npk = s.geom->npk_dec > p->max_npk ? s.geom->npk_dec : p->max_npk;
        InvJobs one;
        inv_jobs(s.geom->out, s.geom->host.bw, s.geom->host.fq, F.coef, scratch, s.planes, s.strides, one);
        for (int k = 0; k < 2; k++) {
            J.rows[k].insert(J.rows[k].end(), one.rows[k].begin(), one.rows[k].end());
            J.vert[k].insert(J.vert[k].end(), one.vert[k].begin(), one.vert[k].end());
        }
        J.last.insert(J.last.end(), one.last.begin(), one.last.end());
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
            p->rmaxw[k] = r.w > p->rmaxw[k] ? r.w : p->rmaxw[k];
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
    if (ok) {
        std::vector<uint32_t> lb;
        uint32_t blk = 0;
        for (auto& f : J.last) {
            f.base = blk;
            lb.push_back(blk);
            blk += f.nblk;
        }
        p->lblocks = blk;
        p->nlast = (int)J.last.size();
        ok = up(J.last, &p->d_last) == hipSuccess && up(lb, &p->d_lbase) == hipSuccess;
    }
    if (!ok) {
        dplan_free(p);
        return nullptr;
    }
    return p;
}

/* A plan of several frames: DPLAN_PARTS parts of consecutive frames (as
 * even as the count allows), each built as a plan of its own over the
 * whole plan's status words. */
static svt_jpeg_xs_amd_decode_plan_t* dplan_build_parts(const std::vector<DecSpec>& specs) {
    const uint32_t n = (uint32_t)specs.size();
    const uint32_t np = n < DPLAN_PARTS ? n : DPLAN_PARTS;
    if (np <= 1)
        return dplan_build(specs, nullptr);
    auto* p = new (std::nothrow) svt_jpeg_xs_amd_decode_plan;
    if (!p)
        return nullptr;
    p->nfr = n;
    bool ok = hipMalloc(&p->d_status, n * 4) == hipSuccess &&
        hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) == hipSuccess;
    uint32_t first = 0;
    for (uint32_t i = 0; i < np && ok; i++) {
        const uint32_t cnt = n / np + (i < n % np ? 1u : 0u);
        std::vector<DecSpec> sub(specs.begin() + first, specs.begin() + first + cnt);
        svt_jpeg_xs_amd_decode_plan_t* q = dplan_build(sub, p->d_status + first);
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        ok = q != nullptr;
        if (q)
            p->parts.push_back(q);
        if (ok && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess) /* default priority */
            p->part_st.push_back(s);
        else
            ok = false;
        if (ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess)
            p->part_done.push_back(e);
        else
            ok = false;
        first += cnt;
    }
    if (!ok) {
        dplan_free(p);
        return nullptr;
    }
    return p;
}

static hipError_t dplan_launch_one(svt_jpeg_xs_amd_decode_plan_t* p, hipStream_t st);

static hipError_t dplan_launch(svt_jpeg_xs_amd_decode_plan_t* p, hipStream_t st) {
    if (p->parts.empty())
        return dplan_launch_one(p, st);
    hipError_t e;
    if ((e = hipEventRecord(p->fork, st)) != hipSuccess)
        return e;
    for (size_t i = 0; i < p->parts.size(); i++) {
        if ((e = hipStreamWaitEvent(p->part_st[i], p->fork, 0)) != hipSuccess ||
            (e = dplan_launch_one(p->parts[i], p->part_st[i])) != hipSuccess ||
            (e = hipEventRecord(p->part_done[i], p->part_st[i])) != hipSuccess ||
            (e = hipStreamWaitEvent(st, p->part_done[i], 0)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

static hipError_t dplan_launch_one(svt_jpeg_xs_amd_decode_plan_t* p, hipStream_t st) {
    hipError_t e;
    if ((e = jxs_launch_zero(p->d_status, p->nfr, st)) != hipSuccess)
        return e;
    if (p->need_scan && (e = jxs_launch_dec_scan(p->d_frames, (int)p->nfr, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_packets(p->d_frames, (int)p->nfr, p->total_prec, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_unpack(p->d_frames, (int)p->nfr, p->max_npk, p->total_prec, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_dec_slice(p->d_frames, (int)p->nfr, p->total_slices, st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_rows(p->d_rows[0], p->d_rbase[0], p->nrows[0], p->rblocks[0], p->rmaxw[0], st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_vert(p->d_vert[0], p->d_vbase[0], p->nvert[0], p->vblocks[0], st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_rows(p->d_rows[1], p->d_rbase[1], p->nrows[1], p->rblocks[1], p->rmaxw[1], st)) != hipSuccess)
        return e;
    if ((e = jxs_launch_inv_vert(p->d_vert[1], p->d_vbase[1], p->nvert[1], p->vblocks[1], st)) != hipSuccess)
        return e;
    return jxs_launch_inv_last(p->d_last, p->d_lbase, p->nlast, p->lblocks, st);
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

/* A decoder's geometry from its first codestream, with the reference's
 * error order: header (ParseHeader.c), layout and proxy rules (Pi.c:189-405,
 * 567-626; Decoder.c:133-139), then what this decoder does not decode. */
static int32_t geom_from_codestream(const uint8_t* cs, size_t len, DecGeom& dg, uint32_t proxy, CsHeader& h,
                                    bool packet_mode) {
    uint32_t used = 0;
    int32_t e = cs_get_header(cs, len, h, &used);
    if (e)
        return e;
    /* packet input needs the frame size from the header (DecHandle.c:197-204) */
    if (packet_mode && h.d.lcod == 0)
        return SvtJxsErrorBadParameter;
    e = cs_pi_check(h, proxy);
    if (!e)
        e = geom_from_header(h, used, dg.host, &dg.hdr_end);
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

/* ---------------- the C-API decoder ----------------
 *
 * Frames go through a small ring of device slots (codestream, precinct
 * offsets, band planes, output planes, a decode plan bound to them).  A send
 * stages the codestream in the slot's page-locked buffer, queues the copy and
 * the decode on the decoder's stream and returns; a finisher thread takes the
 * frames in send order, waits for each, copies the planes into the caller's
 * image buffer, frees the slot and publishes the frame (callbacks as
 * DecThreadInit.c:215-223, 338-341 and DecThreadFinal.c).  A frame whose
 * header or precinct chain does not parse takes no slot; it is published in
 * order with its error, as the reference's init thread forwards header errors
 * to the final thread (DecThreadInit.c:291-330). */

static const int kDecSlots = 3;

struct DecSlot {
    DecGeom geom;             /* the decoder's geometry; its device copy carries this slot's frame's Qpih, Fs, Rm */
    uint32_t cap = 0;         /* codestream bytes the slot holds */
    uint8_t* d_cs = nullptr;
    uint8_t* h_cs = nullptr;  /* page-locked staging of the codestream */
    uint32_t* d_prec = nullptr;
    uint32_t* h_prec = nullptr; /* page-locked precinct offsets */
    void* d_planes = nullptr;
    JxsDecFrame* h_frame = nullptr; /* page-locked copy of the plan's frame descriptor */
    JxsGeom* h_geom = nullptr;      /* page-locked copy of the slot's geometry */
    uint32_t dyn[4] = {0, 0, 0, 0}; /* Qpih, Fs, Rm, short packet headers in the slot's device geometry */
    std::vector<uint32_t> slice_off;
    svt_jpeg_xs_amd_decode_plan_t* plan = nullptr;
    hipEvent_t done = nullptr;
    uint32_t* h_status = nullptr;
    bool busy = false;
    bool made = false; /* buffers allocated (slots past the first are made on first use) */
};

struct DecJob {
    int slot; /* -1: the frame failed before reaching the device */
    svt_jpeg_xs_frame_t frame;
    SvtJxsErrorType_t err;
    /* the error of the first slice the device did not decode (0: it decoded
     * them all); it stands unless a decoded slice failed first */
    SvtJxsErrorType_t tail = SvtJxsErrorNone;
};

struct DecoderPriv {
    svt_jpeg_xs_decoder_api_t* api = nullptr;
    DecGeom geom;
    CsHeader hdr; /* the first codestream's header: later frames must share its configuration */
    uint32_t verbose = 0;
    bool packet_mode = false;
    size_t plane_off[4] = {0, 0, 0, 0};
    uint64_t plane_bytes = 0;
    std::vector<DecSlot> slots;
    hipStream_t stream = nullptr;  /* copies in and decode launches */
    hipStream_t fstream = nullptr; /* the finisher's copies out */
    std::mutex mu;
    std::condition_variable cv_free, cv_job, cv_done;
    std::deque<DecJob> jobs;  /* sent, in send order, not yet published */
    std::deque<DecJob> done;  /* published, for get_frame */
    uint32_t unpublished = 0; /* sent but not yet in done */
    bool eoc = false, stop = false;
    std::thread finisher;
    /* packet input (DecThreadInit.c:347-460): the frame being filled */
    int pk_slot = -1;
    uint32_t pk_filled = 0;
    svt_jpeg_xs_frame_t pk_frame;
    uint32_t slot_cap = 0; /* codestream bytes a new slot holds */
};

static void slot_release_cs(DecSlot& s) {
    (void)hipFree(s.d_cs);
    if (s.h_cs)
        (void)hipHostFree(s.h_cs);
    s.d_cs = nullptr;
    s.h_cs = nullptr;
    s.cap = 0;
}

static void slot_release(DecSlot& s) {
    dplan_free(s.plan);
    s.plan = nullptr;
    slot_release_cs(s);
    (void)hipFree(s.d_prec);
    (void)hipFree(s.d_planes);
    (void)hipFree(s.geom.d_geom);
    for (void* p : {(void*)s.h_prec, (void*)s.h_frame, (void*)s.h_geom, (void*)s.h_status})
        if (p)
            (void)hipHostFree(p);
    if (s.done)
        (void)hipEventDestroy(s.done);
    s = DecSlot();
}

/* (Re)binds the slot's decode plan to its buffers, with room for cap
 * codestream bytes. */
static bool slot_bind(DecoderPriv* d, DecSlot& s, uint32_t cap) {
    if (cap > s.cap) {
        slot_release_cs(s);
        if (hipMalloc(&s.d_cs, cap) != hipSuccess || hipHostMalloc((void**)&s.h_cs, cap, 0) != hipSuccess)
            return false;
        s.cap = cap;
    }
    dplan_free(s.plan);
    DecSpec sp;
    memset(&sp, 0, sizeof(sp));
    sp.geom = &s.geom;
    sp.d_cs = s.d_cs;
    sp.cs_len = s.cap;
    for (uint32_t c = 0; c < s.geom.out.nc; c++) {
        sp.planes[c] = (uint8_t*)s.d_planes + d->plane_off[c];
        sp.strides[c] = s.geom.out.c[c].w;
    }
    sp.prec_in = s.d_prec;
    sp.prec_stride = 1;
    s.plan = dplan_build({sp}, nullptr);
    if (!s.plan)
        return false;
    return hipMemcpy(s.h_frame, s.plan->d_frames, sizeof(JxsDecFrame), hipMemcpyDeviceToHost) == hipSuccess;
}

static bool slot_create(DecoderPriv* d, DecSlot& s, uint32_t cap) {
    const JxsGeom& g = d->geom.host.g;
    s.made = true;
    s.geom = d->geom;
    s.geom.d_geom = nullptr;
    s.dyn[0] = g.qpih;
    s.dyn[1] = g.fs;
    s.dyn[2] = g.rm;
    s.dyn[3] = g.short_hdr;
    s.slice_off.assign(g.slices, 0);
    if (hipMalloc(&s.geom.d_geom, sizeof(JxsGeom)) != hipSuccess ||
        hipMemcpy(s.geom.d_geom, &g, sizeof(JxsGeom), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s.d_prec, (size_t)g.rows * 4) != hipSuccess || hipMalloc(&s.d_planes, d->plane_bytes) != hipSuccess ||
        hipHostMalloc((void**)&s.h_prec, (size_t)g.rows * 4, 0) != hipSuccess ||
        hipHostMalloc((void**)&s.h_frame, sizeof(JxsDecFrame), 0) != hipSuccess ||
        hipHostMalloc((void**)&s.h_geom, sizeof(JxsGeom), 0) != hipSuccess ||
        hipHostMalloc((void**)&s.h_status, 4, 0) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
        return false;
    return slot_bind(d, s, cap);
}

/* Publishes a finished frame: in send order, releasable, with its error. */
static void publish(DecoderPriv* d, DecJob& j) {
    j.frame.image.ready_to_release = 1;
    j.frame.bitstream.ready_to_release = 1;
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->done.push_back(j);
        d->unpublished--;
    }
    d->cv_done.notify_all();
    if (d->api->callback_get_data_available)
        d->api->callback_get_data_available(d->api, d->api->callback_get_data_available_context);
}

static void finisher_loop(DecoderPriv* d) {
    const JxsGeom& og = d->geom.out;
    const uint32_t pix = og.depth <= 8 ? 1 : 2;
    for (;;) {
        DecJob j;
        {
            std::unique_lock<std::mutex> lk(d->mu);
            d->cv_job.wait(lk, [&] { return d->stop || !d->jobs.empty(); });
            if (d->jobs.empty())
                return;
            j = d->jobs.front();
            d->jobs.pop_front();
        }
        if (j.slot >= 0) {
            DecSlot& s = d->slots[j.slot];
            hipError_t he = hipEventSynchronize(s.done);
            if (he != hipSuccess)
                j.err = SvtJxsErrorDecoderInternal;
            else if (*s.h_status)
                j.err = SvtJxsErrorDecoderInvalidBitstream;
            else if (j.tail)
                j.err = j.tail;
            else {
                for (uint32_t c = 0; c < og.nc && he == hipSuccess; c++)
                    he = hipMemcpy2DAsync(j.frame.image.data_yuv[c], (size_t)j.frame.image.stride[c] * pix,
                                          (uint8_t*)s.d_planes + d->plane_off[c], (size_t)og.c[c].w * pix,
                                          (size_t)og.c[c].w * pix, og.c[c].h, hipMemcpyDeviceToHost, d->fstream);
                if (he == hipSuccess)
                    he = hipStreamSynchronize(d->fstream);
                if (he != hipSuccess)
                    j.err = SvtJxsErrorDecoderInternal;
            }
            {
                std::lock_guard<std::mutex> lk(d->mu);
                s.busy = false;
            }
            d->cv_free.notify_all();
            if (d->api->callback_send_data_available)
                d->api->callback_send_data_available(d->api, d->api->callback_send_data_available_context);
        }
        publish(d, j);
    }
}

static void decoder_destroy(DecoderPriv* d) {
    if (!d)
        return;
    if (d->finisher.joinable()) {
        {
            std::lock_guard<std::mutex> lk(d->mu);
            d->stop = true;
        }
        d->cv_job.notify_all();
        d->finisher.join();
    }
    if (d->stream)
        (void)hipStreamSynchronize(d->stream);
    for (auto& s : d->slots)
        slot_release(s);
    (void)hipFree(d->geom.d_geom);
    if (d->stream)
        (void)hipStreamDestroy(d->stream);
    if (d->fstream)
        (void)hipStreamDestroy(d->fstream);
    delete d;
}

/* A free slot (waiting for one when blocking), marked busy; -1 if none.
 * A slot not made yet is made when every made slot is busy, so a decoder
 * that never has two frames in flight holds one slot. */
static int acquire_slot(DecoderPriv* d, bool blocking) {
    std::unique_lock<std::mutex> lk(d->mu);
    for (;;) {
        for (int i = 0; i < (int)d->slots.size(); i++)
            if (d->slots[i].made && !d->slots[i].busy) {
                d->slots[i].busy = true;
                return i;
            }
        for (int i = 0; i < (int)d->slots.size(); i++)
            if (!d->slots[i].made) {
                d->slots[i].busy = true;
                lk.unlock();
                const bool ok = slot_create(d, d->slots[i], d->slot_cap);
                lk.lock();
                if (ok)
                    return i;
                slot_release(d->slots[i]); /* back to not made; another slot frees up in time */
                break;
            }
        if (!blocking)
            return -1;
        d->cv_free.wait(lk);
    }
}

static void free_slot(DecoderPriv* d, int si) {
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->slots[si].busy = false;
    }
    d->cv_free.notify_all();
}

static void enqueue(DecoderPriv* d, const DecJob& j) {
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->jobs.push_back(j);
        d->unpublished++;
    }
    d->cv_job.notify_all();
}

/* A frame's per-frame coding flags (picture_header_dynamic_t, Pi.h:63-80:
 * Qpih, Fs, Rm, Lh may change from frame to frame; Decoder.c:576, Packing.c:
 * 532, 616, 829) over the decoder's geometry.  Precisions and non-linear
 * transforms this decoder does not decode are SvtJxsErrorDecoderInvalidBitstream. */
static int32_t frame_flags(const DecoderPriv* d, const CsHeader& h, JxsGeom& fg) {
    if (h.d.tnlt != 0 || (h.d.bw != 20 && h.d.bw != 18))
        return SvtJxsErrorDecoderInvalidBitstream;
    fg.qpih = h.d.qpih;
    fg.fs = h.d.fs;
    fg.rm = h.d.rm;
    const bool short_ok = ((uint64_t)d->hdr.c.w * d->hdr.c.nc) < 32752;
    fg.short_hdr = (short_ok && !h.d.lh) ? 1u : 0u;
    return SvtJxsErrorNone;
}

/* Queues the decode of the first rows_valid precinct rows of the codestream
 * staged in slot si (fsize bytes, precinct offsets in h_prec, per-frame flags
 * in fg) into frame's image buffer; `tail` is the error of the rows after
 * them. */
static SvtJxsErrorType_t launch_slot(DecoderPriv* d, int si, uint32_t fsize, const JxsGeom& fg,
                                     const svt_jpeg_xs_frame_t& frame, uint32_t rows_valid, SvtJxsErrorType_t tail) {
    DecSlot& s = d->slots[si];
    const JxsGeom& g = d->geom.host.g;
    hipError_t he = hipSuccess;
    if (fg.qpih != s.dyn[0] || fg.fs != s.dyn[1] || fg.rm != s.dyn[2] || fg.short_hdr != s.dyn[3]) {
        *s.h_geom = g;
        s.h_geom->qpih = fg.qpih;
        s.h_geom->fs = fg.fs;
        s.h_geom->rm = fg.rm;
        s.h_geom->short_hdr = fg.short_hdr;
        s.h_geom->pkt_hdr_bytes = fg.short_hdr ? 5 : 7;
        he = hipMemcpyAsync(s.geom.d_geom, s.h_geom, sizeof(JxsGeom), hipMemcpyHostToDevice, d->stream);
        s.dyn[0] = fg.qpih;
        s.dyn[1] = fg.fs;
        s.dyn[2] = fg.rm;
        s.dyn[3] = fg.short_hdr;
    }
    s.h_frame->cs_len = fsize;
    s.h_frame->rows_valid = rows_valid;
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.plan->d_frames, s.h_frame, sizeof(JxsDecFrame), hipMemcpyHostToDevice, d->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.d_cs, s.h_cs, fsize, hipMemcpyHostToDevice, d->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.d_prec, s.h_prec, (size_t)g.rows * 4, hipMemcpyHostToDevice, d->stream);
    if (he == hipSuccess)
        he = dplan_launch(s.plan, d->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.h_status, s.plan->d_status, 4, hipMemcpyDeviceToHost, d->stream);
    if (he == hipSuccess)
        he = hipEventRecord(s.done, d->stream);
    if (he != hipSuccess) {
        s.dyn[0] = ~0u; /* unknown: upload again next time */
        free_slot(d, si);
        return SvtJxsErrorDecoderInternal;
    }
    DecJob j{si, frame, SvtJxsErrorNone};
    j.tail = tail;
    enqueue(d, j);
    return SvtJxsErrorNone;
}

/* Validates the caller's image buffer against the output geometry
 * (DecHandle.c:379-401; the last row may be shorter than the stride). */
static SvtJxsErrorType_t check_image(const DecoderPriv* d, const svt_jpeg_xs_image_buffer_t& im) {
    const JxsGeom& og = d->geom.out;
    const uint32_t pix = og.depth <= 8 ? 1 : 2;
    for (uint32_t c = 0; c < og.nc; c++) {
        if (!im.data_yuv[c] || im.stride[c] < og.c[c].w)
            return SvtJxsErrorBadParameter;
        const uint64_t need = (uint64_t)im.stride[c] * pix * (og.c[c].h - 1) + (uint64_t)og.c[c].w * pix;
        if (im.alloc_size[c] < need)
            return SvtJxsErrorBadParameter;
    }
    return SvtJxsErrorNone;
}

/* The frame staged in slot si (len bytes), as the reference's init thread
 * takes it (DecThreadInit.c:291-338, Decoder.c:377-396, 509-531): its header
 * and configuration, then each slice's precinct walk; the slices before the
 * first one that fails go to the device, and the frame's error is that of
 * the first failing slice in slice order.  A frame that fails before any
 * slice takes no device time and is handed back with its error. */
static SvtJxsErrorType_t submit_staged(DecoderPriv* d, int si, uint32_t len, const svt_jpeg_xs_frame_t& frame) {
    DecSlot& s = d->slots[si];
    const JxsGeom& g = d->geom.host.g;
    CsHeader h;
    uint32_t used = 0;
    JxsGeom fg = g;
    int32_t e = cs_get_header(s.h_cs, len, h, &used);
    if (!e && !cs_same_config(h, d->hdr))
        e = SvtJxsErrorDecoderConfigChange;
    if (!e)
        e = frame_flags(d, h, fg);
    /* Bw and Fq enter the inverse transform jobs: a frame that changes them
     * rebinds the slot's plan */
    if (!e && (h.d.bw != s.geom.host.bw || h.d.fq != s.geom.host.fq)) {
        s.geom.host.bw = h.d.bw;
        s.geom.host.fq = h.d.fq;
        if (!slot_bind(d, s, s.cap)) {
            free_slot(d, si);
            enqueue(d, DecJob{-1, frame, SvtJxsErrorInsufficientResources});
            return SvtJxsErrorNone;
        }
    }
    uint32_t bad = 0, end = 0;
    if (!e) {
        e = cs_slice_walk(s.h_cs, len, used, g.slices, g.pps, g.rows, g.nb_exist, s.h_prec, s.slice_off.data(), &end,
                          &bad);
        /* a slice header whose length field is not 4 fails that slice's decode */
        for (uint32_t sl = 0; sl < bad; sl++)
            if (cs_slice_header_len(s.h_cs, s.slice_off[sl]) != 4) {
                bad = sl;
                e = SvtJxsErrorDecoderInvalidBitstream;
                break;
            }
        if (bad == 0 && !e)
            e = SvtJxsErrorDecoderInvalidBitstream; /* no slice (not reached: a picture has one) */
    }
    if (e && bad == 0) {
        free_slot(d, si);
        enqueue(d, DecJob{-1, frame, (SvtJxsErrorType_t)e});
        return SvtJxsErrorNone;
    }
    const uint32_t rows_valid = std::min(bad * g.pps, g.rows);
    const uint32_t dev_len = bad < g.slices ? s.slice_off[bad] : end;
    return launch_slot(d, si, dev_len, fg, frame, rows_valid, (SvtJxsErrorType_t)e);
}

}  // namespace jxs

using namespace jxs;

extern "C" {

/* ParseHeader.c:652-889: the frame's size from its markers and precinct
 * lengths (Lcod with fast_search), without decoding. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(
    const uint8_t* bitstream_buf, size_t bitstream_buf_size, svt_jpeg_xs_image_config_t* out_image_config,
    uint32_t* frame_size_out, uint32_t fast_search, proxy_mode_t proxy_mode) {
    return (SvtJxsErrorType_t)cs_frame_size(bitstream_buf, bitstream_buf_size, out_image_config, frame_size_out,
                                            fast_search, (uint32_t)proxy_mode);
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
    if (!dec_api || !bitstream_buf || codestream_size == 0)
        return SvtJxsErrorDecoderInvalidPointer;
    dec_api->private_ptr = nullptr;
    if (dec_api->packetization_mode > 1 || (uint32_t)dec_api->proxy_mode >= proxy_mode_max)
        return SvtJxsErrorBadParameter;
    auto* d = new (std::nothrow) DecoderPriv;
    if (!d)
        return SvtJxsErrorInsufficientResources;
    d->api = dec_api;
    d->verbose = dec_api->verbose;
    d->packet_mode = dec_api->packetization_mode == 1;
    int32_t e = geom_from_codestream(bitstream_buf, codestream_size, d->geom, (uint32_t)dec_api->proxy_mode, d->hdr,
                                     d->packet_mode);
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
    d->plane_bytes = planes;
    /* a constant-rate stream carries its frame size (Lcod); otherwise start
     * from the first codestream's size.  Lcod is a header field: a damaged one
     * must not size the staging, which is bounded by twice the raw picture
     * (a codestream never needs more); a slot grows when a frame needs it */
    const uint64_t raw = (uint64_t)g.w * g.h * g.nc * (g.depth <= 8 ? 1u : 2u);
    const uint64_t bound = 2 * raw + 65536;
    uint64_t cap = g.lcod ? std::min<uint64_t>(g.lcod, bound) : (uint64_t)codestream_size;
    cap = std::min<uint64_t>(std::max<uint64_t>(cap, 1), 0xFFFFFFFFull);
    if (hipMalloc(&d->geom.d_geom, sizeof(JxsGeom)) != hipSuccess ||
        hipMemcpy(d->geom.d_geom, &g, sizeof(JxsGeom), hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&d->fstream, hipStreamNonBlocking) != hipSuccess) {
        decoder_destroy(d);
        return SvtJxsErrorInsufficientResources;
    }
    d->slots.resize(kDecSlots);
    d->slot_cap = (uint32_t)cap;
    if (!slot_create(d, d->slots[0], (uint32_t)cap)) {
        decoder_destroy(d);
        return SvtJxsErrorInsufficientResources;
    }
    try {
        d->finisher = std::thread(finisher_loop, d);
    } catch (...) {
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

/* One whole codestream per call (DecHandle.c:363-425).  The codestream is
 * staged before the call returns, so its buffer may be reused at once; the
 * image buffer is written when get_frame hands the frame back.  Without a free
 * slot a non-blocking call returns SvtJxsErrorNoErrorEmptyQueue. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                            svt_jpeg_xs_frame_t* dec_input, uint8_t blocking_flag) {
    if (!dec_api || !dec_api->private_ptr || !dec_input)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    if (d->packet_mode)
        return SvtJxsErrorUndefined; /* DecHandle.c:371-377 */
    SvtJxsErrorType_t e = check_image(d, dec_input->image);
    if (e)
        return e;
    if (!dec_input->bitstream.buffer)
        return SvtJxsErrorDecoderInvalidPointer;
    const size_t len = dec_input->bitstream.used_size ? dec_input->bitstream.used_size
                                                      : dec_input->bitstream.allocation_size;
    const int si = acquire_slot(d, blocking_flag != 0);
    if (si < 0)
        return SvtJxsErrorNoErrorEmptyQueue;
    DecSlot& s = d->slots[si];
    /* the whole buffer is staged: the frame is found in it as the decode
     * walks it, and only the frame's bytes go to the device */
    if (len > s.cap && !slot_bind(d, s, (uint32_t)len)) {
        free_slot(d, si);
        return SvtJxsErrorInsufficientResources;
    }
    memcpy(s.h_cs, dec_input->bitstream.buffer, len);
    return submit_staged(d, si, (uint32_t)len, *dec_input);
}

/* Packet input (DecThreadInit.c:347-460): a call takes at most the rest of
 * the current frame (Lcod bytes from its first packet), so *bytes_used may be
 * less than the chunk and the caller sends the remainder again.  The image
 * buffer of a frame is the one passed with its first packet.  The call returns
 * SvtJxsErrorDecoderBitstreamTooShort until the frame is complete, then
 * queues it; a frame that does not parse is handed back by get_frame with its
 * error. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_packet(svt_jpeg_xs_decoder_api_t* dec_api,
                                                             svt_jpeg_xs_frame_t* dec_input, uint32_t* bytes_used) {
    if (!dec_api || !dec_api->private_ptr || !dec_input || !bytes_used)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    *bytes_used = 0;
    if (!d->packet_mode)
        return SvtJxsErrorUndefined;
    if (!dec_input->bitstream.buffer)
        return SvtJxsErrorDecoderInvalidPointer;
    if (d->pk_slot < 0) {
        SvtJxsErrorType_t e = check_image(d, dec_input->image);
        if (e)
            return e;
        d->pk_slot = acquire_slot(d, true);
        d->pk_filled = 0;
        /* a frame of Lcod bytes is staged whole */
        DecSlot& ns = d->slots[d->pk_slot];
        if (d->geom.host.g.lcod > ns.cap && !slot_bind(d, ns, d->geom.host.g.lcod)) {
            free_slot(d, d->pk_slot);
            d->pk_slot = -1;
            return SvtJxsErrorInsufficientResources;
        }
        d->pk_frame = *dec_input;
        memset(&d->pk_frame.bitstream, 0, sizeof(d->pk_frame.bitstream));
    }
    DecSlot& s = d->slots[d->pk_slot];
    const uint32_t lcod = d->geom.host.g.lcod;
    const uint32_t n = std::min<uint32_t>(lcod - d->pk_filled, dec_input->bitstream.used_size);
    memcpy(s.h_cs + d->pk_filled, dec_input->bitstream.buffer, n);
    d->pk_filled += n;
    *bytes_used = n;
    dec_input->bitstream.ready_to_release = 1;
    dec_input->image.ready_to_release = 0;
    uint32_t fsize = 0;
    const int32_t e = cs_frame_size(s.h_cs, d->pk_filled, nullptr, &fsize, 0, proxy_mode_full);
    if (e == SvtJxsErrorDecoderBitstreamTooShort && d->pk_filled < lcod)
        return SvtJxsErrorDecoderBitstreamTooShort;
    const int si = d->pk_slot;
    const uint32_t filled = d->pk_filled;
    d->pk_slot = -1;
    d->pk_filled = 0;
    return submit_staged(d, si, filled, d->pk_frame);
}

/* Frames in send order (DecHandle.c:435-465); SvtJxsDecoderEndOfCodestream
 * once send_eoc was called and every frame sent before it was handed out. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                           svt_jpeg_xs_frame_t* dec_output, uint8_t blocking_flag) {
    if (!dec_api || !dec_api->private_ptr || !dec_output)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    std::unique_lock<std::mutex> lk(d->mu);
    auto finished = [&] { return d->eoc && d->unpublished == 0 && d->pk_slot < 0; };
    if (d->done.empty()) {
        if (finished())
            return SvtJxsDecoderEndOfCodestream;
        if (!blocking_flag) {
            dec_output->user_prv_ctx_ptr = nullptr;
            return SvtJxsErrorNoErrorEmptyQueue;
        }
        d->cv_done.wait(lk, [&] { return !d->done.empty() || finished(); });
        if (d->done.empty())
            return SvtJxsDecoderEndOfCodestream;
    }
    DecJob r = d->done.front();
    d->done.pop_front();
    *dec_output = r.frame;
    return r.err;
}

/* End of input (DecHandle.c:466-540).  A frame whose packets had not all
 * arrived is handed back with SvtJxsErrorDecoderBitstreamTooShort. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_eoc(svt_jpeg_xs_decoder_api_t* dec_api) {
    if (!dec_api || !dec_api->private_ptr)
        return SvtJxsErrorDecoderInvalidPointer;
    DecoderPriv* d = (DecoderPriv*)dec_api->private_ptr;
    if (d->pk_slot >= 0) {
        const int si = d->pk_slot;
        d->pk_slot = -1;
        d->pk_filled = 0;
        free_slot(d, si);
        enqueue(d, DecJob{-1, d->pk_frame, SvtJxsErrorDecoderBitstreamTooShort});
    }
    {
        std::lock_guard<std::mutex> lk(d->mu);
        d->eoc = true;
    }
    d->cv_done.notify_all();
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
    return dplan_build_parts(specs);
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
