/*
 * jxs_codestream.cpp -- the decoder's codestream checks on the host, with the
 * reference decoder's error codes for every damaged or truncated codestream:
 *
 *   cs_get_header     the header parse of ParseHeader.c:516-634 (markers in
 *                     order, every field range, the mandatory-marker rules).
 *   cs_frame_size     the frame-size walk of ParseHeader.c:652-889 (markers,
 *                     then precincts by their Lprc fields up to EOC).
 *   cs_pi_check       the layout validity rules of pi_compute (Pi.c:189-405)
 *                     and the colour-format rule of decoder init
 *                     (Decoder.c:51-79, 133-139).
 *   cs_slice_walk     the per-slice precinct walk of DecThreadInit.c:35-131
 *                     and the EOC check after the last slice (:160-191).
 *
 * Nothing here touches the device; the decode kernels only ever see precinct
 * offsets these walks produced.
 */
#include <string.h>

#include "SvtJpegxsDec.h"
#include "jxs_codestream.h"

namespace jxs {

namespace {

/* Byte reader with a bit cursor inside the current byte (the reference's
 * bitstream_reader_t, BitstreamReader.c:9-300). */
struct Rd {
    const uint8_t* m;
    uint32_t size, off, bits;
    /* whether n whole bytes are left after the partly read one */
    bool enough(uint32_t n) const { return (int64_t)n <= (int64_t)size - off - (bits ? 1 : 0); }
    uint32_t u8() { return m[off++]; }
    uint32_t u16() {
        const uint32_t v = ((uint32_t)m[off] << 8) | m[off + 1];
        off += 2;
        return v;
    }
    uint32_t u32() {
        const uint32_t v = ((uint32_t)m[off] << 24) | ((uint32_t)m[off + 1] << 16) | ((uint32_t)m[off + 2] << 8) | m[off + 3];
        off += 4;
        return v;
    }
    uint32_t bit() {
        const uint32_t v = (m[off] >> (7 - bits)) & 1u;
        if (++bits == 8) {
            bits = 0;
            off++;
        }
        return v;
    }
    void align() {
        if (bits) {
            bits = 0;
            off++;
        }
    }
};

enum : uint32_t {
    M_SOC = 0xFF10, M_EOC = 0xFF11, M_PIH = 0xFF12, M_CDT = 0xFF13, M_WGT = 0xFF14, M_COM = 0xFF15,
    M_NLT = 0xFF16, M_CWD = 0xFF17, M_CTS = 0xFF18, M_CRG = 0xFF19, M_SLH = 0xFF20, M_CAP = 0xFF50
};

constexpr uint32_t kPihBytes = 26;
constexpr uint32_t kSpecMaxComps = 8;
constexpr uint32_t kPrecMaxBytes = 1048575; /* Lprc < 2^20 */

int32_t get_cap(Rd& r, CsHeader& h) {
    if (h.c.has_cap)
        return SvtJxsErrorDecoderInvalidBitstream;
    h.c.has_cap = 1;
    if (!r.enough(4))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != M_CAP)
        return SvtJxsErrorDecoderInvalidBitstream;
    const uint16_t n = (uint16_t)(r.u16() - 2);
    if (n * 8u > CS_MAX_CAP)
        return SvtJxsErrorDecoderInvalidBitstream;
    if (!r.enough(n))
        return SvtJxsErrorDecoderBitstreamTooShort;
    for (uint32_t i = 0; i < n * 8u; i++)
        h.c.cap[i] = (uint8_t)r.bit();
    r.align();
    return SvtJxsErrorNone;
}

int32_t get_pih(Rd& r, CsHeader& h) {
    if (!h.c.has_cap || h.c.has_pih)
        return SvtJxsErrorDecoderInvalidBitstream;
    h.c.has_pih = 1;
    if (!r.enough(kPihBytes + 2))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != M_PIH || r.u16() != kPihBytes)
        return SvtJxsErrorDecoderInvalidBitstream;
    CsHeader::Dyn& d = h.d;
    CsHeader::Const& c = h.c;
    d.lcod = r.u32();
    c.ppih = r.u16();
    c.plev = r.u16();
    c.w = (int32_t)r.u16();
    c.h = (int32_t)r.u16();
    c.cw = r.u16();
    c.hsl = r.u16();
    if (c.hsl < 1)
        return SvtJxsErrorDecoderInvalidBitstream;
    c.nc = (uint8_t)r.u8();
    if (c.nc < 1 || c.nc > kSpecMaxComps)
        return SvtJxsErrorDecoderInvalidBitstream;
    c.cgs = (int32_t)r.u8();
    if (c.cgs != 4)
        return SvtJxsErrorDecoderInvalidBitstream;
    c.sgs = (int32_t)r.u8();
    if (c.sgs != 8)
        return SvtJxsErrorDecoderInvalidBitstream;
    d.bw = (uint8_t)r.u8();
    uint32_t v = r.u8();
    d.fq = (uint8_t)(v >> 4);
    d.br = (uint8_t)(v & 15);
    if (!(d.fq == 0 || (d.bw == 18 && d.fq == 6) || (d.bw == 20 && d.fq == 8)))
        return SvtJxsErrorDecoderInvalidBitstream;
    if (d.br != 4)
        return SvtJxsErrorDecoderInvalidBitstream;
    v = r.u8();
    d.fslc = (uint8_t)(v >> 7);
    d.ppoc = (uint8_t)((v >> 4) & 7);
    c.cpih = (uint8_t)(v & 15);
    if (d.fslc != 0 || d.ppoc != 0 || (c.cpih != 0 && c.cpih != 1 && c.cpih != 3))
        return SvtJxsErrorDecoderInvalidBitstream;
    v = r.u8();
    c.dh = (int32_t)(v >> 4);
    c.dv = (int32_t)(v & 15);
    v = r.u8();
    d.lh = (uint8_t)(v >> 7);
    d.rl = (uint8_t)((v >> 6) & 1);
    d.qpih = (uint8_t)((v >> 4) & 3);
    d.fs = (uint8_t)((v >> 2) & 3);
    d.rm = (uint8_t)(v & 3);
    if (d.qpih > 1 || d.fs > 1 || d.rm > 1)
        return SvtJxsErrorDecoderInvalidBitstream;
    return SvtJxsErrorNone;
}

int32_t get_cdt(Rd& r, CsHeader& h) {
    CsHeader::Const& c = h.c;
    if (!c.has_cap || !c.has_pih || c.has_cdt)
        return SvtJxsErrorDecoderInvalidBitstream;
    c.has_cdt = 1;
    if (!r.enough(c.nc * 2u + 2))
        return SvtJxsErrorDecoderBitstreamTooShort;
    const uint16_t n = (uint16_t)(r.u16() - 2);
    if (n != 2u * c.nc || c.nc > CS_MAX_COMPS)
        return SvtJxsErrorDecoderInvalidBitstream;
    for (uint32_t k = 0; k < c.nc; k++) {
        c.depth[k] = (uint8_t)r.u8();
        if (c.depth[k] < 8 || c.depth[k] > 16)
            return SvtJxsErrorDecoderInvalidBitstream;
        const uint32_t v = r.u8();
        c.sx[k] = v >> 4;
        c.sy[k] = v & 15;
    }
    return SvtJxsErrorNone;
}

int32_t get_wgt(Rd& r, CsHeader& h) {
    CsHeader::Const& c = h.c;
    if (!c.has_cap || !c.has_pih || c.has_wgt)
        return SvtJxsErrorDecoderInvalidBitstream;
    c.has_wgt = 1;
    if (!r.enough(2))
        return SvtJxsErrorDecoderBitstreamTooShort;
    const uint16_t n = (uint16_t)(r.u16() - 2);
    if (n & 1)
        return SvtJxsErrorDecoderBitstreamTooShort; /* as the reference reports an odd table */
    const uint32_t nb = n / 2u;
    if (nb > CS_MAX_BANDS)
        return SvtJxsErrorDecoderInvalidBitstream;
    if (!r.enough(n))
        return SvtJxsErrorDecoderBitstreamTooShort;
    for (uint32_t b = 0; b < nb; b++) {
        c.gain[b] = (uint8_t)r.u8();
        c.prio[b] = (uint8_t)r.u8();
    }
    return SvtJxsErrorNone;
}

int32_t get_nlt(Rd& r, CsHeader& h) {
    if (!r.enough(3))
        return SvtJxsErrorDecoderBitstreamTooShort;
    r.u16();
    h.d.tnlt = (uint8_t)r.u8();
    if (h.d.tnlt == 1) {
        if (!r.enough(2))
            return SvtJxsErrorDecoderBitstreamTooShort;
        r.u16();
    } else if (h.d.tnlt == 2) {
        if (!r.enough(9))
            return SvtJxsErrorDecoderBitstreamTooShort;
        r.u32();
        r.u32();
        r.u8();
    } else {
        return SvtJxsErrorDecoderInvalidBitstream;
    }
    return SvtJxsErrorNone;
}

int32_t get_cwd(Rd& r, CsHeader& h) {
    if (!r.enough(3))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != 3)
        return SvtJxsErrorDecoderInvalidBitstream;
    h.c.sd = (int32_t)r.u8();
    return SvtJxsErrorNone;
}

int32_t get_cts(Rd& r, CsHeader& h) {
    if (h.d.has_cts)
        return SvtJxsErrorDecoderInvalidBitstream;
    h.d.has_cts = 1;
    if (!r.enough(4))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != 4)
        return SvtJxsErrorDecoderInvalidBitstream;
    r.u8();
    r.u8();
    return SvtJxsErrorNone;
}

int32_t get_crg(Rd& r, CsHeader& h) {
    if (h.d.has_crg)
        return SvtJxsErrorDecoderInvalidBitstream;
    h.d.has_crg = 1;
    const uint32_t n = 2u + h.c.nc * 4u;
    if (!r.enough(n))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != n || h.c.nc > CS_MAX_COMPS)
        return SvtJxsErrorDecoderInvalidBitstream;
    for (uint32_t k = 0; k < h.c.nc; k++)
        r.u32();
    return SvtJxsErrorNone;
}

int32_t get_com(Rd& r) {
    if (!r.enough(4))
        return SvtJxsErrorDecoderBitstreamTooShort;
    const uint16_t n = (uint16_t)(r.u16() - 4);
    r.u16();
    if (!r.enough(n))
        return SvtJxsErrorDecoderBitstreamTooShort;
    r.off += n;
    return SvtJxsErrorNone;
}

inline uint32_t be16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }
inline uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

}  // namespace

int32_t cs_get_header(const uint8_t* cs, size_t len, CsHeader& h, uint32_t* used) {
    memset(&h, 0, sizeof(h));
    Rd r{cs, (uint32_t)len, 0, 0};
    if (!r.enough(2))
        return SvtJxsErrorDecoderBitstreamTooShort;
    if (r.u16() != M_SOC)
        return SvtJxsErrorDecoderInvalidBitstream;
    int32_t e = get_cap(r, h);
    if (!e)
        e = get_pih(r, h);
    if (e)
        return e;
    for (;;) {
        if (!r.enough(2))
            return SvtJxsErrorDecoderBitstreamTooShort;
        const uint32_t m = r.u16();
        switch (m) {
        case M_CDT:
            e = get_cdt(r, h);
            break;
        case M_WGT:
            e = get_wgt(r, h);
            break;
        case M_NLT:
            e = get_nlt(r, h);
            break;
        case M_CWD:
            e = get_cwd(r, h);
            break;
        case M_CTS:
            e = get_cts(r, h);
            break;
        case M_CRG:
            e = get_crg(r, h);
            break;
        case M_COM:
            e = get_com(r);
            break;
        case M_SLH: {
            r.off -= 2;
            const CsHeader::Const& c = h.c;
            if (h.d.fq == 0)
                for (uint32_t k = 1; k < c.nc; k++)
                    if (c.depth[k] != c.depth[0])
                        return SvtJxsErrorDecoderInvalidBitstream;
            if (!c.has_cap || !c.has_pih || !c.has_cdt || !c.has_wgt)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (c.cpih)
                for (uint32_t k = 1; k < c.nc; k++)
                    if (c.sx[k] != c.sx[0] || c.sy[k] != c.sy[0])
                        return SvtJxsErrorDecoderInvalidBitstream;
            if (c.cpih == 3 && (!h.d.has_cts || !h.d.has_crg))
                return SvtJxsErrorDecoderInvalidBitstream;
            if (used)
                *used = r.off;
            return SvtJxsErrorNone;
        }
        default:
            return SvtJxsErrorDecoderInvalidBitstream;
        }
        if (e)
            return e;
    }
}

bool cs_same_config(const CsHeader& a, const CsHeader& b) { return memcmp(&a.c, &b.c, sizeof(a.c)) == 0; }

uint32_t cs_slice_header_len(const uint8_t* buf, uint32_t slice_off) { return be16(buf + slice_off + 2); }

/* The colour format named by the components and their sampling (Decoder.c:51-79). */
static ColourFormat_t format_of(uint32_t nc, const uint32_t sx[], const uint32_t sy[]) {
    if (nc == 3) {
        if (sx[0] == 1 && sx[1] == 1 && sx[2] == 1 && sy[0] == 1 && sy[1] == 1 && sy[2] == 1)
            return COLOUR_FORMAT_PLANAR_YUV444_OR_RGB;
        if (sx[0] == 1 && sx[1] == 2 && sx[2] == 2 && sy[0] == 1 && sy[1] == 1 && sy[2] == 1)
            return COLOUR_FORMAT_PLANAR_YUV422;
        if (sx[0] == 1 && sx[1] == 2 && sx[2] == 2 && sy[0] == 1 && sy[1] == 2 && sy[2] == 2)
            return COLOUR_FORMAT_PLANAR_YUV420;
        return COLOUR_FORMAT_INVALID;
    }
    if (nc == 4)
        return COLOUR_FORMAT_PLANAR_4_COMPONENTS;
    if (nc == 1)
        return COLOUR_FORMAT_GRAY;
    return COLOUR_FORMAT_INVALID;
}

int32_t cs_frame_size(const uint8_t* buf, size_t size, svt_jpeg_xs_image_config_t* cfg, uint32_t* frame_size,
                      uint32_t fast_search, uint32_t proxy) {
    if (!frame_size)
        return SvtJxsErrorDecoderInvalidPointer;
    if (proxy >= proxy_mode_max)
        return SvtJxsErrorBadParameter;
    *frame_size = 0;
    int32_t nc = -1, sd = 0, dh = -1, dv = -1, bands = -1;
    uint32_t sx[8], sy[8];
    for (int k = 0; k < 8; k++)
        sx[k] = sy[k] = 99; /* unset entries fail the sampling check */
    uint32_t lcod = 0, off = 0, hs = 0, W = 0, H = 0, depth = 0;
    bool in_slices = false;
    if (cfg)
        memset(cfg, 0, sizeof(*cfg));
    if (!buf || size == 0)
        return SvtJxsErrorDecoderInvalidPointer;
    for (;;) {
        if (off + 2 > size)
            return SvtJxsErrorDecoderBitstreamTooShort;
        const uint32_t m = be16(buf + off);
        switch (m) {
        case M_SOC:
            off += 2;
            break;
        case M_PIH:
            if (in_slices)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 28 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            hs = be16(buf + off + 2);
            if (off + hs + 2 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            lcod = be32(buf + off + 4);
            W = be16(buf + off + 12);
            H = be16(buf + off + 14);
            nc = buf[off + 20];
            dh = buf[off + 26] >> 4;
            dv = buf[off + 26] & 15;
            off += 2 + hs;
            break;
        case M_CWD:
            if (in_slices)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 4 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            hs = be16(buf + off + 2);
            sd = off + 4 < size ? buf[off + 4] : 0; /* a byte the buffer does not hold reads as 0 */
            off += 2 + hs;
            break;
        case M_CDT: {
            if (in_slices)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 4 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            hs = be16(buf + off + 2);
            if (off + hs + 2 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            if (nc < 1 || nc > (int32_t)kSpecMaxComps)
                return SvtJxsErrorDecoderInvalidBitstream;
            if ((uint32_t)(2 * nc + 2) > hs)
                return SvtJxsErrorDecoderInvalidBitstream;
            uint32_t o = off + 4;
            for (int32_t k = 0; k < nc; k++) {
                if (k == 0)
                    depth = buf[o];
                sx[k] = buf[o + 1] >> 4;
                sy[k] = buf[o + 1] & 15;
                o += 2;
            }
            off += 2 + hs;
            break;
        }
        case M_CAP:
        case M_WGT:
        case M_NLT:
        case M_CTS:
        case M_CRG:
        case M_COM:
            if (in_slices)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 4 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            hs = be16(buf + off + 2);
            off += 2 + hs;
            break;
        case M_SLH:
            if (off + 6 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            hs = be16(buf + off + 2);
            off += 2 + hs;
            if (!in_slices) {
                in_slices = true;
                if (nc < 1 || dh < 0 || dv < 0 || sd >= nc)
                    return SvtJxsErrorDecoderInvalidBitstream;
                bands = sd;
                for (int32_t k = 0; k < nc - sd; k++) {
                    if (sx[k] < 1 || sy[k] < 1 || sx[k] > 2 || sy[k] > sx[k])
                        return SvtJxsErrorDecoderInvalidBitstream;
                    const int32_t nx = dh, ny = dv - ((int32_t)sy[k] - 1);
                    if (ny < 0 || ny > nx)
                        return SvtJxsErrorDecoderInvalidBitstream;
                    bands += 2 * ny + nx + 1;
                }
                if (cfg) {
                    cfg->width = W;
                    cfg->height = H;
                    cfg->bit_depth = (uint8_t)depth;
                    cfg->components_num = (uint8_t)nc;
                    cfg->format = format_of((uint32_t)nc, sx, sy);
                    int32_t k = 0;
                    for (; k < nc - sd && k < MAX_COMPONENTS_NUM; k++) {
                        cfg->components[k].width = W >> (sx[k] - 1);
                        cfg->components[k].height = H >> (sy[k] - 1);
                    }
                    for (; k < nc && k < MAX_COMPONENTS_NUM; k++) {
                        cfg->components[k].width = W;
                        cfg->components[k].height = H;
                    }
                    const uint32_t ncc = nc < MAX_COMPONENTS_NUM ? (uint32_t)nc : MAX_COMPONENTS_NUM;
                    if (proxy != proxy_mode_full) {
                        const uint32_t ss = proxy == proxy_mode_half ? 1u : 2u;
                        cfg->width = (W + (1u << ss) - 1) >> ss;
                        cfg->height = (H + (1u << ss) - 1) >> ss;
                        for (uint32_t q = 0; q < ncc; q++) {
                            cfg->components[q].width = (cfg->components[q].width + (1u << ss) - 1) >> ss;
                            cfg->components[q].height = (cfg->components[q].height + (1u << ss) - 1) >> ss;
                        }
                    }
                    for (uint32_t q = 0; q < ncc; q++) {
                        const uint64_t b = (uint64_t)cfg->components[q].width * cfg->components[q].height * (depth <= 8 ? 1u : 2u);
                        if (b > 0xFFFFFFFFull)
                            return SvtJxsErrorDecoderInvalidBitstream;
                        cfg->components[q].byte_size = (uint32_t)b;
                    }
                }
                if (fast_search && lcod != 0) {
                    *frame_size = lcod;
                    return SvtJxsErrorNone;
                }
            }
            break;
        case M_EOC:
            off += 2;
            *frame_size = off;
            return SvtJxsErrorNone;
        default: {
            if (!in_slices || bands < 0 || (m >> 12) != 0)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 3 > size)
                return SvtJxsErrorDecoderBitstreamTooShort;
            const uint32_t lprc = (m << 8) | buf[off + 2];
            if (lprc > kPrecMaxBytes)
                return SvtJxsErrorDecoderInvalidBitstream;
            off += lprc + 5 + ((uint32_t)bands * 2 + 7) / 8;
        }
        }
    }
}

int32_t cs_pi_check(const CsHeader& h, uint32_t proxy) {
    const CsHeader::Const& c = h.c;
    const uint32_t nc = c.nc, dh = (uint32_t)c.dh, dv = (uint32_t)c.dv, sd = (uint32_t)c.sd;
    const uint32_t slice_h = c.hsl << dv; /* hdr_Hsl * 2^NL,y, as a 32-bit product */
    if (sd > nc || dv > dh || dh > 5 || dv > 2 || nc > CS_MAX_COMPS || slice_h == 0)
        return SvtJxsErrorBadParameter;
    const uint32_t ph = 1u << dv;
    if (slice_h < (uint32_t)c.h && slice_h % ph)
        return SvtJxsErrorBadParameter;
    const uint32_t ndec = nc - sd;
    uint32_t cdv_of[CS_MAX_COMPS] = {0, 0, 0, 0}, cdh_of[CS_MAX_COMPS] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < nc; k++) {
        if (c.sx[k] < 1 || c.sy[k] < 1 || ph / c.sy[k] == 0)
            return SvtJxsErrorBadParameter;
        uint32_t cdv = 0, cdh = 0;
        if (k < ndec) {
            if (dv < c.sy[k] / 2)
                return SvtJxsErrorBadParameter;
            cdv = dv - c.sy[k] / 2;
            cdh = dh;
        }
        cdv_of[k] = cdv;
        cdh_of[k] = cdh;
        const uint32_t nb = 2 * cdv + cdh + 1;
        if (nb > 10)
            return SvtJxsErrorBadParameter;
        if (k >= ndec && nb != 1)
            return SvtJxsErrorBadParameter;
        /* every band of the component has a width and a height */
        uint32_t w = (uint32_t)c.w / c.sx[k], hh = (uint32_t)c.h / c.sy[k];
        for (uint32_t d = 0; d < cdv; d++) {
            if (w / 2 == 0 || hh / 2 == 0)
                return SvtJxsErrorBadParameter;
            w -= w / 2;
            hh -= hh / 2;
        }
        for (uint32_t d = cdv; d < cdh; d++) {
            if (w / 2 == 0)
                return SvtJxsErrorBadParameter;
            w -= w / 2;
        }
        if (w == 0 || hh == 0)
            return SvtJxsErrorBadParameter;
    }
    /* proxy decoding (Pi.c:567-626) */
    if (proxy != proxy_mode_full) {
        if (proxy >= proxy_mode_max || proxy > dv || proxy > dh)
            return SvtJxsErrorBadParameter;
        for (uint32_t k = 0; k < nc; k++)
            if (proxy > cdv_of[k] || proxy > cdh_of[k])
                return SvtJxsErrorBadParameter;
    }
    uint32_t sx[4] = {c.sx[0], c.sx[1], c.sx[2], c.sx[3]}, sy[4] = {c.sy[0], c.sy[1], c.sy[2], c.sy[3]};
    if (format_of(nc, sx, sy) == COLOUR_FORMAT_INVALID)
        return SvtJxsErrorDecoderInvalidBitstream;
    return SvtJxsErrorNone;
}

int32_t cs_slice_walk(const uint8_t* buf, size_t size, uint32_t hdr_end, uint32_t slices, uint32_t pps, uint32_t rows,
                      uint32_t nb_exist, uint32_t* prec_off, uint32_t* slice_off, uint32_t* end, uint32_t* bad_slice) {
    const uint32_t spb = (nb_exist * 2 + 7) / 8;
    uint64_t base = hdr_end;
    for (uint32_t s = 0; s < slices; s++) {
        *bad_slice = s;
        slice_off[s] = (uint32_t)base;
        const uint32_t first = s * pps;
        const uint32_t count = s == slices - 1 ? rows - first : pps;
        uint32_t found = 0;
        const uint8_t* b = buf + base;
        const uint64_t n = base <= size ? size - base : 0;
        uint64_t off = 0;
        bool in_slice = false, first_marker = true;
        if (n == 0)
            return SvtJxsErrorDecoderBitstreamTooShort;
        for (;;) {
            if (off + 2 > n)
                return SvtJxsErrorDecoderBitstreamTooShort;
            const uint32_t m = be16(b + off);
            if (m == M_SLH) {
                if (!first_marker)
                    return SvtJxsErrorDecoderInvalidBitstream;
                first_marker = false;
                if (off + 6 > n)
                    return SvtJxsErrorDecoderBitstreamTooShort;
                const uint32_t hs = be16(b + off + 2), idx = be16(b + off + 4);
                off += 2 + hs;
                if (idx != s)
                    return SvtJxsErrorDecoderInvalidBitstream;
                in_slice = true;
                continue;
            }
            if ((m >= M_SOC && m <= M_CRG) || m == M_CAP)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (!in_slice || (m >> 12) != 0)
                return SvtJxsErrorDecoderInvalidBitstream;
            if (off + 3 > n)
                return SvtJxsErrorDecoderBitstreamTooShort;
            const uint32_t lprc = (m << 8) | b[off + 2];
            if (lprc > kPrecMaxBytes)
                return SvtJxsErrorDecoderInvalidBitstream;
            prec_off[first + found] = (uint32_t)(base + off);
            off += lprc + 5 + spb;
            if (off > n)
                return SvtJxsErrorDecoderBitstreamTooShort;
            if (++found == count)
                break;
        }
        base += off;
    }
    *bad_slice = slices - 1;
    if (base + 1 < size) {
        if (be16(buf + base) != M_EOC)
            return SvtJxsErrorDecoderInvalidBitstream;
    } else {
        return SvtJxsErrorDecoderBitstreamTooShort;
    }
    *bad_slice = slices;
    *end = (uint32_t)(base + 2);
    return SvtJxsErrorNone;
}

}  // namespace jxs
