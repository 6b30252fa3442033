/*
 * jxs_codestream.h -- host-side codestream checks of the decoder: header
 * parse, frame-size walk, layout validity and per-slice precinct walk, each
 * giving the reference decoder's error code for a damaged or truncated
 * codestream (ParseHeader.c, DecThreadInit.c:35-191, Pi.c:189-405,
 * Decoder.c:51-142).
 */
#ifndef JXS_CODESTREAM_H
#define JXS_CODESTREAM_H

#include <stddef.h>
#include <stdint.h>

#include "SvtJpegxs.h"

namespace jxs {

#define CS_MAX_COMPS 4
#define CS_MAX_BANDS 40 /* 10 bands per component x 4 components */
#define CS_MAX_CAP 16   /* capability bits kept */

/* A parsed header.  `c` holds every field whose change between frames needs a
 * new decoder (picture_header_const_t, Pi.h:31-61): two frames belong to one
 * decoder when their `c` parts are byte-equal.  `d` holds the fields a frame
 * may change (picture_header_dynamic_t, Pi.h:63-80). */
struct CsHeader {
    struct Const {
        uint8_t has_cap, has_pih, has_cdt, has_wgt;
        uint32_t ppih, plev;
        int32_t w, h;
        uint32_t cw, hsl;
        uint8_t nc;
        int32_t dh, dv, sd;
        uint8_t cpih;
        uint32_t sx[CS_MAX_COMPS], sy[CS_MAX_COMPS];
        uint8_t depth[CS_MAX_COMPS];
        uint8_t cap[CS_MAX_CAP];
        int32_t cgs, sgs;
        uint8_t gain[CS_MAX_BANDS], prio[CS_MAX_BANDS];
    } c;
    struct Dyn {
        uint8_t has_cts, has_crg;
        uint32_t lcod;
        uint8_t bw, fq, br, fslc, ppoc, lh, rl, qpih, fs, rm, tnlt;
    } d;
};

/* The header of the codestream at cs (len bytes): 0 or the reference's error
 * code; *used = the offset of the first slice header. */
int32_t cs_get_header(const uint8_t* cs, size_t len, CsHeader& h, uint32_t* used);

/* Whether two headers describe the same decoder configuration. */
bool cs_same_config(const CsHeader& a, const CsHeader& b);

/* svt_jpeg_xs_decoder_get_single_frame_size_with_proxy: the size of the frame
 * at buf, from its markers and precinct lengths; cfg (optional) receives the
 * image configuration at the proxy resolution. */
int32_t cs_frame_size(const uint8_t* buf, size_t size, svt_jpeg_xs_image_config_t* cfg, uint32_t* frame_size,
                      uint32_t fast_search, uint32_t proxy);

/* The layout rules a decoder applies to a parsed header before it decodes
 * anything (band counts, sampling, slice height, non-empty bands, proxy
 * levels, colour format). */
int32_t cs_pi_check(const CsHeader& h, uint32_t proxy);

/* The precincts of every slice found by their Lprc fields, slice after slice
 * from hdr_end, then the EOC marker.  prec_off[rows] receives each precinct's
 * header offset, slice_off[slices] each slice header's offset, *end the
 * offset after EOC.  On an error, *bad_slice is the slice it belongs to (the
 * last slice for a missing EOC); on success it is `slices`.  The walk checks
 * the slice marker and index but reads past the header by its length field;
 * the slice decode then requires that field to be 4 (cs_slice_header_len). */
int32_t cs_slice_walk(const uint8_t* buf, size_t size, uint32_t hdr_end, uint32_t slices, uint32_t pps, uint32_t rows,
                      uint32_t nb_exist, uint32_t* prec_off, uint32_t* slice_off, uint32_t* end, uint32_t* bad_slice);

/* The length field of the slice header at slice_off (6 bytes there). */
uint32_t cs_slice_header_len(const uint8_t* buf, uint32_t slice_off);

}  // namespace jxs

#endif
