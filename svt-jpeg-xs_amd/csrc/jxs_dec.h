/*
 * jxs_dec.h -- decoder descriptors shared by the host code and the gfx950
 * decode kernels.
 */
#ifndef JXS_DEC_H
#define JXS_DEC_H

#include <stdint.h>

#include "jxs_geom.h"

/* One codestream of a decode launch. */
struct JxsDecFrame {
    const JxsGeom* g;
    const uint8_t* cs; /* codestream in HBM */
    uint32_t cs_len;
    uint32_t hdr_end; /* first slice header */
    uint32_t* prec_off; /* rows: byte offset of each precinct header */
    int16_t* coef;      /* band planes (JxsBand::coeff_off), two's complement */
    uint32_t* status;   /* non-zero: invalid codestream */
    uint32_t prec_base; /* first precinct of this frame in the launch */
    const uint32_t* prec_in; /* known precinct offsets (stride prec_in_stride), or NULL: walk the codestream */
    uint32_t prec_in_stride;
    uint32_t pkt_stride; /* u32 words per precinct in pkt */
    uint32_t npk_dec;    /* packets of each precinct decoded (fewer for a proxy decode) */
    uint32_t slice_base; /* first slice of this frame in the launch */
    uint32_t* pkt;       /* per precinct: JXS_DPKT_HDR words of band modes and flags, then 5 words per packet */
    uint8_t* gcli;       /* decoded bit-plane counts (JxsBand::gcli_off layout), for slices decoded in order */
    uint32_t rows_valid; /* precinct rows decoded; the rows after them are skipped (their slices failed on the host) */
};

/* Packet table of one precinct (written by k_dec_packets, read by
 * k_dec_unpack): words 0..19 hold one u16 per (component, band):
 * D-mode | GTLI << 8; word 20 is non-zero when the precinct is invalid;
 * word 21 is non-zero when the precinct uses vertical prediction or the
 * codestream sends signs apart (its slice is then decoded in order by
 * k_dec_slice); then per packet {significance bit, GCLI bit, data bit, data
 * end bit, flags} with flags bit 0 = the packet exists in this precinct,
 * bit 1 = raw mode. */
#define JXS_DPKT_BAD 20
#define JXS_DPKT_ORDERED 21
#define JXS_DPKT_HDR 24
#define JXS_DPKT_WORDS 5

/* Inverse horizontal job: one block per row.  Inputs are the final low band
 * (int16 band or int32 plane) and the high bands of `levels` levels (finest
 * first); the output is an int32 row or an output-plane row. */
struct JxsInvRowJob {
    const int16_t* lf16; /* low band (if lf32 == NULL), stride lf_stride */
    const int32_t* lf32;
    uint32_t lf_stride;
    uint32_t levels, w, rows, base;
    const int16_t* hf[5];
    uint32_t hf_stride[5];
    int32_t* out32; /* int32 output rows (stride out_stride), or NULL */
    void* out;      /* output plane (u8 / u16), when out32 == NULL */
    uint32_t out_stride, depth;
    uint32_t fq; /* fractional bits of the coefficients (Fq): band samples enter the lifting scaled by 2^fq */
    uint32_t bw; /* nominal coefficient precision (Bw) of the output scaling */
};

/* Final vertical level of a component with the finest horizontal level of
 * its low and high rows.  Low row i at level-1 resolution comes from low32
 * (int32, the coarser levels already inverted, or the next vertical level's
 * plane) or low16 (a band, scaled by 2^fq); hl its level-1 high band; high row
 * i from bands lh (low) and hh (high).  One thread per output column pair, a
 * block of 256 pairs by `seg` row pairs. */
struct JxsInvLastJob {
    const int32_t* low32;
    const int16_t* low16;
    const int16_t* hl;
    const int16_t* lh;
    const int16_t* hh;
    void* out;
    uint32_t low_stride, hl_stride, lh_stride, hh_stride;
    uint32_t out_stride, depth, bw, fq;
    uint32_t w, h, seg, tiles_x, nblk, base;
};

/* Inverse vertical job: int32 low rows then high rows (stride w) -> int32
 * plane or output plane. */
struct JxsInvVertJob {
    const int32_t* src;
    int32_t* out32;
    void* out;
    uint32_t w, h, out_stride, depth;
    uint32_t tiles_x, tiles, base, bw;
};

#endif
