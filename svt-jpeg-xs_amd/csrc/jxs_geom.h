/*
 * jxs_geom.h -- per-stream geometry and per-launch job descriptors shared by
 * the host code and the gfx950 kernels.
 *
 * A JxsGeom holds everything the kernels need to know about one coding
 * configuration (band layout, packet inclusion, weights, rate-control
 * bounds, frame header, buffer offsets).  It is built once on the host at
 * encoder/decoder init and kept resident in HBM; batched launches carry one
 * geometry pointer per frame, so frames of different resolution, format and
 * depth share a launch.
 */
#ifndef JXS_GEOM_H
#define JXS_GEOM_H

#include <stdint.h>

#define JXS_MAXC 4
#define JXS_NBC 10  /* bands per component (2*V + H + 1 <= 10) */
#define JXS_NGB 40  /* global bands */
#define JXS_NPK 30  /* packets per precinct */
#define JXS_NBL 64  /* band lines per precinct (<= 4 components x 13) */
#define JXS_NQ 32   /* quantisation values 0..31 */
#define JXS_NR 32   /* refinement values 0..31 (<= bands - 1) */
#define JXS_LUT 52  /* u32 per band-line rate-control table */
#define JXS_HDR_MAX 256
#define JXS_MAX_ROW 7680 /* widest component row the row kernels stage in LDS */

/* Offsets inside one band-line rate-control table (u32 units). */
#define LUT_CNT 0   /* [16] groups with gcli <= g */
#define LUT_DATA 16 /* [16] sum over gcli > g of (gcli - g + 1) */
#define LUT_SIG 32  /* [16] full significance groups with max <= g */
#define LUT_TAIL 48 /* max of the partial significance group, or -1 */

enum { JXS_SRC_U8 = 0, JXS_SRC_U16 = 1, JXS_SRC_I32 = 2 };

struct JxsBand {
    uint32_t w, h, hl, gid;
    uint32_t gw, sw; /* groups of 4 coefficients, significance groups of 8 groups */
    uint32_t gain, prio;
    uint32_t lines[2]; /* band lines in a normal / the last precinct row */
    uint32_t bl;       /* first band-line index of this band inside a precinct */
    uint32_t pad_;
    uint64_t coeff_off; /* u16 elements into the frame's coefficient buffer */
    uint64_t gcli_off;  /* bytes into the frame's GCLI buffer */
    uint64_t smax_off;  /* bytes into the frame's significance-max buffer */
};

struct JxsComp {
    uint32_t sx, sy, w, h, dv, dh, nb, pad_;
    JxsBand b[JXS_NBC];
};

struct JxsGeom {
    uint32_t nc, w, h, depth, format, packed;
    uint32_t dv, dh, rows, pps, slices;
    uint32_t nb_all, nb_exist, npk, short_hdr, qpih, sig;
    uint32_t maxq, maxr, hdr_len, lcod;
    uint32_t prec_hdr_bytes; /* precinct header: ceil((40 + 2 * nb_exist) / 8) */
    uint32_t pkt_hdr_bytes;  /* 5 (short) or 7 (long) */
    uint32_t rc_hdr_bytes[2];
    uint32_t n_bl; /* band lines per precinct */
    uint32_t pk_exist[2];
    uint32_t empty_mask[JXS_NQ]; /* bit r set when (q, r) truncates every band to 15 */
    JxsComp c[JXS_MAXC];
    uint8_t gc[JXS_NGB], gb[JXS_NGB];
    uint32_t pk_bs[JXS_NPK], pk_be[JXS_NPK], pk_line[JXS_NPK];
    uint32_t raw_bytes[2][JXS_NPK];
    uint64_t n_coeff, n_gcli, n_smax; /* buffer sizes per frame */
    uint64_t n_i32;                   /* int32 transform scratch per frame */
    uint8_t hdr[JXS_HDR_MAX];
};

/* Per-precinct rate-control result. */
struct JxsPrec {
    uint32_t q, r, total, pad, offset, err;
};

/* Vertical 5/3 lifting job: columns of a w x h plane -> low rows then high
 * rows (int32, stride w) at dst. */
struct JxsVertJob {
    const void* src;
    int32_t* dst;
    uint32_t kind, stride, step, w, h;
    uint32_t shift, mask;
    int32_t offset;
    uint32_t tiles_x, tiles, base; /* blocks of this job: [base, base + tiles) */
    uint32_t pad_;
};

struct JxsBandOut {
    uint16_t* coeff;
    uint8_t* gcli;
    uint8_t* smax;
    uint32_t w, gw, sw, pad_;
};

/* Horizontal multi-level 5/3 lifting job: one block per row. */
struct JxsHorzJob {
    const void* src;
    uint32_t kind, stride, step, w, rows, levels;
    uint32_t shift, mask;
    int32_t offset;
    uint32_t base; /* first block of this job */
    uint32_t lf_is_band;
    uint32_t lf_stride;
    int32_t* lf_i32;
    JxsBandOut lf;
    JxsBandOut hf[5];
};

/* One frame of a (possibly heterogeneous) launch. */
struct JxsFrameDesc {
    const JxsGeom* g;          /* device copy of the geometry */
    const uint32_t* slice_off; /* codestream offset of each slice header */
    const uint32_t* slice_bud; /* precinct budget of each slice */
    uint16_t* coeff;
    uint8_t* gcli;
    uint8_t* smax;
    uint32_t* lut;   /* rows x n_bl x JXS_LUT */
    uint32_t* table; /* rows x JXS_NQ x JXS_NR */
    JxsPrec* prec;   /* rows */
    uint8_t* out;    /* codestream */
    uint32_t* status;
    uint32_t rows, npk;
    uint32_t prec_base; /* first precinct of this frame in the launch */
    uint32_t slice_base;
};

#endif
