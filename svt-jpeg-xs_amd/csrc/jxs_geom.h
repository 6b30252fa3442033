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
#define JXS_HDR_MAX 256
#define JXS_NU 64  /* packing units (band lines) per precinct: one per lane of a wave */
#define JXS_MAX_ROW 8192 /* widest component row: the row kernels stage one in LDS, the transform covers
                            one with 256 threads of 32 columns */
#define JXS_SHADOW_UNITS 32768 /* packing units of a launch above which boundary words go through shadow words */

/* Band-line histogram record (u16 units), written by the transform for
 * every band line of every precinct:
 *   [0, 16)   groups with GCLI == v
 *   [16, 31)  full significance groups (8 GCLIs) with maximum == v, v < 15
 *   31        maximum of the trailing partial significance group, or 0xFFFF */
#define JXS_HREC 32
#define HREC_TAIL 31

enum { JXS_SRC_U8 = 0, JXS_SRC_U16 = 1, JXS_SRC_I32 = 2 };

struct JxsBand {
    uint32_t w, h, hl, gid;
    uint32_t gw, sw; /* groups of 4 coefficients, significance groups of 8 groups */
    uint32_t gain, prio;
    uint32_t lines[2]; /* band lines in a normal / the last precinct row */
    uint32_t bl;       /* first band-line index of this band inside a precinct */
    uint32_t cs;       /* coefficient row stride in elements (4 * gw: every group 8-byte aligned) */
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
    uint8_t ex_gid[JXS_NGB]; /* global band index of the e-th existing band */
    uint32_t pk_exist[2];
    uint32_t empty_mask[JXS_NQ]; /* bit r set when (q, r) truncates every band to 15 */
    JxsComp c[JXS_MAXC];
    uint8_t gc[JXS_NGB], gb[JXS_NGB];
    uint32_t pk_bs[JXS_NPK], pk_be[JXS_NPK], pk_line[JXS_NPK];
    uint32_t raw_bytes[2][JXS_NPK];
    /* packing units of a precinct: the band lines in packet order (u_p
     * packet, u_gi global band, u_info the unit word below) */
    uint32_t n_units, pad_u[3];
    uint8_t u_p[JXS_NU], u_gi[JXS_NU];
    uint8_t u_rank[JXS_NU]; /* place of the unit in the frame's unit array: widest lines first */
    uint32_t u_info[JXS_NU];
    uint32_t ln_info[JXS_NBL]; /* per band line of a precinct, the line word below */
    /* the rate control's walk over the band lines present in a normal / the
     * last precinct row, in packet order: walk words (below) and the RAW
     * bytes of each line's packet; zero past the end */
    uint32_t rq[2][JXS_NU], rq_raw[2][JXS_NU];
    uint32_t rq_n[2];
    /* the walk of a normal / the last precinct row cut into 2^lg runs of
     * whole packets (lg 0..2): run k is walk entries [rq_seg[t][lg][k],
     * rq_seg[t][lg][k + 1]) */
    uint8_t rq_seg[2][3][8];
    uint32_t rc_mode; /* rate control: 0 budget per precinct, 1 unused padding moved to the next precinct,
                         2 / 3 budgets from one (Q, R) per slice (3: with a minimum precinct size) */
    uint32_t fs;      /* Fs: signs in sub-packets of their own */
    uint32_t rm;      /* Rm: vertically predicted significance groups are runs of zero coefficients */
    uint32_t vpred;   /* encoder: coding_vertical_prediction_mode (0 off, 1 zero residuals, 2 zero coefficients) */
    uint32_t signs;   /* encoder: coding_signs_handling (0 off, 1 fast, 2 full) */
    uint32_t pad_t;
    uint64_t n_coeff, n_gcli, n_smax; /* buffer sizes per frame */
    uint64_t n_i32;                   /* int32 transform scratch per frame */
    uint8_t hdr[JXS_HDR_MAX];
};

/* Unit word (u_info): band line inside the precinct (bits 0-7), band gain
 * (8-15) and priority (16-23), line present in a normal / in the last
 * precinct row (24, 25), last band line of its packet (26), packet (27-31). */
#define UI_NORMAL 24
#define UI_LAST 26
#define UI_PKT 27
/* Walk word (rq): the line's offset in a rate table, line * JXS_RT_STRIDE
 * (bits 0-10), band gain (11-18) and priority (19-26), last present line of
 * its packet (27). */
#define JXS_RT_STRIDE 17 /* words per band line of a rate table: truncations 0..15, then a zero */
#define RQ_GAIN 11
#define RQ_PRIO 19
#define RQ_END 27
/* Line word (ln_info): groups of 4 coefficients (bits 0-11), second line of
 * its band (12), lines of its band in a normal / the last precinct row
 * (13-14, 15-16), significance groups (17-24). */
#define LI_SECOND 12
#define LI_NL 13
#define LI_SW 17

/* kernels stage the geometry in LDS with 16-byte copies */
static_assert(sizeof(JxsGeom) % 16 == 0, "JxsGeom size must be a multiple of 16 bytes");

/* Per-precinct rate-control result. */
struct JxsPrec {
    uint32_t q, r, total, pad, offset, err;
};

/* What the rate control of a precinct needs from its rate table, for the
 * in-order pass over a slice (rate control modes 0 and 1): the precinct's
 * bytes at (q, 0) for every q, at (q0, r) and (q0 - 1, r) for every r, where
 * q0 is the Q its budget gets without a padding carry, the (Q, R) = (q0, r0)
 * and bytes at that budget, and the smallest of the entries the two binary
 * searches look at that exceeds that budget (any budget below it gets the
 * same (Q, R); 0 when the searches failed there). */
struct JxsRateRec {
    uint32_t col[32];
    uint32_t row[64];
    uint32_t q0, r0, data0, thr;
};

/* Packet flags of the packet plan: the packet has a band line, and its
 * significance / bit-plane-count sub-packets are sent RAW. */
enum { PKT_ANY = 1u, PKT_RAW = 2u };

/* One packing unit = one band line of one precinct, self-contained
 * (written by the packet plan, read by the packing kernel with a single
 * load): where the line's significance, bit-plane-count and data segments
 * go and what produces them; the first band line of a packet also carries
 * the packet header. */
struct JxsUnit {
    const uint8_t* gl;  /* GCLI row of the band line */
    const uint8_t* sl;  /* significance-maximum row */
    const uint16_t* cl; /* coefficient row */
    uint32_t* status;   /* the frame's error flags */
    uint32_t* ob[3];    /* output dword holding the first bit of each segment */
    uint32_t* hob;      /* output dword holding the packet header's first bit */
    uint64_t hdr;       /* packet header bits */
    uint32_t lo[3], len[3]; /* first bit inside *ob (0..31) and bits of each segment */
    uint32_t hlo, hlen;     /* the same for the packet header (hlen 0: none) */
    uint32_t gt, ms, raw, qpih;
    uint32_t w, gw, sw, pad_;
    int64_t sdelta; /* the shadow word of an output dword lies sdelta bytes after it */
};

struct JxsBandOut {
    uint16_t* coeff;
    uint8_t* gcli;
    uint8_t* smax;
    uint16_t* hist; /* frame's histogram records, or null */
    uint32_t w, gw, sw, cs;
    uint32_t hl, bl, n_bl, pad_; /* record of band row y: (y / hl) * n_bl + bl + y % hl */
};

/* One forward-transform step of one component plane (w x h samples at src):
 * with `vert` set, one vertical 5/3 level followed by horizontal levels of
 * the low rows (lev_l) and one horizontal level of the high rows; without,
 * horizontal levels of every row.  Blocks [base, base + nblk), rpb output
 * rows (pairs) per block, one thread per `chunk` columns. */
struct JxsDwtJob {
    const void* src;
    uint32_t kind, stride, step, w, h;
    uint32_t shift, mask;
    int32_t offset;
    uint32_t vert, lev_l;
    uint32_t base, rows; /* first block; output rows (row pairs with `vert`) */
    uint32_t chunk;      /* columns per thread: 16 or 32 (>= 2^lev_l) */
    uint32_t tpr;        /* threads per row: a multiple of 64, >= w / chunk */
    uint32_t rpb;        /* rows per block: 256 / tpr */
    uint32_t nblk;       /* blocks of the job */
    uint32_t lf_is_band; /* low rows' low-pass output: band, or int32 plane */
    uint32_t lf_stride;
    uint32_t vec; /* 1: rows readable 16 bytes per load (step 1, plane and row pitch 16-byte aligned); 2 + c: packed 16-bit rows so readable, component c */
    /* il > 1: jobs il_first .. il_first + il - 1 of the launch's list (the
     * components of one packed plane) share their blocks interleaved, block
     * b of the group working on job b % il, row b / il, so that the blocks
     * reading the same packed rows run together on one XCD */
    uint32_t il, il_first;
    int32_t* lf_i32;
    JxsBandOut lf, hf[5]; /* low rows: final low-pass and high-pass of each level */
    JxsBandOut lh, hh;    /* high rows: low-pass and high-pass of their level */
};

/* One frame of a (possibly heterogeneous) launch. */
struct JxsFrameDesc {
    const JxsGeom* g;          /* device copy of the geometry */
    const uint32_t* slice_off; /* codestream offset of each slice header */
    const uint32_t* slice_bud; /* precinct budget of each slice */
    uint16_t* coeff;
    uint8_t* gcli;
    uint8_t* smax;
    uint16_t* hist;  /* rows x n_bl x JXS_HREC */
    JxsPrec* prec;   /* rows */
    JxsRateRec* rrec; /* rows */
    JxsUnit* units;  /* packing units: unit u of precinct row r at u_rank[u] * rows + r */
    uint8_t* out;    /* codestream */
    uint32_t* status;
    uint32_t* shadow; /* one word per codestream dword: bits of words several packing units share */
    uint32_t rows, npk;
    uint32_t prec_base; /* first precinct of this frame in the launch */
    uint32_t slice_base;
    uint32_t unit_base; /* first packing unit (rows x n_units numbering) of this frame in the launch */
    uint32_t n_units;
    uint32_t pps, slices, n_bl; /* copies of the geometry's, readable before it is staged */
    uint32_t ordered;           /* coded by k_enc_slice (vertical prediction or sign coding), not k_rate / k_pack */
    uint32_t slice_rc;          /* rate control modes 2 / 3: one (Q, R) per slice */
};

#endif
