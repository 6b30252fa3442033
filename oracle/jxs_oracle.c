/*
 * jxs_oracle.c -- CPU restatement of the SVT-JPEG-XS hot path (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see jxs_oracle.h).  Scalar, single-threaded,
 * written for clarity and exactness, not speed.  The forward and inverse
 * 5/3 transforms are computed on whole components (the full-frame form the
 * reference's own tests prove equal to its precinct-streamed form:
 * tests/UnitTests/TestDwtFrame.cc:938-1040, TestInvDwtFrame.cc:652-720).
 * Everything else follows the reference per precinct.
 *
 * Pinning: see tests/test_oracle_*.py -- this file is checked against the
 * reference library built from source (oracle/Makefile.ref) on byte-exact
 * codestreams and bit-exact decoded planes, against the quantisation
 * known-answer vectors of tests/UnitTests/TestQuant.cc:27-136, against the
 * binary-search == linear-search property of TestReateControl.cc, and the
 * transform against perfect reconstruction.
 */
#include "jxs_oracle.h"

#include <stdlib.h>
#include <string.h>

#define E_NONE 0
#define E_BAD_PARAM ((int32_t)0x80001005)
#define E_UNDEFINED ((int32_t)0x80001001)
#define E_ENCODE_FRAME ((int32_t)0x80002035)
#define E_INVALID_BS ((int32_t)0x80003001)
#define E_TOO_SHORT ((int32_t)0x80003003)
#define E_INSUFFICIENT ((int32_t)0x80001000)

#define NBC 10 /* bands per component */
#define NGB 40 /* global bands */
#define NPK 30 /* packets */
#define NE 0xFF
#define TRUNC_MAX 15
#define BW 20
#define FQ 8

enum { FMT_INVALID = 0, FMT_YUV400 = 1, FMT_YUV420 = 2, FMT_YUV422 = 3, FMT_YUV444 = 4, FMT_4COMP = 5, FMT_GRAY = 6,
       FMT_PACKED444 = 21 };

typedef struct {
    uint32_t w, h, hl, gid;
    uint8_t gain, prio;
} ob_band;

typedef struct {
    uint32_t sx, sy, w, h, dv, dh, nb;
    ob_band b[NBC];
} ob_comp;

typedef struct {
    uint32_t nc, w, h, dv, dh, nb_all, nb_exist, prec_h, slice_h, pps, rows, slices, short_hdr;
    ob_comp c[4];
    uint8_t gc[NGB], gb[NGB];
    uint32_t npk, pk_bs[NPK], pk_be[NPK], pk_line[NPK];
    uint32_t lines[2][4][NBC]; /* band lines in a normal / the last precinct row */
    int32_t pk_exist[2];
    uint32_t raw_bytes[2][NPK];
    uint8_t maxq, maxr;
    /* encoder side */
    uint32_t depth, format, lcod, qpih, sig, hdr_len;
    uint32_t rc_mode; /* 0: budget per precinct; 1: unused padding moves to the next precinct of the slice;
                         2, 3: one Q and R for the slice sizes the precinct budgets (3: with a minimum size) */
    uint32_t vpred;   /* vertical prediction: 0 off, 1 zero residuals (Rm = 0), 2 zero coefficients (Rm = 1) */
    uint32_t signs;   /* sign coding: 0 with the data (Fs = 0), 1 fast, 2 full (Fs = 1) */
    uint8_t hdr[512];
    uint32_t* slice_sizes;
} ob_pi;

static inline uint32_t gcli_w(uint32_t w) { return (w + 3) / 4; }
static inline uint32_t sig_w(uint32_t w) { return (gcli_w(w) + 7) / 8; }
static inline uint32_t bytes_of_bits(uint32_t b) { return (b + 7) >> 3; }

/* ----------------------------------------------------------------------
 * Sampling factors: Pi.c:521-565 (format_get_sampling_factory)
 * ---------------------------------------------------------------------- */
static int sampling(uint32_t fmt, uint32_t* nc, uint32_t sx[4], uint32_t sy[4]) {
    switch (fmt) {
    case FMT_YUV422: *nc = 3; sx[0] = 1; sx[1] = sx[2] = 2; sy[0] = sy[1] = sy[2] = 1; return 0;
    case FMT_YUV420: *nc = 3; sx[0] = sy[0] = 1; sx[1] = sx[2] = sy[1] = sy[2] = 2; return 0;
    case FMT_YUV444:
    case FMT_PACKED444: *nc = 3; sx[0] = sx[1] = sx[2] = sy[0] = sy[1] = sy[2] = 1; return 0;
    case FMT_4COMP: *nc = 4; for (int i = 0; i < 4; i++) sx[i] = sy[i] = 1; return 0;
    case FMT_GRAY: *nc = 1; sx[0] = sy[0] = 1; return 0;
    default: return -1;
    }
}

/* ----------------------------------------------------------------------
 * Band topology, packets and precinct geometry: Pi.c:189-406 (pi_compute),
 * Pi.c:131-187 (calc_packet_inclusion), Pi.c:21-129 (calc_precinct_dimension,
 * the Cw = 0 branch).  One precinct column, no suppressed components.
 * ---------------------------------------------------------------------- */
static int pi_build(ob_pi* pi, uint32_t nc, const uint32_t sx[4], const uint32_t sy[4], uint32_t W, uint32_t H, uint32_t dh,
                    uint32_t dv, uint32_t slice_h) {
    if (dv > dh || dh > 5 || dv > 2 || nc > 4 || slice_h == 0)
        return -1;
    pi->nc = nc;
    pi->w = W;
    pi->h = H;
    pi->dh = dh;
    pi->dv = dv;
    pi->prec_h = 1u << dv;
    if (slice_h < H && (slice_h % pi->prec_h))
        return -1;
    pi->slice_h = slice_h;
    pi->pps = (slice_h + pi->prec_h - 1) / pi->prec_h;
    pi->slices = (H + slice_h - 1) / slice_h;
    pi->nb_exist = 0;
    for (uint32_t c = 0; c < nc; c++) {
        ob_comp* cp = &pi->c[c];
        cp->sx = sx[c];
        cp->sy = sy[c];
        cp->w = W / sx[c];
        cp->h = H / sy[c];
        if (pi->prec_h / sy[c] == 0)
            return -1;
        if (dv < sy[c] / 2)
            return -1;
        cp->dh = dh;
        cp->dv = dv - sy[c] / 2;
        cp->nb = 2 * cp->dv + cp->dh + 1;
        if (cp->nb > NBC)
            return -1;
        pi->nb_exist += cp->nb;
    }
    pi->nb_all = pi->c[0].nb * nc;
    const uint32_t nb_global = 2 * dv + dh + 1;
    memset(pi->gc, NE, sizeof(pi->gc));
    memset(pi->gb, NE, sizeof(pi->gb));
    for (uint32_t c = 0; c < nc; c++) {
        ob_comp* cp = &pi->c[c];
        uint32_t ignore = nb_global - cp->nb;
        uint32_t at_end = 3 * cp->dv;
        uint32_t gid = c;
        for (uint32_t i = 0; i < cp->nb; i++) {
            cp->b[i].gid = gid;
            pi->gc[gid] = (uint8_t)c;
            pi->gb[gid] = (uint8_t)i;
            gid += nc;
            if (ignore > 0 && i == cp->nb - at_end + 1 - ignore)
                gid += nc * ignore;
        }
    }
    /* band sizes: each vertical level splits the current low band in four,
     * each horizontal-only level in two, with ceil/floor halves. */
    for (uint32_t c = 0; c < nc; c++) {
        ob_comp* cp = &pi->c[c];
        uint32_t lvl = cp->nb - 1;
        uint32_t w = cp->w, h = cp->h;
        for (uint32_t d = 0; d < cp->dv; d++) {
            uint32_t w2 = w / 2, w1 = w - w2, h2 = h / 2, h1 = h - h2;
            cp->b[lvl].w = w2;     cp->b[lvl].h = h2;     /* HH */
            cp->b[lvl - 1].w = w1; cp->b[lvl - 1].h = h2; /* LH: low horizontal, high vertical */
            cp->b[lvl - 2].w = w2; cp->b[lvl - 2].h = h1; /* HL */
            lvl -= 3;
            w = w1;
            h = h1;
        }
        for (uint32_t d = cp->dv; d < cp->dh; d++) {
            uint32_t w2 = w / 2, w1 = w - w2;
            cp->b[lvl].w = w2;
            cp->b[lvl].h = h;
            lvl -= 1;
            w = w1;
        }
        cp->b[0].w = w;
        cp->b[0].h = h;
        for (uint32_t b = 0; b < cp->nb; b++) {
            if (cp->b[b].w == 0 || cp->b[b].h == 0)
                return -1;
            cp->b[b].hl = (cp->dv == 2 && b >= cp->nb - 3) ? 2 : 1;
        }
    }
    /* packets */
    uint32_t n = 0;
    uint32_t first_stop = (dh > dv ? dh - dv : dv - dh) + 1;
    pi->pk_line[n] = 0;
    pi->pk_bs[n] = 0;
    pi->pk_be[n] = first_stop * nc;
    n++;
    for (uint32_t lv = first_stop; lv < pi->c[0].nb; lv++) {
        pi->pk_line[n] = 0;
        pi->pk_bs[n] = lv * nc;
        pi->pk_be[n] = lv * nc;
        for (uint32_t i = lv * nc; i < (lv + 1) * nc; i++)
            if (pi->gb[i] != NE)
                pi->pk_be[n]++;
        n++;
    }
    for (uint32_t lv = first_stop; lv < pi->c[0].nb; lv++) {
        if (pi->c[0].b[lv].hl == 1)
            continue;
        pi->pk_line[n] = 1;
        pi->pk_bs[n] = lv * nc;
        pi->pk_be[n] = lv * nc;
        for (uint32_t i = lv * nc; i < (lv + 1) * nc; i++)
            if (pi->gb[i] != NE && pi->c[pi->gc[i]].b[pi->gb[i]].hl > 1)
                pi->pk_be[n]++;
        n++;
    }
    pi->npk = n;
    pi->short_hdr = (W * nc) < 32752;
    /* precinct rows and the lines of each band in a normal / last row */
    pi->rows = (H + pi->prec_h - 1) / pi->prec_h;
    for (uint32_t c = 0; c < nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            pi->lines[0][c][b] = pi->c[c].b[b].hl;
            pi->lines[1][c][b] = pi->c[c].b[b].h - (pi->rows - 1) * pi->c[c].b[b].hl;
        }
    for (int t = 0; t < 2; t++) {
        int32_t exist = 0;
        for (uint32_t p = 0; p < pi->npk; p++) {
            uint32_t raw_bits = 0;
            int any = 0;
            for (uint32_t g = pi->pk_bs[p]; g < pi->pk_be[p]; g++) {
                uint32_t c = pi->gc[g], b = pi->gb[g];
                if (pi->pk_line[p] < pi->lines[t][c][b]) {
                    any = 1;
                    raw_bits += gcli_w(pi->c[c].b[b].w) * 4;
                }
            }
            exist += any;
            pi->raw_bytes[t][p] = bytes_of_bits(raw_bits);
        }
        pi->pk_exist[t] = t == 0 ? (int32_t)pi->npk : exist;
    }
    return 0;
}

/* ----------------------------------------------------------------------
 * Weight tables: ISO/IEC 21122 default tables as carried by
 * WeightTable.c:18-74, selection WeightTable.c:336-413, and the derivation
 * for 4:2:2 / 4:4:4 at other V/H, WeightTable.c:128-334.
 * ---------------------------------------------------------------------- */
static const uint8_t g444_05[] = {3, 2, 2, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0};
static const uint8_t p444_05[] = {6, 8, 7, 1, 4, 5, 12, 14, 15, 0, 2, 3, 9, 11, 10, 13, 16, 17};
static const uint8_t g444_15[] = {4, 2, 2, 3, 2, 2, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
static const uint8_t p444_15[] = {21, 1, 0, 15, 19, 18, 5, 9, 8, 14, 17, 16, 2, 4, 3, 7, 13, 11, 6, 12, 10, 20, 23, 22};
static const uint8_t g444_25[] = {4, 3, 3, 3, 2, 2, 3, 2, 2, 2, 1, 1, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
static const uint8_t p444_25[] = {12, 15, 14, 3, 11, 10, 24, 26, 27, 0, 4, 5, 18, 21, 20,
                                  19, 23, 22, 13, 16, 17, 2, 9, 6, 1, 7, 8, 25, 28, 29};
static const uint8_t g422_05[] = {2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t p422_05[] = {8, 7, 6, 5, 3, 4, 13, 12, 14, 1, 0, 2, 9, 11, 10, 16, 15, 17};
static const uint8_t g422_15[] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t p422_15[] = {0, 1, 2, 18, 19, 20, 6, 7, 8, 15, 16, 17, 3, 4, 5, 10, 12, 14, 9, 11, 13, 21, 22, 23};
static const uint8_t g422_25[] = {3, 3, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t p422_25[] = {14, 13, 12, 9, 11, 10, 25, 24, 26, 0, 1, 2, 19, 20, 18,
                                  23, 22, 21, 17, 15, 16, 4, 8, 5, 3, 6, 7, 28, 27, 29};
static const uint8_t g420_11[] = {1, 1, 1, 1, 0, 0, 1, NE, NE, 0, NE, NE};
static const uint8_t p420_11[] = {2, 5, 4, 6, 1, 0, 7, NE, NE, 3, NE, NE};
static const uint8_t g420_12[] = {1, 1, 1, 1, 0, 0, 1, 0, 0, 1, NE, NE, 0, NE, NE};
static const uint8_t p420_12[] = {4, 7, 6, 8, 1, 0, 10, 3, 2, 9, NE, NE, 5, NE, NE};
static const uint8_t g420_13[] = {2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, NE, NE, 0, NE, NE};
static const uint8_t p420_13[] = {11, 3, 2, 6, 9, 8, 10, 1, 0, 13, 5, 4, 12, NE, NE, 7, NE, NE};
static const uint8_t g420_14[] = {2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, NE, NE, 0, NE, NE};
static const uint8_t p420_14[] = {7, 12, 11, 14, 3, 2, 6, 10, 9, 13, 1, 0, 16, 5, 4, 15, NE, NE, 8, NE, NE};
static const uint8_t g420_15[] = {3, 3, 3, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, NE, NE, 0, NE, NE};
static const uint8_t p420_15[] = {9, 16, 15, 7, 13, 12, 17, 3, 2, 6, 11, 10, 14, 1, 0, 19, 5, 4, 18, NE, NE, 8, NE, NE};
static const uint8_t g420_22[] = {2, 1, 1, 1, 1, 1, 1, NE, NE, 0, NE, NE, 1, 1, 1, 1, 1, 1, 0, 0, 0};
static const uint8_t p420_22[] = {10, 4, 3, 2, 9, 8, 1, NE, NE, 0, NE, NE, 16, 14, 12, 15, 13, 11, 7, 6, 5};
static const uint8_t g420_23[] = {2, 2, 2, 2, 1, 1, 1, 1, 1, 1, NE, NE, 0, NE, NE, 1, 1, 1, 1, 1, 1, 0, 0, 0};
static const uint8_t p420_23[] = {5, 12, 11, 13, 4, 3, 2, 10, 9, 1, NE, NE, 0, NE, NE, 19, 17, 15, 18, 16, 14, 8, 7, 6};
static const uint8_t g420_24[] = {3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, NE, NE, 0, NE, NE, 1, 1, 1, 1, 1, 1, 0, 0, 0};
static const uint8_t p420_24[] = {22, 6, 5, 7, 14, 13, 15, 4, 3, 2, 12, 11, 1, NE,
                                  NE, 0, NE, NE, 21, 19, 17, 20, 18, 16, 10, 9, 8};
static const uint8_t g420_25[] = {3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, NE, NE, 0, NE, NE, 1, 1, 1, 1, 1, 1, 0, 0, 0};
static const uint8_t p420_25[] = {0, 13, 12, 25, 7, 6, 8, 17, 16, 18, 5, 4, 3, 15, 14,
                                  2, NE, NE, 1, NE, NE, 24, 22, 20, 23, 21, 19, 11, 10, 9};

typedef struct {
    uint8_t g[NGB], p[NGB];
    uint32_t n;
} wtab;

static int wtab_pick(uint32_t fmt, uint32_t v, uint32_t h, wtab* t) {
    const uint8_t *g = NULL, *p = NULL;
    uint32_t n = 0;
#define PICK(G, P) do { g = G; p = P; n = sizeof(G); } while (0)
    if (h == 5 && (fmt == FMT_YUV444 || fmt == FMT_PACKED444)) {
        if (v == 0) PICK(g444_05, p444_05);
        else if (v == 1) PICK(g444_15, p444_15);
        else PICK(g444_25, p444_25);
    } else if (h == 5 && fmt == FMT_YUV422) {
        if (v == 0) PICK(g422_05, p422_05);
        else if (v == 1) PICK(g422_15, p422_15);
        else PICK(g422_25, p422_25);
    } else if (fmt == FMT_YUV420 && v >= 1 && h >= v) {
        static const uint8_t* const G1[6] = {NULL, g420_11, g420_12, g420_13, g420_14, g420_15};
        static const uint8_t* const P1[6] = {NULL, p420_11, p420_12, p420_13, p420_14, p420_15};
        static const uint32_t N1[6] = {0, 12, 15, 18, 21, 24};
        static const uint8_t* const G2[6] = {NULL, NULL, g420_22, g420_23, g420_24, g420_25};
        static const uint8_t* const P2[6] = {NULL, NULL, p420_22, p420_23, p420_24, p420_25};
        static const uint32_t N2[6] = {0, 0, 21, 24, 27, 30};
        if (v == 1) { g = G1[h]; p = P1[h]; n = N1[h]; }
        else { g = G2[h]; p = P2[h]; n = N2[h]; }
    }
#undef PICK
    if (!g)
        return -1;
    memcpy(t->g, g, n);
    memcpy(t->p, p, n);
    t->n = n;
    return 0;
}

/* Derived tables for 4:2:2 and 4:4:4 (WeightTable.c:128-235, 282-334): start
 * from the V2/H5 table, drop the groups of three entries that belong to the
 * removed vertical levels, then the removed horizontal levels, then close the
 * gaps in the priority order one step at a time. */
static void wtab_remove_group(wtab* t, uint32_t grp) {
    for (uint32_t i = 3 * grp; i + 3 < t->n; i++) {
        t->g[i] = t->g[i + 3];
        t->p[i] = t->p[i + 3];
    }
    t->n -= 3;
}

static int wtab_derive(uint32_t fmt, uint32_t dv, uint32_t dh, wtab* t) {
    if (fmt != FMT_YUV422 && fmt != FMT_YUV444 && fmt != FMT_PACKED444)
        return -1;
    if (wtab_pick(fmt, 2, 5, t))
        return -1;
    int v = 2, h = 5;
    if (h < v || (int)dv < 0)
        return -1;
    while (v > (int)dv) {
        v--;
        wtab_remove_group(t, (uint32_t)(h - 1));
        wtab_remove_group(t, 0);
    }
    if (h < (int)dv || (int)dh < (int)dv)
        return -1;
    while (h > (int)dh) {
        h--;
        if (h != 1)
            wtab_remove_group(t, 0);
        else
            t->n -= 3;
    }
    for (uint32_t i = 0; i < t->n; i++) {
        for (;;) {
            int found = 0;
            for (uint32_t j = 0; j < t->n; j++)
                if (t->p[j] == i) {
                    found = 1;
                    break;
                }
            if (found)
                break;
            for (uint32_t j = 0; j < t->n; j++)
                if (t->p[j] > i) {
                    t->p[j]--;
                    break;
                }
        }
    }
    return 0;
}

static int weights_assign(ob_pi* pi, uint32_t fmt) {
    wtab t;
    int have = 0;
    if (pi->nc == 3)
        have = wtab_pick(fmt, pi->dv, pi->dh, &t) == 0;
    if (!have) {
        if (pi->nc != 3 || wtab_derive(fmt, pi->dv, pi->dh, &t))
            return -1;
    }
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            if (pi->c[c].b[b].gid >= t.n)
                return -1;
            pi->c[c].b[b].gain = t.g[pi->c[c].b[b].gid];
            pi->c[c].b[b].prio = t.p[pi->c[c].b[b].gid];
        }
    return 0;
}

/* ----------------------------------------------------------------------
 * Truncation: EncDec.c:8-21 (compute_truncation); maximum Q / R:
 * PiEnc.c:11-53.
 * ---------------------------------------------------------------------- */
static inline uint8_t truncation(uint8_t gain, uint8_t prio, uint8_t q, uint8_t r) {
    uint8_t pump = prio < r ? 1 : 0;
    if (q < gain + pump)
        return 0;
    uint32_t t = (uint32_t)q - gain - pump;
    return (uint8_t)(t > TRUNC_MAX ? TRUNC_MAX : t);
}

/* 1 when every band would be truncated to TRUNC_MAX (an "empty" precinct). */
static int all_truncated(const ob_pi* pi, uint8_t q, uint8_t r) {
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++)
            if (truncation(pi->c[c].b[b].gain, pi->c[c].b[b].prio, q, r) != TRUNC_MAX)
                return 0;
    return 1;
}

static void max_qr(ob_pi* pi) {
    uint32_t mr = 0;
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++)
            if (pi->c[c].b[b].prio > mr)
                mr = pi->c[c].b[b].prio;
    if (mr < 255)
        mr++;
    if (mr > pi->nb_all - 1)
        mr = pi->nb_all - 1;
    pi->maxr = (uint8_t)mr;
    uint32_t mq = 31;
    while (mq > 0) {
        if (!all_truncated(pi, (uint8_t)mq, (uint8_t)mr))
            break;
        mq--;
    }
    pi->maxq = (uint8_t)mq;
}

/* ----------------------------------------------------------------------
 * Bit writer / reader: MSB-first, as BitstreamWriter.c / BitstreamReader.c.
 * The writer ORs into a zeroed buffer.
 * ---------------------------------------------------------------------- */
typedef struct {
    uint8_t* m;
    uint64_t bit;
} bw_t;

static inline void bw_put(bw_t* w, uint32_t v, uint32_t n) {
    for (int i = (int)n - 1; i >= 0; i--) {
        if ((v >> i) & 1)
            w->m[w->bit >> 3] |= (uint8_t)(0x80u >> (w->bit & 7));
        w->bit++;
    }
}
static inline void bw_align(bw_t* w) { w->bit = (w->bit + 7) & ~(uint64_t)7; }

typedef struct {
    const uint8_t* m;
    uint64_t bit, end_bit;
    int err;
} br_t;

static inline uint32_t br_get(br_t* r, uint32_t n) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (r->bit >= r->end_bit) {
            r->err = 1;
            return 0;
        }
        v = (v << 1) | ((r->m[r->bit >> 3] >> (7 - (r->bit & 7))) & 1);
        r->bit++;
    }
    return v;
}
static inline void br_align(br_t* r) { r->bit = (r->bit + 7) & ~(uint64_t)7; }

/* ----------------------------------------------------------------------
 * Codestream header: PackHeaders.c:11-140, written once per encoder
 * (PreRcStageProcess.c:23-29).
 * ---------------------------------------------------------------------- */
static uint32_t write_frame_header(ob_pi* pi, uint8_t* out) {
    bw_t w = {out, 0};
    memset(out, 0, 512);
    bw_put(&w, 0xFF10, 16); /* SOC */
    bw_put(&w, 0xFF50, 16); /* CAP */
    bw_put(&w, 4, 16);      /* Lcap: 16 + 9 capability bits, rounded to bytes */
    for (int i = 0; i < 9; i++) {
        uint32_t bit = 0;
        if (i == 4)
            bit = pi->format == FMT_YUV420;
        if (i == 8)
            bit = 1; /* Rl */
        bw_put(&w, bit, 1);
    }
    bw_align(&w);
    bw_put(&w, 0xFF12, 16); /* PIH */
    bw_put(&w, 26, 16);
    bw_put(&w, pi->lcod, 32);
    bw_put(&w, 0, 16); /* Ppih */
    bw_put(&w, 0, 16); /* Plev */
    bw_put(&w, pi->w, 16);
    bw_put(&w, pi->h, 16);
    bw_put(&w, 0, 16); /* Cw */
    bw_put(&w, pi->pps, 16);
    bw_put(&w, pi->nc, 8);
    bw_put(&w, 4, 8);  /* Ng */
    bw_put(&w, 8, 8);  /* Ss */
    bw_put(&w, BW, 8); /* Bw */
    bw_put(&w, FQ, 4);
    bw_put(&w, 4, 4); /* Br */
    bw_put(&w, 0, 8); /* Fslc, Ppoc, Cpih */
    bw_put(&w, pi->dh, 4);
    bw_put(&w, pi->dv, 4);
    bw_put(&w, !pi->short_hdr, 1);
    bw_put(&w, 1, 1); /* Rl */
    bw_put(&w, pi->qpih, 2);
    bw_put(&w, pi->signs ? 1u : 0u, 2);     /* Fs (EncHandle.c:274-278) */
    bw_put(&w, pi->vpred == 2 ? 1u : 0u, 2); /* Rm (EncHandle.c:250-253) */
    bw_put(&w, 0xFF13, 16); /* CDT */
    bw_put(&w, 2 * pi->nc + 2, 16);
    for (uint32_t c = 0; c < pi->nc; c++) {
        bw_put(&w, pi->depth, 8);
        bw_put(&w, pi->c[c].sx, 4);
        bw_put(&w, pi->c[c].sy, 4);
    }
    bw_put(&w, 0xFF14, 16); /* WGT */
    bw_put(&w, 2 * pi->nb_exist + 2, 16);
    for (uint32_t g = 0; g < pi->nb_all; g++) {
        if (pi->gc[g] == NE)
            continue;
        const ob_band* b = &pi->c[pi->gc[g]].b[pi->gb[g]];
        bw_put(&w, b->gain, 8);
        bw_put(&w, b->prio, 8);
    }
    bw_align(&w);
    return (uint32_t)(w.bit >> 3);
}

/* ----------------------------------------------------------------------
 * Encoder configuration: EncHandle.c:128-454 (validation, Lcod), 760-809
 * (slice sizes).
 * ---------------------------------------------------------------------- */
void jxso_default_params(jxso_params_t* p) {
    memset(p, 0, sizeof(*p));
    p->bpp_den = 1;
    p->decomp_v = 2;
    p->decomp_h = 5;
    p->quantization = 0;
    p->slice_height = 16;
    p->significance = 1;
    p->rc_mode = 1; /* RC_CBR_PER_PRECINCT_MOVE_PADDING (EncHandle.c:519) */
    p->vpred = 0;
    p->signs = 0;
}

static int32_t enc_configure(const jxso_params_t* prm, ob_pi* pi) {
    memset(pi, 0, sizeof(*pi));
    uint32_t nc, sx[4], sy[4];
    if (prm->depth < 8 || prm->depth > 14)
        return E_BAD_PARAM;
    if (prm->rc_mode > 3)
        return E_BAD_PARAM;
    /* EncHandle.c:242-278 */
    if (prm->vpred > 2 || prm->signs > 2)
        return E_BAD_PARAM;
    /* Not restated: full sign coding with the per-slice modes, where the
     * reference's per-band cache of two truncations carries sizes from the
     * slice-wide search (computed with fast sign coding,
     * RateControl.c:1336-1342) into the first precinct's own search
     * (RateControl.c:1285-1290 resets it only below another precinct). */
    if (prm->signs == 2 && prm->rc_mode >= 2)
        return E_BAD_PARAM;
    if (sampling(prm->format, &nc, sx, sy))
        return E_BAD_PARAM;
    if (prm->format == FMT_YUV400)
        return E_BAD_PARAM;
    if (prm->decomp_v == 0 && prm->format == FMT_YUV420)
        return E_BAD_PARAM;
    uint32_t slice_h = prm->slice_height;
    if (slice_h == 0)
        return E_BAD_PARAM;
    if (slice_h >= prm->height)
        slice_h = prm->height;
    else if (slice_h % (1u << prm->decomp_v))
        return E_BAD_PARAM;
    if (prm->quantization > 1 || prm->bpp_den == 0 || prm->bpp_num == 0)
        return E_BAD_PARAM;
    uint64_t bpf = ((uint64_t)prm->width * prm->height * prm->bpp_num / prm->bpp_den + 7) / 8;
    if (bpf >= ((uint64_t)1 << 32) || bpf == 0)
        return E_BAD_PARAM;
    if (prm->width < 4)
        return E_BAD_PARAM;
    if ((prm->format == FMT_YUV422 || prm->format == FMT_YUV420) && (prm->width % 2))
        return E_BAD_PARAM;
    if (prm->format == FMT_YUV420 && (prm->height % 2))
        return E_BAD_PARAM;
    if (prm->decomp_v > 2 || prm->decomp_h > 5 || prm->decomp_h < prm->decomp_v)
        return E_BAD_PARAM;
    if (prm->decomp_h == 0 && prm->decomp_v == 0)
        return E_BAD_PARAM;
    /* zero-sized bands (EncHandle.c:362-409) */
    uint32_t mw = prm->width, mh = prm->height, mv = prm->decomp_v;
    if (prm->format == FMT_YUV420) {
        mw >>= 1;
        mh >>= 1;
        mv -= 1;
    }
    if (prm->format == FMT_YUV422)
        mw >>= 1;
    for (uint32_t i = 0; i < prm->decomp_h; i++) {
        if (mw > 1)
            mw -= mw / 2;
        else {
            mw = 0;
            break;
        }
    }
    for (uint32_t i = 0; i < mv; i++) {
        if (mh > 1)
            mh -= mh / 2;
        else {
            mh = 0;
            break;
        }
    }
    if (mw == 0 || mh == 0)
        return E_BAD_PARAM;
    if (pi_build(pi, nc, sx, sy, prm->width, prm->height, prm->decomp_h, prm->decomp_v, slice_h))
        return E_BAD_PARAM;
    pi->depth = prm->depth;
    pi->format = prm->format;
    pi->lcod = (uint32_t)bpf;
    pi->qpih = prm->quantization;
    pi->sig = !!prm->significance;
    pi->rc_mode = prm->rc_mode;
    pi->vpred = prm->vpred;
    pi->signs = prm->signs;
    if (weights_assign(pi, prm->format))
        return E_BAD_PARAM;
    max_qr(pi);
    pi->hdr_len = write_frame_header(pi, pi->hdr);
    uint32_t tags = pi->hdr_len + 2 + 6 * pi->slices;
    if (tags >= pi->lcod)
        return E_BAD_PARAM;
    pi->slice_sizes = (uint32_t*)malloc(sizeof(uint32_t) * pi->slices);
    {
        uint32_t all = pi->lcod - tags;
        uint32_t last_prec = pi->pps - (pi->slices * pi->pps - pi->rows);
        uint32_t per_prec = all / pi->rows;
        uint32_t per_slice = per_prec * pi->pps;
        uint32_t last_slice = per_prec * last_prec;
        uint32_t left = all - last_slice - per_slice * (pi->slices - 1);
        if (pi->slices > 1)
            per_slice += (left * (pi->slices - 1) * pi->pps / pi->rows) / (pi->slices - 1);
        last_slice += left * last_prec / pi->rows;
        left = all - last_slice - per_slice * (pi->slices - 1);
        uint32_t remain = pi->lcod - pi->hdr_len;
        for (uint32_t i = 0; i < pi->slices; i++) {
            if (i != pi->slices - 1) {
                uint32_t s = per_slice + (i < left ? 1 : 0) + 6;
                pi->slice_sizes[i] = s;
                remain -= s;
            } else {
                pi->slice_sizes[i] = remain;
            }
        }
    }
    return E_NONE;
}

uint32_t jxso_codestream_bytes(const jxso_params_t* p) {
    ob_pi pi;
    if (enc_configure(p, &pi))
        return 0;
    free(pi.slice_sizes);
    return pi.lcod;
}

/* ----------------------------------------------------------------------
 * 5/3 lifting with symmetric extension.  Forward: Dwt.c:15-40 /
 * CodeDeprecated.cc:84-103.  Inverse: CodeDeprecated.cc:339-375 / Idwt.c.
 * Strided so the same code runs along rows and columns.
 * ---------------------------------------------------------------------- */
static void fwd53(const int32_t* x, size_t sx, int32_t* lf, size_t slf, int32_t* hf, size_t shf, uint32_t n) {
    uint32_t nh = n / 2, nl = n - nh;
    for (uint32_t i = 0; i < nh; i++) {
        if (2 * i + 2 < n)
            hf[i * shf] = x[(2 * i + 1) * sx] - ((x[2 * i * sx] + x[(2 * i + 2) * sx]) >> 1);
        else
            hf[i * shf] = x[(2 * i + 1) * sx] - x[2 * i * sx];
    }
    lf[0] = x[0] + ((hf[0] + 1) >> 1);
    for (uint32_t i = 1; i < nl; i++) {
        if (i < nh)
            lf[i * slf] = x[2 * i * sx] + ((hf[(i - 1) * shf] + hf[i * shf] + 2) >> 2);
        else
            lf[i * slf] = x[2 * i * sx] + ((hf[(i - 1) * shf] + 1) >> 1);
    }
}

static void inv53(const int32_t* lf, size_t slf, const int32_t* hf, size_t shf, int32_t* x, size_t sx, uint32_t n) {
    uint32_t nh = n / 2, nl = n - nh;
    x[0] = lf[0] - ((hf[0] + 1) >> 1);
    for (uint32_t i = 1; i < nl; i++) {
        if (i < nh)
            x[2 * i * sx] = lf[i * slf] - ((hf[(i - 1) * shf] + hf[i * shf] + 2) >> 2);
        else
            x[2 * i * sx] = lf[i * slf] - ((hf[(i - 1) * shf] + 1) >> 1);
    }
    for (uint32_t i = 0; i < nh; i++) {
        if (2 * i + 2 < n)
            x[(2 * i + 1) * sx] = hf[i * shf] + ((x[2 * i * sx] + x[(2 * i + 2) * sx]) >> 1);
        else
            x[(2 * i + 1) * sx] = hf[i * shf] + x[2 * i * sx];
    }
}

/* Band planes of one component: int32 values, band-major, each band w x h. */
typedef struct {
    int32_t* band[NBC];
} comp_bands32;

/* Forward transform of one component, order as
 * TestDwtFrame.cc:938-1040: for each vertical level, vertical lifting on the
 * current low band, then horizontal lifting of its low rows (-> next low band
 * and HL) and high rows (-> LH and HH); then the horizontal-only levels. */
static void dwt_component(const ob_comp* cp, int32_t* plane, comp_bands32* out) {
    uint32_t w = cp->w, h = cp->h;
    int32_t* cur = plane; /* w x h, stride w */
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
    int32_t* low = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
    uint32_t lvl = cp->nb - 1;
    for (uint32_t d = 0; d < cp->dv; d++) {
        uint32_t h2 = h / 2, h1 = h - h2, w2 = w / 2, w1 = w - w2;
        /* vertical: L rows into tmp[0..h1), H rows into tmp[h1..h) */
        for (uint32_t x = 0; x < w; x++)
            fwd53(cur + x, w, tmp + x, w, tmp + (size_t)h1 * w + x, w, h);
        int32_t* hl = out->band[lvl - 2];
        int32_t* lh = out->band[lvl - 1];
        int32_t* hh = out->band[lvl];
        for (uint32_t y = 0; y < h1; y++)
            fwd53(tmp + (size_t)y * w, 1, low + (size_t)y * w1, 1, hl + (size_t)y * w2, 1, w);
        for (uint32_t y = 0; y < h2; y++)
            fwd53(tmp + (size_t)(h1 + y) * w, 1, lh + (size_t)y * w1, 1, hh + (size_t)y * w2, 1, w);
        memcpy(plane, low, sizeof(int32_t) * (size_t)w1 * h1);
        cur = plane;
        w = w1;
        h = h1;
        lvl -= 3;
    }
    for (uint32_t d = cp->dv; d < cp->dh; d++) {
        uint32_t w2 = w / 2, w1 = w - w2;
        int32_t* hb = out->band[lvl];
        for (uint32_t y = 0; y < h; y++)
            fwd53(cur + (size_t)y * w, 1, low + (size_t)y * w1, 1, hb + (size_t)y * w2, 1, w);
        memcpy(plane, low, sizeof(int32_t) * (size_t)w1 * h);
        w = w1;
        lvl -= 1;
    }
    memcpy(out->band[0], plane, sizeof(int32_t) * (size_t)w * h);
    free(tmp);
    free(low);
}

/* Inverse of dwt_component (TestInvDwtFrame.cc:652-707). */
static void idwt_component(const ob_comp* cp, comp_bands32* in, int32_t* plane) {
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
    int32_t* tmp2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
    uint32_t w = cp->b[0].w, h = cp->b[0].h;
    memcpy(plane, in->band[0], sizeof(int32_t) * (size_t)w * h);
    uint32_t bi = 1;
    for (uint32_t d = 0; d < cp->dh - cp->dv; d++) {
        uint32_t w2 = cp->b[bi].w, n = w + w2;
        for (uint32_t y = 0; y < h; y++)
            inv53(plane + (size_t)y * w, 1, in->band[bi] + (size_t)y * w2, 1, tmp + (size_t)y * n, 1, n);
        memcpy(plane, tmp, sizeof(int32_t) * (size_t)n * h);
        w = n;
        bi++;
    }
    for (uint32_t d = 0; d < cp->dv; d++) {
        uint32_t w2 = cp->b[bi].w, h2 = cp->b[bi + 1].h, n = w + w2;
        /* rows: low rows from (plane, HL), high rows from (LH, HH) */
        for (uint32_t y = 0; y < h; y++)
            inv53(plane + (size_t)y * w, 1, in->band[bi] + (size_t)y * w2, 1, tmp + (size_t)y * n, 1, n);
        for (uint32_t y = 0; y < h2; y++)
            inv53(in->band[bi + 1] + (size_t)y * w, 1, in->band[bi + 2] + (size_t)y * w2, 1, tmp + (size_t)(h + y) * n, 1, n);
        for (uint32_t x = 0; x < n; x++)
            inv53(tmp + x, n, tmp + (size_t)h * n + x, n, tmp2 + x, n, h + h2);
        memcpy(plane, tmp2, sizeof(int32_t) * (size_t)n * (h + h2));
        w = n;
        h = h + h2;
        bi += 3;
    }
    free(tmp);
    free(tmp2);
}

/* ----------------------------------------------------------------------
 * Encoder frame state
 * ---------------------------------------------------------------------- */
typedef struct {
    uint16_t* coeff[4][NBC]; /* sign-magnitude, w x h per band */
    uint8_t* gcli[4][NBC];   /* gcli_w x h */
    uint8_t* smax[4][NBC];   /* sig_w x h */
} enc_frame;

/* Input scaling: NltEnc.c:33-71 (Tnlt = 0).  Packed RGB is split per line
 * first (GcStageProcess.c:396-425). */
static void load_component(const ob_pi* pi, uint32_t c, const void* const planes[4], const uint32_t strides[4], int32_t* dst) {
    const ob_comp* cp = &pi->c[c];
    const uint32_t shift = BW - pi->depth;
    const int32_t off = 1 << (BW - 1);
    const int packed = pi->format == FMT_PACKED444;
    const uint32_t mask = (1u << pi->depth) - 1;
    for (uint32_t y = 0; y < cp->h; y++) {
        for (uint32_t x = 0; x < cp->w; x++) {
            uint32_t v;
            size_t idx = packed ? (size_t)y * strides[0] + 3 * (size_t)x + c : (size_t)y * strides[c] + x;
            const void* base = packed ? planes[0] : planes[c];
            if (pi->depth <= 8)
                v = ((const uint8_t*)base)[idx];
            else
                v = ((const uint16_t*)base)[idx] & mask;
            dst[(size_t)y * cp->w + x] = (int32_t)(v << shift) - off;
        }
    }
}

/* image_shift: NltEnc.c:12-31 (Fq = 8) */
static inline uint16_t to_sign_mag(int32_t v) {
    if (v >= 0)
        return (uint16_t)((v + (1 << (FQ - 1))) >> FQ);
    uint32_t m = (uint32_t)((-v + (1 << (FQ - 1))) >> FQ);
    return (uint16_t)(m ? (m | 0x8000u) : 0);
}

/* GCLI per group of 4 and max per significance group of 8:
 * GcStageProcess.c:22-97. */
static inline uint8_t bitlen16(uint32_t v) {
    uint8_t n = 0;
    while (v) {
        n++;
        v >>= 1;
    }
    return n;
}

static void gcli_line(const uint16_t* coef, uint32_t w, uint8_t* g) {
    for (uint32_t i = 0; i < gcli_w(w); i++) {
        uint32_t o = 0;
        for (uint32_t k = 4 * i; k < 4 * i + 4 && k < w; k++)
            o |= coef[k] & 0x7FFFu;
        g[i] = bitlen16(o);
    }
}

static void smax_line(const uint8_t* g, uint32_t gw, uint8_t* s) {
    for (uint32_t i = 0; i < (gw + 7) / 8; i++) {
        uint8_t m = 0;
        for (uint32_t k = 8 * i; k < 8 * i + 8 && k < gw; k++)
            if (g[k] > m)
                m = g[k];
        s[i] = m;
    }
}

static int32_t enc_transform(const ob_pi* pi, const void* const planes[4], const uint32_t strides[4], enc_frame* f) {
    for (uint32_t c = 0; c < pi->nc; c++) {
        const ob_comp* cp = &pi->c[c];
        int32_t* plane = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
        comp_bands32 b32;
        for (uint32_t b = 0; b < cp->nb; b++)
            b32.band[b] = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->b[b].w * cp->b[b].h);
        load_component(pi, c, planes, strides, plane);
        dwt_component(cp, plane, &b32);
        for (uint32_t b = 0; b < cp->nb; b++) {
            const ob_band* bd = &cp->b[b];
            size_t n = (size_t)bd->w * bd->h;
            f->coeff[c][b] = (uint16_t*)malloc(sizeof(uint16_t) * n);
            for (size_t i = 0; i < n; i++)
                f->coeff[c][b][i] = to_sign_mag(b32.band[b][i]);
            f->gcli[c][b] = (uint8_t*)malloc((size_t)gcli_w(bd->w) * bd->h);
            f->smax[c][b] = (uint8_t*)malloc((size_t)sig_w(bd->w) * bd->h);
            for (uint32_t y = 0; y < bd->h; y++) {
                gcli_line(f->coeff[c][b] + (size_t)y * bd->w, bd->w, f->gcli[c][b] + (size_t)y * gcli_w(bd->w));
                smax_line(f->gcli[c][b] + (size_t)y * gcli_w(bd->w), gcli_w(bd->w), f->smax[c][b] + (size_t)y * sig_w(bd->w));
            }
            free(b32.band[b]);
        }
        free(plane);
    }
    return 0;
}

static void enc_frame_free(const ob_pi* pi, enc_frame* f) {
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            free(f->coeff[c][b]);
            free(f->gcli[c][b]);
            free(f->smax[c][b]);
        }
}

/* ----------------------------------------------------------------------
 * Rate control.  Per band line: histograms of GCLI and of the maxima of the
 * full significance groups (RateControl.c:18-154).  Sizes of a precinct for
 * one (Q, R): RateControl.c:457-781 with vertical prediction and sign
 * coding off.
 * ---------------------------------------------------------------------- */
typedef struct {
    uint32_t cnt_le[16];   /* number of groups with gcli <= g */
    uint32_t data[16];     /* sum over gcli > g of (gcli - g + 1) */
    uint32_t sig_le[16];   /* number of full sig groups with max <= g */
    int32_t tail_max;      /* max of the partial sig group, -1 if none */
} line_lut;

typedef struct {
    uint32_t row, type;
    line_lut lut[4][NBC][2];
    /* vertical prediction: the precinct above in the slice, its truncations */
    int has_top;
    uint8_t top_gtli[4][NBC];
    /* result of the last evaluation */
    uint8_t gtli[4][NBC];
    uint8_t method[4][NBC]; /* CodingMethodBand: 0 zero prediction, 1 + significance, 2 vertical, 3 vertical + significance */
    uint32_t bits_data[4][NBC][2], bits_gcli[4][NBC][2], bits_sig[4][NBC][2], bits_sign[4][NBC][2];
    uint8_t raw[NPK];
    uint32_t pk_data[NPK], pk_gcli[NPK], pk_sig[NPK], pk_sign[NPK];
} prec_state;

static void build_luts(const ob_pi* pi, const enc_frame* f, prec_state* ps) {
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            const ob_band* bd = &pi->c[c].b[b];
            uint32_t gw = gcli_w(bd->w), sw = sig_w(bd->w);
            for (uint32_t l = 0; l < pi->lines[ps->type][c][b]; l++) {
                uint32_t y = ps->row * bd->hl + l;
                const uint8_t* g = f->gcli[c][b] + (size_t)y * gw;
                const uint8_t* s = f->smax[c][b] + (size_t)y * sw;
                line_lut* L = &ps->lut[c][b][l];
                uint32_t hist[16] = {0}, shist[16] = {0};
                for (uint32_t i = 0; i < gw; i++)
                    hist[g[i]]++;
                uint32_t full = gw / 8;
                for (uint32_t i = 0; i < full; i++)
                    shist[s[i]]++;
                L->tail_max = full < sw ? s[full] : -1;
                uint32_t acc = 0, sacc = 0;
                for (int t = 0; t < 16; t++) {
                    acc += hist[t];
                    sacc += shist[t];
                    L->cnt_le[t] = acc;
                    L->sig_le[t] = sacc;
                    uint32_t d = 0;
                    for (int i = t + 1; i < 16; i++)
                        d += (uint32_t)(i - t + 1) * hist[i];
                    L->data[t] = d;
                }
            }
        }
}

/* ----------------------------------------------------------------------
 * Quantisation (Quant.c:9-45) applied while packing: the value packed for a
 * coefficient of a group with gcli > gtli.
 * ---------------------------------------------------------------------- */
static inline uint16_t quantize(uint16_t v, uint8_t gcli, uint8_t gtli, uint32_t qpih) {
    if (gcli <= gtli)
        return 0;
    uint16_t sign = v & 0x8000u, m = v & 0x7FFFu, o;
    if (gtli == 0)
        return v; /* QuantStageProcess.c:20-22 skips bands with gtli 0 */
    if (qpih == 0) {
        o = (uint16_t)((m >> gtli) << gtli);
    } else {
        uint8_t sc = gcli - gtli + 1;
        uint16_t d = (uint16_t)((((uint32_t)m << sc) - m + (1u << gcli)) >> (gcli + 1));
        o = (uint16_t)(d << gtli);
    }
    return o ? (uint16_t)(o | sign) : 0;
}

/* Bits of a vertical-prediction code without its closing zero:
 * vlc_encode_get_bits (PackPrecinct.h:84-104). */
static inline uint32_t vlc_bits(int32_t x, int32_t r, int32_t t) {
    int32_t mx = r - t;
    if (mx < 0)
        mx = 0;
    if (x > mx)
        return (uint32_t)(x + mx);
    x *= 2;
    return (uint32_t)(x < 0 ? -x - 1 : x);
}

/* The line above band line l of a precinct and its truncation: the previous
 * line of the band, or the last line of the band in the precinct above
 * (RateControl.c:603-613; Packing.c:265-276). */
static const uint8_t* top_line(const ob_pi* pi, const enc_frame* f, const prec_state* ps, uint32_t c, uint32_t b, uint32_t l,
                               uint8_t* gtli_top) {
    const ob_band* bd = &pi->c[c].b[b];
    const uint32_t gw = gcli_w(bd->w);
    if (l > 0) {
        *gtli_top = ps->gtli[c][b];
        return f->gcli[c][b] + (size_t)(ps->row * bd->hl + l - 1) * gw;
    }
    *gtli_top = ps->top_gtli[c][b];
    return f->gcli[c][b] + (size_t)((ps->row - 1) * bd->hl + pi->lines[0][c][b] - 1) * gw;
}

/* Vertical prediction of one band line: each group's code bits (without the
 * closing zero) and, per significance group, whether the whole group is
 * predicted exactly (Rm = 0) or below the truncation (Rm = 1):
 * rate_control_calc_vpred_cost_nosigf_c / _sigf_c (RateControl.c:291-361). */
static void vpred_line(const uint8_t* g, const uint8_t* top, uint32_t gw, uint8_t gtli, uint8_t gtli_top, uint32_t rm,
                       uint8_t* bits, uint8_t* flag) {
    const uint8_t gmax = gtli > gtli_top ? gtli : gtli_top;
    for (uint32_t i = 0; i < gw; i++) {
        int32_t mt = top[i] > gmax ? top[i] : gmax;
        int32_t gc = g[i] > gtli ? g[i] : gtli;
        bits[i] = (uint8_t)vlc_bits(gc - mt, mt, gtli);
    }
    for (uint32_t s = 0; s < (gw + 7) / 8; s++) {
        uint8_t fl = 1;
        for (uint32_t i = 8 * s; i < 8 * s + 8 && i < gw; i++) {
            if (rm) {
                if (g[i] > gtli)
                    fl = 0;
            } else {
                int32_t mt = top[i] > gmax ? top[i] : gmax;
                int32_t gc = g[i] > gtli ? g[i] : gtli;
                if (mt != gc)
                    fl = 0;
            }
        }
        flag[s] = fl;
    }
}

/* Signs a band line sends with full sign coding: its coefficients that stay
 * non-zero after quantisation (rate_control_sign_coding_get_sum_nonzero_coeff_c,
 * RateControl.c:363-455). */
static uint32_t nonzero_coeffs(const ob_pi* pi, const enc_frame* f, const prec_state* ps, uint32_t c, uint32_t b, uint32_t l,
                               uint8_t gtli) {
    const ob_band* bd = &pi->c[c].b[b];
    const uint32_t y = ps->row * bd->hl + l;
    const uint8_t* g = f->gcli[c][b] + (size_t)y * gcli_w(bd->w);
    const uint16_t* v = f->coeff[c][b] + (size_t)y * bd->w;
    uint32_t n = 0;
    for (uint32_t k = 0; k < bd->w; k++)
        n += quantize(v[k], g[k / 4], gtli, pi->qpih) != 0;
    return n;
}

/* Returns the bytes of all packets (without precinct/packet headers) for
 * (q, r) and records the per band / per packet choices in ps:
 * precinct_get_budget_bytes (RateControl.c:704-781) over
 * rate_control_calculate_band_best_method (RateControl.c:457-689).  vp: the
 * vertical prediction mode in force (0 in the slice-wide searches), sg: the
 * sign coding (fast in place of full there). */
static uint32_t precinct_bytes(const ob_pi* pi, const enc_frame* f, prec_state* ps, uint8_t q, uint8_t r, uint32_t vp,
                               uint32_t sg) {
    uint8_t vbits[4096], vflag[512];
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            const ob_band* bd = &pi->c[c].b[b];
            uint8_t gtli = truncation(bd->gain, bd->prio, q, r);
            ps->gtli[c][b] = gtli;
        }
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            const ob_band* bd = &pi->c[c].b[b];
            const uint8_t gtli = ps->gtli[c][b];
            uint32_t nl = pi->lines[ps->type][c][b];
            if (nl == 0)
                continue;
            uint32_t gw = gcli_w(bd->w), sw = sig_w(bd->w);
            uint32_t nosig[2], best = 0;
            for (uint32_t l = 0; l < nl; l++) {
                const line_lut* L = &ps->lut[c][b][l];
                const uint32_t coded = gw - L->cnt_le[gtli]; /* groups above the truncation */
                ps->bits_data[c][b][l] = (sg ? L->data[gtli] - coded : L->data[gtli]) * 4;
                ps->bits_sign[c][b][l] = sg == 0 ? 0 : sg == 1 ? 4 * coded : nonzero_coeffs(pi, f, ps, c, b, l, gtli);
                nosig[l] = L->data[gtli] + L->cnt_le[gtli];
                ps->bits_gcli[c][b][l] = nosig[l];
                ps->bits_sig[c][b][l] = 0;
                best += nosig[l];
            }
            ps->method[c][b] = 0;
            if (pi->sig) {
                uint32_t budget = sw * nl, withsig[2];
                for (uint32_t l = 0; l < nl; l++) {
                    const line_lut* L = &ps->lut[c][b][l];
                    uint32_t zeroed = L->sig_le[gtli] * 8;
                    if (L->tail_max >= 0 && (uint32_t)L->tail_max <= gtli)
                        zeroed += gw - (gw / 8) * 8;
                    withsig[l] = nosig[l] - zeroed;
                    budget += withsig[l];
                }
                if (best > budget) {
                    best = budget;
                    ps->method[c][b] = 1;
                    for (uint32_t l = 0; l < nl; l++) {
                        ps->bits_sig[c][b][l] = sw;
                        ps->bits_gcli[c][b][l] = withsig[l];
                    }
                }
            }
            if (vp && ps->has_top) {
                uint32_t vnosig[2], vsig[2], bud_nosig = 0, bud_sig = 0;
                for (uint32_t l = 0; l < nl; l++) {
                    uint8_t gtop;
                    const uint8_t* top = top_line(pi, f, ps, c, b, l, &gtop);
                    const uint8_t* g = f->gcli[c][b] + (size_t)(ps->row * bd->hl + l) * gw;
                    vpred_line(g, top, gw, gtli, gtop, pi->vpred == 2, vbits, vflag);
                    uint32_t sum = gw, red = 0;
                    for (uint32_t i = 0; i < gw; i++)
                        sum += vbits[i];
                    for (uint32_t s = 0; s < sw; s++)
                        if (vflag[s])
                            for (uint32_t i = 8 * s; i < 8 * s + 8 && i < gw; i++)
                                red += vbits[i] + 1u;
                    vnosig[l] = sum;
                    vsig[l] = sum - red;
                    bud_nosig += sum;
                    bud_sig += sum - red + sw;
                }
                if (best > bud_nosig) {
                    best = bud_nosig;
                    ps->method[c][b] = 2;
                }
                if (pi->sig && best > bud_sig) {
                    best = bud_sig;
                    ps->method[c][b] = 3;
                }
                if (ps->method[c][b] >= 2)
                    for (uint32_t l = 0; l < nl; l++) {
                        ps->bits_sig[c][b][l] = ps->method[c][b] == 3 ? sw : 0;
                        ps->bits_gcli[c][b][l] = ps->method[c][b] == 3 ? vsig[l] : vnosig[l];
                    }
            }
        }
    uint32_t total = 0;
    for (uint32_t p = 0; p < pi->npk; p++) {
        uint32_t d = 0, g = 0, s = 0, sn = 0;
        for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
            uint32_t c = pi->gc[gi], b = pi->gb[gi], l = pi->pk_line[p];
            if (l < pi->lines[ps->type][c][b]) {
                d += ps->bits_data[c][b][l];
                g += ps->bits_gcli[c][b][l];
                s += ps->bits_sig[c][b][l];
                sn += ps->bits_sign[c][b][l];
            }
        }
        uint32_t db = bytes_of_bits(d), gb = bytes_of_bits(g), sb = bytes_of_bits(s), nb = bytes_of_bits(sn);
        ps->pk_data[p] = db;
        ps->pk_sign[p] = nb;
        total += db + nb;
        if (sb + gb > pi->raw_bytes[ps->type][p]) {
            ps->raw[p] = 1;
            ps->pk_sig[p] = 0;
            ps->pk_gcli[p] = pi->raw_bytes[ps->type][p];
            total += pi->raw_bytes[ps->type][p];
        } else {
            ps->raw[p] = 0;
            ps->pk_sig[p] = sb;
            ps->pk_gcli[p] = gb;
            total += sb + gb;
        }
    }
    return total;
}

/* BinarySearch.c:9-78 */
void jxso_bsearch_init(jxso_bsearch_t* b, uint32_t beg, uint32_t end, uint8_t below, uint32_t step) {
    b->beg = (int32_t)beg;
    b->end = (int32_t)end;
    b->below = below;
    b->best = -1;
    b->last = -1;
    b->step = step ? (int32_t)step : (b->end - b->beg) / 2;
}

int32_t jxso_bsearch_next(jxso_bsearch_t* b, int32_t result, uint32_t* next) {
    if (result == 1) {
        if (b->below && (b->best == -1 || b->best < b->last))
            b->best = b->last;
        if (b->end >= b->last + 1)
            b->beg = b->last + 1;
        else if (b->best == -1)
            return 2;
        else {
            *next = (uint32_t)b->best;
            return 1;
        }
    } else if (result == 2 || result == 3) {
        if (result == 2 && !b->below && (b->best == -1 || b->best > b->last))
            b->best = b->last;
        if (b->beg <= b->last - 1)
            b->end = b->last - 1;
        else if (b->best == -1)
            return 2;
        else {
            *next = (uint32_t)b->best;
            return 1;
        }
    }
    if (b->step > b->end - b->beg)
        b->step = (b->end - b->beg + 1) / 2;
    b->last = b->step + b->beg;
    *next = (uint32_t)b->last;
    return 0;
}

/* rate_control_precinct (RateControl.c:1259-1333) with the binary searches
 * of RateControl.c:889-1083.  On success fills q, r and the data bytes. */
static int32_t rc_precinct(const ob_pi* pi, const enc_frame* f, prec_state* ps, uint32_t budget, uint8_t* oq, uint8_t* or_,
                           uint32_t* odata, uint32_t* ohdr) {
    const uint32_t vp = pi->vpred, sg = pi->signs;
    if (budget > 1048575u)
        return E_ENCODE_FRAME;
    uint32_t hdr = bytes_of_bits(40 + pi->nb_exist * 2) + bytes_of_bits((pi->short_hdr ? 40u : 56u) * (uint32_t)pi->pk_exist[ps->type]);
    *ohdr = hdr;
    if (budget <= hdr)
        return E_UNDEFINED;
    uint32_t avail = budget - hdr;
    /* Q: smallest acceptable */
    jxso_bsearch_t s;
    uint32_t step = 6;
    if (step > pi->maxq)
        step = 0;
    jxso_bsearch_init(&s, 0, pi->maxq, 0, step);
    int32_t res, next = 0;
    uint32_t q;
    while ((res = jxso_bsearch_next(&s, next, &q)) == 0) {
        if (all_truncated(pi, (uint8_t)q, 0)) {
            next = 3;
            continue;
        }
        next = precinct_bytes(pi, f, ps, (uint8_t)q, 0, vp, sg) > avail ? 1 : 2;
    }
    if (res != 1)
        return E_ENCODE_FRAME;
    /* R: largest acceptable */
    step = 6;
    if (step >= pi->maxr)
        step = 0;
    jxso_bsearch_init(&s, 0, pi->maxr, 1, step);
    next = 0;
    uint32_t r;
    int32_t last_tested = -1;
    uint32_t last_bytes = 0;
    while ((res = jxso_bsearch_next(&s, next, &r)) == 0) {
        last_tested = -1;
        if (all_truncated(pi, (uint8_t)q, (uint8_t)r)) {
            next = 3;
            continue;
        }
        uint32_t by = precinct_bytes(pi, f, ps, (uint8_t)q, (uint8_t)r, vp, sg);
        last_tested = (int32_t)r;
        last_bytes = by;
        next = by <= avail ? 1 : 2;
    }
    if (res != 1)
        return E_ENCODE_FRAME;
    if ((int32_t)r != last_tested)
        last_bytes = precinct_bytes(pi, f, ps, (uint8_t)q, (uint8_t)r, vp, sg);
    *oq = (uint8_t)q;
    *or_ = (uint8_t)r;
    *odata = last_bytes;
    return E_NONE;
}

/* Overwrites n bits at bit position pos (MSB first). */
static void bw_set(bw_t* w, uint64_t pos, uint32_t v, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t p = pos + i;
        const uint8_t m = (uint8_t)(0x80u >> (p & 7));
        if ((v >> (n - 1 - i)) & 1)
            w->m[p >> 3] |= m;
        else
            w->m[p >> 3] &= (uint8_t)~m;
    }
}

/* pack_precinct (PackPrecinct.c:235-572).  total_bytes: the precinct's bytes
 * with headers; pad_bytes: zero bytes after it.  With fast sign coding the
 * sign sub-packets take only the signs sent, each packet header gets their
 * true length, and the bytes so saved either shorten the precinct (cut: Lprc
 * is lowered) or pad it; they are returned (PackPrecinct.c:466-571). */
static uint32_t pack_precinct(const ob_pi* pi, const enc_frame* f, prec_state* ps, uint8_t q, uint8_t r, uint32_t total_bytes,
                              uint32_t pad_bytes, int cut, bw_t* w) {
    uint8_t vbits[4096], vflag[512];
    uint32_t hlen = bytes_of_bits(40 + 2 * pi->nb_exist);
    const uint64_t lprc_pos = w->bit;
    bw_put(w, total_bytes - hlen, 24);
    bw_put(w, q, 8);
    bw_put(w, r, 8);
    for (uint32_t g = 0; g < pi->nb_all; g++) {
        if (pi->gb[g] == NE)
            continue;
        uint32_t c = pi->gc[g], b = pi->gb[g];
        static const uint8_t dpb[4] = {0, 2, 1, 3}; /* D[p,b]: bit 0 vertical prediction, bit 1 significance */
        uint32_t d = 0;
        if (pi->lines[ps->type][c][b] > 0)
            d = dpb[ps->method[c][b]];
        bw_put(w, d, 2);
    }
    bw_align(w);
    uint32_t retrieve = 0;
    for (uint32_t p = 0; p < pi->npk; p++) {
        uint32_t l = pi->pk_line[p];
        int any = 0;
        for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++)
            if (l < pi->lines[ps->type][pi->gc[gi]][pi->gb[gi]])
                any = 1;
        if (!any)
            continue;
        /* packet header: PackHeaders.c:98-127 */
        const uint32_t sign_field = pi->short_hdr ? 11 : 15;
        bw_put(w, ps->raw[p], 1);
        if (pi->short_hdr) {
            bw_put(w, ps->pk_data[p], 15);
            bw_put(w, ps->pk_gcli[p], 13);
        } else {
            bw_put(w, ps->pk_data[p], 20);
            bw_put(w, ps->pk_gcli[p], 20);
        }
        const uint64_t sign_pos = w->bit;
        bw_put(w, ps->pk_sign[p], sign_field);
        if (ps->raw[p]) {
            for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                uint32_t c = pi->gc[gi], b = pi->gb[gi];
                if (l >= pi->lines[ps->type][c][b])
                    continue;
                const ob_band* bd = &pi->c[c].b[b];
                const uint8_t* g = f->gcli[c][b] + (size_t)(ps->row * bd->hl + l) * gcli_w(bd->w);
                for (uint32_t i = 0; i < gcli_w(bd->w); i++)
                    bw_put(w, g[i], 4);
            }
            bw_align(w);
        } else {
            /* significance sub-packet */
            for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                uint32_t c = pi->gc[gi], b = pi->gb[gi];
                const uint8_t m = ps->method[c][b];
                if (l >= pi->lines[ps->type][c][b] || (m != 1 && m != 3))
                    continue;
                const ob_band* bd = &pi->c[c].b[b];
                const uint32_t y = ps->row * bd->hl + l, gw = gcli_w(bd->w);
                if (m == 1) {
                    const uint8_t* s = f->smax[c][b] + (size_t)y * sig_w(bd->w);
                    for (uint32_t i = 0; i < sig_w(bd->w); i++)
                        bw_put(w, s[i] <= ps->gtli[c][b], 1);
                } else {
                    uint8_t gtop;
                    const uint8_t* top = top_line(pi, f, ps, c, b, l, &gtop);
                    vpred_line(f->gcli[c][b] + (size_t)y * gw, top, gw, ps->gtli[c][b], gtop, pi->vpred == 2, vbits,
                               vflag);
                    for (uint32_t i = 0; i < sig_w(bd->w); i++)
                        bw_put(w, vflag[i], 1);
                }
            }
            bw_align(w);
            /* bit-plane count sub-packet */
            for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                uint32_t c = pi->gc[gi], b = pi->gb[gi];
                if (l >= pi->lines[ps->type][c][b])
                    continue;
                const ob_band* bd = &pi->c[c].b[b];
                uint32_t gw = gcli_w(bd->w);
                const uint32_t y = ps->row * bd->hl + l;
                const uint8_t* g = f->gcli[c][b] + (size_t)y * gw;
                const uint8_t* s = f->smax[c][b] + (size_t)y * sig_w(bd->w);
                const uint8_t gt = ps->gtli[c][b], m = ps->method[c][b];
                if (m >= 2) {
                    uint8_t gtop;
                    const uint8_t* top = top_line(pi, f, ps, c, b, l, &gtop);
                    vpred_line(g, top, gw, gt, gtop, pi->vpred == 2, vbits, vflag);
                }
                for (uint32_t i = 0; i < gw; i++) {
                    uint32_t n;
                    if (m == 1 && s[i / 8] <= gt)
                        continue;
                    if (m == 3 && vflag[i / 8])
                        continue;
                    if (m >= 2)
                        n = vbits[i];
                    else
                        n = g[i] > gt ? g[i] - gt : 0;
                    bw_put(w, ((1u << n) - 1) << 1, n + 1); /* n ones, then a zero */
                }
            }
            bw_align(w);
        }
        /* data sub-packet: per group the four signs (unless sent apart),
         * then bit-planes gcli - 1 .. gtli (PackPrecinct.c:125-198) */
        for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
            uint32_t c = pi->gc[gi], b = pi->gb[gi];
            if (l >= pi->lines[ps->type][c][b])
                continue;
            const ob_band* bd = &pi->c[c].b[b];
            uint32_t y = ps->row * bd->hl + l;
            const uint8_t* g = f->gcli[c][b] + (size_t)y * gcli_w(bd->w);
            const uint16_t* v = f->coeff[c][b] + (size_t)y * bd->w;
            uint8_t gt = ps->gtli[c][b];
            for (uint32_t i = 0; i < gcli_w(bd->w); i++) {
                if (g[i] <= gt)
                    continue;
                uint16_t qv[4];
                for (int k = 0; k < 4; k++)
                    qv[k] = (4 * i + k < bd->w) ? quantize(v[4 * i + k], g[i], gt, pi->qpih) : 0;
                if (!pi->signs)
                    for (int k = 0; k < 4; k++)
                        bw_put(w, qv[k] >> 15, 1);
                for (int bit = g[i] - 1; bit >= gt; bit--)
                    for (int k = 0; k < 4; k++)
                        bw_put(w, (qv[k] >> bit) & 1, 1);
            }
        }
        bw_align(w);
        if (pi->signs) {
            /* sign sub-packet: one bit per coefficient that stays non-zero
             * (pack_sign_c, PackPrecinct.c:200-233) */
            const uint64_t s0 = w->bit;
            for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                uint32_t c = pi->gc[gi], b = pi->gb[gi];
                if (l >= pi->lines[ps->type][c][b])
                    continue;
                const ob_band* bd = &pi->c[c].b[b];
                uint32_t y = ps->row * bd->hl + l;
                const uint8_t* g = f->gcli[c][b] + (size_t)y * gcli_w(bd->w);
                const uint16_t* v = f->coeff[c][b] + (size_t)y * bd->w;
                for (uint32_t k = 0; k < bd->w; k++) {
                    uint16_t qv = quantize(v[k], g[k / 4], ps->gtli[c][b], pi->qpih);
                    if (qv)
                        bw_put(w, qv >> 15, 1);
                }
            }
            bw_align(w);
            const uint32_t sent = (uint32_t)((w->bit - s0) >> 3);
            if (pi->signs == 1) {
                bw_set(w, sign_pos, sent, sign_field);
                retrieve += ps->pk_sign[p] - sent;
            }
        }
    }
    if (pi->signs == 1) {
        if (cut)
            bw_set(w, lprc_pos, total_bytes - hlen - retrieve, 24);
        else
            w->bit += 8ull * retrieve;
    }
    w->bit += 8ull * pad_bytes; /* zero padding */
    return pi->signs == 1 ? retrieve : 0;
}

/* Precinct budgets of one slice for rate control modes 2 and 3
 * (PackStageProcess.c:340-481): the data budget (the slice's bytes less
 * every precinct's headers, RateControl.c:258-271) decides one Q and then one
 * R for all the slice's precincts with the per-precinct searches' binary
 * search on the summed bytes (RateControl.c:1086-1210, 1336-1425); each
 * precinct gets its headers and data at (Q, R), the last one the leftover.
 * Mode 3 raises every precinct to a quarter of the largest and takes the
 * bytes back from the leftover, else one byte at a time from the first
 * precincts on.  The leftover then goes 10 % of the first precinct's bytes
 * to the first one and the rest evenly, the first precincts one byte more. */
static int32_t slice_budgets(const ob_pi* pi, const enc_frame* f, uint32_t first_row, uint32_t nprec, uint32_t sbudget,
                             uint32_t* bud) {
    if ((sbudget + nprec - 1) / nprec > 1048575u)
        return E_ENCODE_FRAME;
    prec_state* st = (prec_state*)calloc(nprec, sizeof(prec_state));
    uint32_t* pad = (uint32_t*)calloc(nprec, sizeof(uint32_t));
    /* the slice-wide searches run without vertical prediction and with fast
     * sign coding in place of full (RateControl.c:1103-1106, 1120-1121) */
    const uint32_t sg = pi->signs ? 1u : 0u;
    int32_t err = E_NONE;
    uint32_t hsum = 0;
    for (uint32_t i = 0; i < nprec; i++) {
        st[i].row = first_row + i;
        st[i].type = (st[i].row + 1 >= pi->rows) ? 1 : 0;
        build_luts(pi, f, &st[i]);
        bud[i] = bytes_of_bits(40 + pi->nb_exist * 2) +
            bytes_of_bits((pi->short_hdr ? 40u : 56u) * (uint32_t)pi->pk_exist[st[i].type]);
        hsum += bud[i];
    }
    uint32_t q = 0, r = 0;
    if (sbudget <= hsum) {
        err = E_ENCODE_FRAME;
    } else {
        const uint64_t avail = sbudget - hsum;
        jxso_bsearch_t s;
        uint32_t step = 6;
        if (step > pi->maxq)
            step = 0;
        jxso_bsearch_init(&s, 0, pi->maxq, 0, step);
        int32_t res, next = 0;
        while ((res = jxso_bsearch_next(&s, next, &q)) == 0) {
            if (all_truncated(pi, (uint8_t)q, 0)) {
                next = 3;
                continue;
            }
            uint64_t sum = 0;
            for (uint32_t i = 0; i < nprec; i++)
                sum += precinct_bytes(pi, f, &st[i], (uint8_t)q, 0, 0, sg);
            next = sum > avail ? 1 : 2;
        }
        if (res != 1)
            err = E_ENCODE_FRAME;
        if (!err) {
            step = 6;
            if (step >= pi->maxr)
                step = 0;
            jxso_bsearch_init(&s, 0, pi->maxr, 1, step);
            next = 0;
            while ((res = jxso_bsearch_next(&s, next, &r)) == 0) {
                if (all_truncated(pi, (uint8_t)q, (uint8_t)r)) {
                    next = 3;
                    continue;
                }
                uint64_t sum = 0;
                for (uint32_t i = 0; i < nprec; i++)
                    sum += precinct_bytes(pi, f, &st[i], (uint8_t)q, (uint8_t)r, 0, sg);
                next = sum <= avail ? 1 : 2;
            }
            if (res != 1)
                err = E_ENCODE_FRAME;
        }
        if (!err) {
            uint32_t left = (uint32_t)avail;
            for (uint32_t i = 0; i < nprec; i++) {
                const uint32_t d = precinct_bytes(pi, f, &st[i], (uint8_t)q, (uint8_t)r, 0, sg);
                left -= d;
                bud[i] += d;
            }
            const uint32_t last = nprec - 1;
            pad[last] = left;
            bud[last] += left;
            if (pi->rc_mode == 3) {
                uint32_t mx = 0, get = 0;
                for (uint32_t i = 0; i < nprec; i++)
                    if (bud[i] - pad[i] > mx)
                        mx = bud[i] - pad[i];
                const uint32_t mn = mx / 4;
                for (uint32_t i = 0; i < nprec; i++) {
                    const uint32_t ds = bud[i] - pad[i];
                    if (ds < mn) {
                        bud[i] += mn - ds;
                        get += mn - ds;
                    }
                }
                if (get) {
                    if (pad[last] >= get) {
                        pad[last] -= get;
                        bud[last] -= get;
                    } else {
                        get -= pad[last];
                        bud[last] -= pad[last];
                        pad[last] = 0;
                        while (get)
                            for (uint32_t i = 0; i < nprec && get; i++) {
                                get--;
                                bud[i]--;
                            }
                    }
                }
            }
            if (pad[last] > 0) {
                /* 20 % with fast sign coding (PackStageProcess.c:440-450) */
                uint32_t add_first = bud[0] * (pi->signs == 1 ? 20u : 10u) / 100;
                if (add_first > pad[last])
                    add_first = pad[last];
                pad[last] -= add_first;
                bud[last] -= add_first;
                bud[0] += add_first;
                if (pad[last]) {
                    const uint32_t lb = pad[last];
                    bud[last] -= lb;
                    pad[last] = 0;
                    const uint32_t mb = lb / nprec, rem = lb - mb * nprec;
                    for (uint32_t i = 0; i < nprec; i++)
                        bud[i] += mb + (i < rem ? 1u : 0u);
                }
            }
        }
    }
    free(pad);
    free(st);
    return err;
}

/* One whole frame: PreRcStageProcess.c:31-116 (slice byte ranges),
 * PackStageProcess.c:567-749 (per-slice precinct budgets; with rate control
 * mode 1 the first precinct gets 10 % more and each precinct's unused bytes
 * move to the next one of its slice, PackStageProcess.c:621-640, 293-307;
 * mode 0 splits the slice evenly and pads every precinct to its budget),
 * PackStageProcess.c:274-338 (per precinct). */
static int32_t enc_code(const ob_pi* pi, const enc_frame* f, uint8_t* out, uint8_t* qr_dump) {
    memset(out, 0, pi->lcod);
    memcpy(out, pi->hdr, pi->hdr_len);
    uint64_t pos = pi->hdr_len;
    prec_state* ps = (prec_state*)calloc(1, sizeof(prec_state));
    int32_t err = E_NONE;
    for (uint32_t s = 0; s < pi->slices && !err; s++) {
        uint32_t first_row = s * pi->pps;
        uint32_t nprec = pi->pps;
        uint32_t sbudget;
        if (s == pi->slices - 1) {
            nprec = pi->rows - first_row;
            sbudget = pi->slice_sizes[s] - 6 - 2;
        } else {
            sbudget = pi->slice_sizes[s] - 6;
        }
        bw_t w = {out, pos * 8};
        bw_put(&w, 0xFF20, 16);
        bw_put(&w, 4, 16);
        bw_put(&w, s, 16);
        uint32_t minb = sbudget / nprec;
        uint32_t first = minb, left = sbudget - minb * nprec;
        if (pi->rc_mode == 1) {
            /* 10 % more for the first precinct, 20 % with fast sign coding
             * (PackStageProcess.c:621-640) */
            first = minb * (pi->signs == 1 ? 120u : 110u) / 100;
            if (first > sbudget)
                first = sbudget;
            if (nprec > 1) {
                minb = (sbudget - first) / (nprec - 1);
                left = sbudget - minb * (nprec - 1) - first;
            } else {
                left = 0;
            }
        }
        uint32_t* sbud = NULL;
        if (pi->rc_mode >= 2) {
            sbud = (uint32_t*)calloc(nprec, sizeof(uint32_t));
            if (slice_budgets(pi, f, first_row, nprec, sbudget, sbud)) {
                free(sbud);
                err = E_ENCODE_FRAME;
                break;
            }
        }
        uint32_t carry = 0;
        for (uint32_t i = 0; i < nprec; i++) {
            uint32_t budget = (sbud ? sbud[i] : (i == 0 ? first : minb) + (i < left ? 1 : 0)) + carry;
            /* the precinct above in the slice, for vertical prediction
             * (PackStageProcess.c:655-660, 355-358) */
            if (i > 0)
                memcpy(ps->top_gtli, ps->gtli, sizeof(ps->gtli));
            ps->has_top = pi->vpred && i > 0;
            ps->row = first_row + i;
            ps->type = (ps->row + 1 >= pi->rows) ? 1 : 0;
            build_luts(pi, f, ps);
            uint8_t q = 0, r = 0;
            uint32_t data = 0, hdr = 0;
            err = rc_precinct(pi, f, ps, budget, &q, &r, &data, &hdr);
            if (err) {
                /* any slice error surfaces as a frame error (FinalStageProcess.c:120,
                 * EncHandle.c:1043-1046) */
                err = E_ENCODE_FRAME;
                break;
            }
            uint32_t pad = budget - hdr - data;
            uint32_t total = budget;
            int cut = 0;
            if (pi->rc_mode != 0 && i + 1 < nprec) {
                total -= pad;
                carry = pad;
                pad = 0;
                cut = pi->signs == 1; /* PackStageProcess.c:305-307, 533-535 */
            }
            if (qr_dump) {
                qr_dump[2 * ps->row] = q;
                qr_dump[2 * ps->row + 1] = r;
            }
            const uint32_t retrieved = pack_precinct(pi, f, ps, q, r, total, pad, cut, &w);
            if (cut)
                carry += retrieved; /* PackStageProcess.c:326-333, 556-560 */
        }
        free(sbud);
        pos = w.bit >> 3;
        if (s == pi->slices - 1 && !err) {
            out[pos] = 0xFF;
            out[pos + 1] = 0x11;
            pos += 2;
        }
    }
    free(ps);
    if (!err && pos != pi->lcod)
        err = E_ENCODE_FRAME;
    return err;
}

int32_t jxso_encode(const jxso_params_t* p, const void* const planes[4], const uint32_t strides[4], uint8_t* out,
                    uint32_t out_cap, uint32_t* out_len) {
    ob_pi pi;
    int32_t e = enc_configure(p, &pi);
    if (e)
        return e;
    if (out_cap < pi.lcod) {
        free(pi.slice_sizes);
        return E_BAD_PARAM;
    }
    enc_frame f;
    memset(&f, 0, sizeof(f));
    enc_transform(&pi, planes, strides, &f);
    e = enc_code(&pi, &f, out, NULL);
    if (out_len)
        *out_len = pi.lcod;
    enc_frame_free(&pi, &f);
    free(pi.slice_sizes);
    return e;
}

int32_t jxso_encode_dump(const jxso_params_t* p, const void* const planes[4], const uint32_t strides[4], uint16_t* coeff,
                         uint8_t* gcli, uint8_t* qr, uint32_t sizes[3]) {
    ob_pi pi;
    int32_t e = enc_configure(p, &pi);
    if (e)
        return e;
    enc_frame f;
    memset(&f, 0, sizeof(f));
    enc_transform(&pi, planes, strides, &f);
    size_t nc = 0, ng = 0;
    for (uint32_t c = 0; c < pi.nc; c++)
        for (uint32_t b = 0; b < pi.c[c].nb; b++) {
            const ob_band* bd = &pi.c[c].b[b];
            size_t n = (size_t)bd->w * bd->h, m = (size_t)gcli_w(bd->w) * bd->h;
            if (coeff)
                memcpy(coeff + nc, f.coeff[c][b], n * 2);
            if (gcli)
                memcpy(gcli + ng, f.gcli[c][b], m);
            nc += n;
            ng += m;
        }
    uint8_t* tmp = (uint8_t*)malloc(pi.lcod);
    e = enc_code(&pi, &f, tmp, qr);
    free(tmp);
    sizes[0] = (uint32_t)nc;
    sizes[1] = (uint32_t)ng;
    sizes[2] = 2 * pi.rows;
    enc_frame_free(&pi, &f);
    free(pi.slice_sizes);
    return e;
}

/* ----------------------------------------------------------------------
 * Decoder: header parsing (ParseHeader.c:20-618, the markers this encoder
 * writes plus COM skipping), precinct unpacking (Packing.c:513-868, with
 * vertical prediction and sign sub-packets), dequantisation and sign conversion
 * (Dequant.c:9-57, Precinct.c:9-29), inverse transform, linear output
 * scaling (NltDec.c:242-265).
 * ---------------------------------------------------------------------- */
typedef struct {
    ob_pi pi;
    uint32_t hdr_end;
    uint32_t depth, bw, fq, qpih, lcod;
    uint32_t fs, rm; /* separate sign sub-packets; run mode of vertically predicted significance groups */
} dec_hdr;

static int32_t parse_header(const uint8_t* cs, uint32_t len, dec_hdr* d) {
    memset(d, 0, sizeof(*d));
    br_t r = {cs, 0, (uint64_t)len * 8, 0};
    if (br_get(&r, 16) != 0xFF10)
        return E_INVALID_BS;
    int have_pih = 0, have_cdt = 0, have_wgt = 0;
    uint32_t W = 0, H = 0, nc = 0, dh = 0, dv = 0, hsl = 0, cw = 0, lh = 0;
    uint32_t sx[4] = {0}, sy[4] = {0};
    uint8_t gains[NGB], prios[NGB];
    uint32_t nw = 0;
    for (;;) {
        if (r.bit + 16 > r.end_bit)
            return E_TOO_SHORT;
        uint32_t marker = br_get(&r, 16);
        if (marker == 0xFF20) {
            r.bit -= 16;
            break;
        }
        uint32_t seglen = br_get(&r, 16);
        uint64_t seg_start = r.bit - 16;
        if (r.err)
            return E_TOO_SHORT;
        switch (marker) {
        case 0xFF12: /* PIH */
            d->lcod = br_get(&r, 32);
            br_get(&r, 16);
            br_get(&r, 16);
            W = br_get(&r, 16);
            H = br_get(&r, 16);
            cw = br_get(&r, 16);
            hsl = br_get(&r, 16);
            nc = br_get(&r, 8);
            if (br_get(&r, 8) != 4 || br_get(&r, 8) != 8)
                return E_INVALID_BS;
            d->bw = br_get(&r, 8);
            d->fq = br_get(&r, 4);
            if (br_get(&r, 4) != 4)
                return E_INVALID_BS;
            if (br_get(&r, 8) != 0) /* Fslc, Ppoc, Cpih */
                return E_INVALID_BS;
            dh = br_get(&r, 4);
            dv = br_get(&r, 4);
            lh = br_get(&r, 1);
            br_get(&r, 1); /* Rl */
            d->qpih = br_get(&r, 2);
            d->fs = br_get(&r, 2);
            d->rm = br_get(&r, 2);
            /* ParseHeader.c:238-251 */
            if (d->qpih > 1 || d->fs > 1 || d->rm > 1)
                return E_INVALID_BS;
            have_pih = 1;
            break;
        case 0xFF13: /* CDT */
            for (uint32_t c = 0; c < (seglen - 2) / 2 && c < 4; c++) {
                d->depth = br_get(&r, 8);
                sx[c] = br_get(&r, 4);
                sy[c] = br_get(&r, 4);
            }
            have_cdt = 1;
            break;
        case 0xFF14: /* WGT */
            nw = (seglen - 2) / 2;
            if (nw > NGB)
                return E_INVALID_BS;
            for (uint32_t i = 0; i < nw; i++) {
                gains[i] = (uint8_t)br_get(&r, 8);
                prios[i] = (uint8_t)br_get(&r, 8);
            }
            have_wgt = 1;
            break;
        case 0xFF50: /* CAP */
        case 0xFF15: /* COM */
            break;
        default:
            return E_INVALID_BS; /* NLT, CWD, CTS, CRG: tools this path does not decode */
        }
        r.bit = seg_start + 8ull * seglen;
        if (r.err || r.bit > r.end_bit)
            return E_TOO_SHORT;
    }
    if (!have_pih || !have_cdt || !have_wgt || cw != 0 || nc == 0 || nc > 4 || hsl == 0)
        return E_INVALID_BS;
    /* sampling factors this restatement lays out (a damaged header can carry
     * any 4-bit value; Pi.c:230-231 refuses 0) */
    for (uint32_t c = 0; c < nc; c++)
        if (sx[c] < 1 || sx[c] > 2 || sy[c] < 1 || sy[c] > 2)
            return E_INVALID_BS;
    if (pi_build(&d->pi, nc, sx, sy, W, H, dh, dv, hsl << dv))
        return E_INVALID_BS;
    if (d->pi.short_hdr != !lh)
        return E_INVALID_BS;
    if (nw != d->pi.nb_exist)
        return E_INVALID_BS;
    uint32_t k = 0;
    for (uint32_t g = 0; g < d->pi.nb_all; g++) {
        if (d->pi.gc[g] == NE)
            continue;
        d->pi.c[d->pi.gc[g]].b[d->pi.gb[g]].gain = gains[k];
        d->pi.c[d->pi.gc[g]].b[d->pi.gb[g]].prio = prios[k];
        k++;
    }
    d->pi.depth = d->depth;
    d->hdr_end = (uint32_t)(r.bit >> 3);
    return E_NONE;
}

/* Proxy decoding, pi_update_proxy_mode (Pi.c:567-626): half (1) or quarter
 * (2) resolution drops that many vertical and horizontal levels from the
 * picture and from every component; only the first packets of each
 * precinct, which carry the remaining bands, are decoded (one packet, or
 * four at half resolution with two vertical levels).  The remaining bands
 * are the first 2V' + H' + 1 of each component, unchanged. */
static int32_t proxy_geometry(const ob_pi* pi, uint32_t proxy, ob_comp out[4], uint32_t* npk_dec, uint32_t* w,
                              uint32_t* h) {
    *npk_dec = pi->npk;
    *w = pi->w;
    *h = pi->h;
    for (uint32_t c = 0; c < pi->nc; c++)
        out[c] = pi->c[c];
    if (proxy == 0)
        return E_NONE;
    if (proxy > 2)
        return E_BAD_PARAM;
    *npk_dec = proxy == 1 && pi->dv != 1 ? 4u : 1u;
    if (proxy > pi->dv || proxy > pi->dh)
        return E_BAD_PARAM;
    *w = (pi->w + (1u << proxy) - 1) >> proxy;
    *h = (pi->h + (1u << proxy) - 1) >> proxy;
    for (uint32_t c = 0; c < pi->nc; c++) {
        if (proxy > out[c].dv || proxy > out[c].dh)
            return E_BAD_PARAM;
        out[c].w = (out[c].w + (1u << proxy) - 1) >> proxy;
        out[c].h = (out[c].h + (1u << proxy) - 1) >> proxy;
        out[c].dv -= proxy;
        out[c].dh -= proxy;
        out[c].nb = 2 * out[c].dv + out[c].dh + 1;
    }
    return E_NONE;
}

static void fill_info(const dec_hdr* d, const ob_comp comps[4], uint32_t w, uint32_t h, jxso_image_info_t* info) {
    info->width = w;
    info->height = h;
    info->depth = d->depth;
    info->comps = d->pi.nc;
    for (uint32_t c = 0; c < d->pi.nc; c++) {
        info->comp_w[c] = comps[c].w;
        info->comp_h[c] = comps[c].h;
    }
}

int32_t jxso_probe_proxy(const uint8_t* cs, uint32_t len, uint32_t proxy, jxso_image_info_t* info) {
    dec_hdr d;
    int32_t e = parse_header(cs, len, &d);
    if (e)
        return e;
    ob_comp comps[4];
    uint32_t npk_dec, w, h;
    e = proxy_geometry(&d.pi, proxy, comps, &npk_dec, &w, &h);
    if (e)
        return e;
    fill_info(&d, comps, w, h, info);
    return E_NONE;
}

int32_t jxso_probe(const uint8_t* cs, uint32_t len, jxso_image_info_t* info) { return jxso_probe_proxy(cs, len, 0, info); }

int32_t jxso_decode(const uint8_t* cs, uint32_t len, void* const planes[4], jxso_image_info_t* info) {
    return jxso_decode_proxy(cs, len, 0, planes, info);
}

int32_t jxso_decode_proxy(const uint8_t* cs, uint32_t len, uint32_t proxy, void* const planes[4],
                          jxso_image_info_t* info) {
    dec_hdr d;
    int32_t e = parse_header(cs, len, &d);
    if (e)
        return e;
    ob_pi* pi = &d.pi;
    ob_comp ocomp[4];
    uint32_t npk_dec, ow, oh;
    e = proxy_geometry(pi, proxy, ocomp, &npk_dec, &ow, &oh);
    if (e)
        return e;
    if (info)
        fill_info(&d, ocomp, ow, oh, info);
    uint16_t* coef[4][NBC];
    uint8_t* gcl[4][NBC];
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            const ob_band* bd = &pi->c[c].b[b];
            coef[c][b] = (uint16_t*)calloc((size_t)bd->w * bd->h, 2);
            gcl[c][b] = (uint8_t*)calloc((size_t)gcli_w(bd->w) * bd->h, 1);
        }
    uint8_t* sflag = (uint8_t*)malloc(4096);
    uint64_t pos = d.hdr_end;
    e = E_NONE;
    for (uint32_t s = 0; s < pi->slices && !e; s++) {
        br_t r = {cs, pos * 8, (uint64_t)len * 8, 0};
        if (br_get(&r, 16) != 0xFF20 || br_get(&r, 16) != 4 || br_get(&r, 16) != s) {
            e = E_INVALID_BS;
            break;
        }
        uint32_t first_row = s * pi->pps;
        uint32_t nprec = s == pi->slices - 1 ? pi->rows - first_row : pi->pps;
        uint8_t prev_gtli[4][NBC] = {{0}};
        for (uint32_t i = 0; i < nprec && !e; i++) {
            uint32_t row = first_row + i, type = row + 1 >= pi->rows ? 1 : 0;
            uint64_t pstart = r.bit;
            uint32_t lprc = br_get(&r, 24);
            uint8_t q = (uint8_t)br_get(&r, 8), rr = (uint8_t)br_get(&r, 8);
            uint8_t dmode[4][NBC] = {{0}};
            for (uint32_t g = 0; g < pi->nb_all; g++) {
                if (pi->gb[g] == NE)
                    continue;
                dmode[pi->gc[g]][pi->gb[g]] = (uint8_t)br_get(&r, 2);
            }
            br_align(&r);
            uint64_t pend = pstart + 8ull * (bytes_of_bits(40 + 2 * pi->nb_exist) + lprc);
            if (r.err || pend > r.end_bit) {
                e = E_TOO_SHORT;
                break;
            }
            uint8_t gtli[4][NBC];
            for (uint32_t c = 0; c < pi->nc; c++)
                for (uint32_t b = 0; b < pi->c[c].nb; b++) {
                    gtli[c][b] = truncation(pi->c[c].b[b].gain, pi->c[c].b[b].prio, q, rr);
                    /* the first precinct of a slice has nothing above it
                     * (Packing.c:536-542) */
                    if ((dmode[c][b] & 1) && i == 0)
                        e = E_INVALID_BS;
                }
            for (uint32_t p = 0; p < npk_dec && !e; p++) {
                uint32_t l = pi->pk_line[p];
                int any = 0;
                for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++)
                    if (l < pi->lines[type][pi->gc[gi]][pi->gb[gi]])
                        any = 1;
                if (!any)
                    continue;
                uint32_t raw = br_get(&r, 1), ld, lg, ls;
                if (pi->short_hdr) {
                    ld = br_get(&r, 15);
                    lg = br_get(&r, 13);
                    ls = br_get(&r, 11);
                } else {
                    ld = br_get(&r, 20);
                    lg = br_get(&r, 20);
                    ls = br_get(&r, 15);
                }
                if (raw) {
                    uint64_t gend = r.bit + 8ull * lg;
                    for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                        uint32_t c = pi->gc[gi], b = pi->gb[gi];
                        if (l >= pi->lines[type][c][b])
                            continue;
                        const ob_band* bd = &pi->c[c].b[b];
                        uint8_t* g = gcl[c][b] + (size_t)(row * bd->hl + l) * gcli_w(bd->w);
                        for (uint32_t k = 0; k < gcli_w(bd->w); k++)
                            g[k] = (uint8_t)br_get(&r, 4);
                    }
                    r.bit = gend;
                } else {
                    /* significance sub-packet, then bit-plane counts */
                    uint64_t sstart = r.bit;
                    uint32_t nsig = 0;
                    uint32_t soff[4][NBC];
                    for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                        uint32_t c = pi->gc[gi], b = pi->gb[gi];
                        if (l >= pi->lines[type][c][b] || !(dmode[c][b] & 2))
                            continue;
                        soff[c][b] = nsig;
                        for (uint32_t k = 0; k < sig_w(pi->c[c].b[b].w); k++)
                            sflag[nsig++] = (uint8_t)br_get(&r, 1);
                    }
                    br_align(&r);
                    uint32_t sbytes = (uint32_t)((r.bit - sstart) >> 3);
                    uint64_t gend = r.bit + 8ull * lg;
                    for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                        uint32_t c = pi->gc[gi], b = pi->gb[gi];
                        if (l >= pi->lines[type][c][b])
                            continue;
                        const ob_band* bd = &pi->c[c].b[b];
                        uint32_t gw = gcli_w(bd->w);
                        uint8_t* g = gcl[c][b] + (size_t)(row * bd->hl + l) * gw;
                        const uint8_t gt = gtli[c][b];
                        /* vertical prediction (Packing.c:249-385): from the
                         * line above, the band's previous line or the last
                         * line of the precinct above */
                        const int vp = dmode[c][b] & 1;
                        const uint8_t* top = NULL;
                        uint8_t t = gt;
                        if (vp) {
                            const uint8_t gtop = l > 0 ? gt : prev_gtli[c][b];
                            t = gtop > gt ? gtop : gt;
                            top = l > 0 ? g - gw : gcl[c][b] + (size_t)((row - 1) * bd->hl + pi->lines[0][c][b] - 1) * gw;
                        }
                        for (uint32_t k = 0; k < gw; k++) {
                            const int32_t mt = vp ? (top[k] > t ? top[k] : t) : 0;
                            if ((dmode[c][b] & 2) && sflag[soff[c][b] + k / 8]) {
                                g[k] = (uint8_t)(vp && !d.rm ? mt : gt);
                                continue;
                            }
                            uint32_t n = 0;
                            while (br_get(&r, 1) == 1 && !r.err && n < 64)
                                n++;
                            if (!vp) {
                                g[k] = (uint8_t)(gt + n);
                                continue;
                            }
                            const int32_t th = mt - gt > 0 ? mt - gt : 0;
                            int32_t dm = 0;
                            if ((int32_t)n > 2 * th)
                                dm = (int32_t)n - th;
                            else if (n > 0)
                                dm = (n & 1) ? -(int32_t)((n + 1) / 2) : (int32_t)(n / 2);
                            g[k] = (uint8_t)(mt + dm);
                        }
                    }
                    if (r.bit > gend)
                        e = E_INVALID_BS;
                    r.bit = gend;
                    (void)sbytes;
                }
                uint64_t dend = r.bit + 8ull * ld;
                uint32_t pad_signs[4][NBC] = {{0}};
                for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p] && !e; gi++) {
                    uint32_t c = pi->gc[gi], b = pi->gb[gi];
                    if (l >= pi->lines[type][c][b])
                        continue;
                    const ob_band* bd = &pi->c[c].b[b];
                    uint32_t y = row * bd->hl + l;
                    uint8_t* g = gcl[c][b] + (size_t)y * gcli_w(bd->w);
                    uint16_t* v = coef[c][b] + (size_t)y * bd->w;
                    uint8_t gt = gtli[c][b];
                    for (uint32_t k = 0; k < gcli_w(bd->w); k++) {
                        if (g[k] <= gt || g[k] > 15)
                            continue;
                        uint16_t val[4] = {0};
                        uint32_t sg = d.fs ? 0 : br_get(&r, 4);
                        for (int bit = g[k] - 1; bit >= gt; bit--) {
                            uint32_t nib = br_get(&r, 4);
                            for (int t = 0; t < 4; t++)
                                val[t] |= (uint16_t)(((nib >> (3 - t)) & 1) << bit);
                        }
                        for (int t = 0; t < 4; t++) {
                            if (4 * k + t >= bd->w) {
                                /* sign bits of non-zero padding coefficients are
                                 * skipped (Packing.c:205-209, 240-246) */
                                pad_signs[c][b] += val[t] != 0;
                                continue;
                            }
                            uint16_t m = val[t];
                            /* deadzone / uniform reconstruction (Dequant.c) */
                            if (gt > 0 && m) {
                                if (d.qpih == 0) {
                                    m |= (uint16_t)(1u << (gt - 1));
                                } else {
                                    uint8_t sc = g[k] - gt + 1;
                                    uint16_t acc = 0;
                                    for (uint16_t x = m; x > 0; x = (uint16_t)(x >> sc))
                                        acc = (uint16_t)(acc + x);
                                    m = acc;
                                }
                            }
                            v[4 * k + t] = (uint16_t)(m | (((sg >> (3 - t)) & 1) ? 0x8000u : 0));
                        }
                    }
                }
                if (r.bit > dend || r.err)
                    e = E_INVALID_BS;
                r.bit = dend;
                if (d.fs && !e) {
                    /* sign sub-packet: one bit per non-zero coefficient
                     * (unpack_sign, Packing.c:221-247) */
                    uint64_t send = r.bit + 8ull * ls;
                    for (uint32_t gi = pi->pk_bs[p]; gi < pi->pk_be[p]; gi++) {
                        uint32_t c = pi->gc[gi], b = pi->gb[gi];
                        if (l >= pi->lines[type][c][b])
                            continue;
                        const ob_band* bd = &pi->c[c].b[b];
                        uint16_t* v = coef[c][b] + (size_t)(row * bd->hl + l) * bd->w;
                        for (uint32_t k = 0; k < bd->w; k++)
                            if (v[k])
                                v[k] |= (uint16_t)(br_get(&r, 1) << 15);
                        r.bit += pad_signs[c][b];
                    }
                    br_align(&r);
                    if (r.bit > send || r.err)
                        e = E_INVALID_BS;
                    r.bit = send;
                }
            }
            memcpy(prev_gtli, gtli, sizeof(gtli));
            r.bit = pend;
        }
        pos = r.bit >> 3;
    }
    free(sflag);
    if (!e) {
        const uint32_t dz = d.bw - d.depth;
        const int32_t maxv = (1 << d.depth) - 1;
        for (uint32_t c = 0; c < pi->nc; c++) {
            const ob_comp* cp = &ocomp[c]; /* the component as output (proxy: fewer levels, the same first bands) */
            comp_bands32 b32;
            for (uint32_t b = 0; b < cp->nb; b++) {
                size_t n = (size_t)cp->b[b].w * cp->b[b].h;
                b32.band[b] = (int32_t*)malloc(sizeof(int32_t) * n);
                for (size_t i = 0; i < n; i++) {
                    uint16_t v = coef[c][b][i];
                    int16_t sv = (int16_t)((v & 0x8000u) ? -(int32_t)(v & 0x7FFFu) : (int32_t)v);
                    b32.band[b][i] = (int32_t)((uint32_t)(int32_t)sv << d.fq);
                }
            }
            int32_t* plane = (int32_t*)malloc(sizeof(int32_t) * (size_t)cp->w * cp->h);
            idwt_component(cp, &b32, plane);
            for (size_t i = 0; i < (size_t)cp->w * cp->h; i++) {
                int32_t v = plane[i] + ((1 << d.bw) >> 1);
                v = (int32_t)(((int64_t)v + (((int64_t)1 << dz) >> 1)) >> dz);
                if (v < 0)
                    v = 0;
                if (v > maxv)
                    v = maxv;
                if (d.depth <= 8)
                    ((uint8_t*)planes[c])[i] = (uint8_t)v;
                else
                    ((uint16_t*)planes[c])[i] = (uint16_t)v;
            }
            free(plane);
            for (uint32_t b = 0; b < cp->nb; b++)
                free(b32.band[b]);
        }
    }
    for (uint32_t c = 0; c < pi->nc; c++)
        for (uint32_t b = 0; b < pi->c[c].nb; b++) {
            free(coef[c][b]);
            free(gcl[c][b]);
        }
    return e;
}

/* ----------------------------------------------------------------------
 * Test hooks: one line of quantization() (Quant.c:9-60, no gtli == 0 skip)
 * and the 1-D lifting steps, for the known-answer and transform tests.
 * ---------------------------------------------------------------------- */
void jxso_quantize(uint16_t* buf, uint32_t n, const uint8_t* gclis, uint8_t gtli, uint32_t uniform) {
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t gcli = gclis[i / 4];
        if (gcli <= gtli) {
            buf[i] = 0;
            continue;
        }
        const uint16_t sign = buf[i] & 0x8000u, m = buf[i] & 0x7FFFu;
        uint16_t o;
        if (!uniform) {
            o = (uint16_t)((m >> gtli) << gtli);
        } else {
            const uint8_t sc = gcli - gtli + 1;
            const uint16_t d = (uint16_t)((((uint32_t)m << sc) - m + (1u << gcli)) >> (gcli + 1));
            o = (uint16_t)(d << gtli);
        }
        buf[i] = o ? (uint16_t)(o | sign) : 0;
    }
}

void jxso_fwd53(const int32_t* x, int32_t* lf, int32_t* hf, uint32_t n) { fwd53(x, 1, lf, 1, hf, 1, n); }

void jxso_inv53(const int32_t* lf, const int32_t* hf, int32_t* x, uint32_t n) { inv53(lf, 1, hf, 1, x, 1, n); }
