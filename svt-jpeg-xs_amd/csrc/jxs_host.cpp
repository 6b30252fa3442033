/*
 * jxs_host.cpp -- host-side coding configuration for the MI355X path.
 *
 * Restates, for the default coding tools (RC mode 1, no vertical prediction,
 * no separate signs, one precinct column), the reference's init-time
 * computations: parameter checks (EncHandle.c:128-454), band layout and
 * packet inclusion (Pi.c:21-406), weight tables (WeightTable.c), Q/R bounds
 * (PiEnc.c:11-53), frame header (PackHeaders.c:11-140) and slice budgets
 * (EncHandle.c:760-809).  The result is a JxsGeom the kernels read from HBM.
 */
#include "jxs_host.h"

#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

namespace jxs {

enum : int32_t {
    E_NONE = 0,
    E_BAD_PARAM = (int32_t)0x80001005,
    E_INVALID_BS = (int32_t)0x80003001,
    E_TOO_SHORT = (int32_t)0x80003003,
};

enum { F_YUV400 = 1, F_YUV420 = 2, F_YUV422 = 3, F_YUV444 = 4, F_4COMP = 5, F_GRAY = 6, F_PACKED444 = 21 };
static const uint8_t NE = 0xFF;

static int sampling(uint32_t fmt, uint32_t& nc, uint32_t sx[4], uint32_t sy[4]) {
    for (int i = 0; i < 4; i++)
        sx[i] = sy[i] = 1;
    switch (fmt) {
    case F_YUV422: nc = 3; sx[1] = sx[2] = 2; return 0;
    case F_YUV420: nc = 3; sx[1] = sx[2] = sy[1] = sy[2] = 2; return 0;
    case F_YUV444:
    case F_PACKED444: nc = 3; return 0;
    case F_4COMP: nc = 4; return 0;
    case F_GRAY: nc = 1; return 0;
    default: return -1;
    }
}

static inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

int build_layout(JxsGeom& g, uint32_t nc, const uint32_t sx[4], const uint32_t sy[4], uint32_t W, uint32_t H, uint32_t dh,
                 uint32_t dv, uint32_t slice_h) {
    if (dv > dh || dh > 5 || dv > 2 || nc == 0 || nc > JXS_MAXC || slice_h == 0)
        return -1;
    g.nc = nc;
    g.w = W;
    g.h = H;
    g.dh = dh;
    g.dv = dv;
    const uint32_t ph = 1u << dv;
    if (slice_h < H && (slice_h % ph))
        return -1;
    g.pps = div_up(slice_h, ph);
    g.slices = div_up(H, slice_h);
    g.nb_exist = 0;
    for (uint32_t c = 0; c < nc; c++) {
        JxsComp& cp = g.c[c];
        cp.sx = sx[c];
        cp.sy = sy[c];
        cp.w = W / sx[c];
        cp.h = H / sy[c];
        if (ph / sy[c] == 0 || dv < sy[c] / 2)
            return -1;
        cp.dh = dh;
        cp.dv = dv - sy[c] / 2;
        cp.nb = 2 * cp.dv + cp.dh + 1;
        if (cp.nb > JXS_NBC)
            return -1;
        g.nb_exist += cp.nb;
    }
    g.nb_all = g.c[0].nb * nc;
    /* global band order with gaps for components with fewer vertical levels
     * (Pi.c:268-303) */
    const uint32_t nb_global = 2 * dv + dh + 1;
    memset(g.gc, NE, sizeof(g.gc));
    memset(g.gb, NE, sizeof(g.gb));
    for (uint32_t c = 0; c < nc; c++) {
        JxsComp& cp = g.c[c];
        const uint32_t skip = nb_global - cp.nb, tail = 3 * cp.dv;
        uint32_t gid = c;
        for (uint32_t i = 0; i < cp.nb; i++) {
            cp.b[i].gid = gid;
            g.gc[gid] = (uint8_t)c;
            g.gb[gid] = (uint8_t)i;
            gid += nc;
            if (skip > 0 && i == cp.nb - tail + 1 - skip)
                gid += nc * skip;
        }
    }
    /* band sizes (Pi.c:318-368) */
    for (uint32_t c = 0; c < nc; c++) {
        JxsComp& cp = g.c[c];
        uint32_t lvl = cp.nb - 1, w = cp.w, h = cp.h;
        for (uint32_t d = 0; d < cp.dv; d++) {
            const uint32_t w2 = w / 2, w1 = w - w2, h2 = h / 2, h1 = h - h2;
            cp.b[lvl].w = w2;
            cp.b[lvl].h = h2;
            cp.b[lvl - 1].w = w1;
            cp.b[lvl - 1].h = h2;
            cp.b[lvl - 2].w = w2;
            cp.b[lvl - 2].h = h1;
            lvl -= 3;
            w = w1;
            h = h1;
        }
        for (uint32_t d = cp.dv; d < cp.dh; d++) {
            const uint32_t w2 = w / 2;
            cp.b[lvl].w = w2;
            cp.b[lvl].h = h;
            lvl -= 1;
            w -= w2;
        }
        cp.b[0].w = w;
        cp.b[0].h = h;
        for (uint32_t b = 0; b < cp.nb; b++) {
            JxsBand& bd = cp.b[b];
            if (bd.w == 0 || bd.h == 0)
                return -1;
            bd.hl = (cp.dv == 2 && b >= cp.nb - 3) ? 2 : 1;
            bd.gw = div_up(bd.w, 4);
            bd.sw = div_up(bd.gw, 8);
        }
    }
    /* packets (Pi.c:131-187) */
    uint32_t n = 0;
    const uint32_t first = (dh > dv ? dh - dv : dv - dh) + 1;
    g.pk_line[n] = 0;
    g.pk_bs[n] = 0;
    g.pk_be[n] = first * nc;
    n++;
    for (uint32_t lv = first; lv < g.c[0].nb; lv++, n++) {
        g.pk_line[n] = 0;
        g.pk_bs[n] = g.pk_be[n] = lv * nc;
        for (uint32_t i = lv * nc; i < (lv + 1) * nc; i++)
            if (g.gb[i] != NE)
                g.pk_be[n]++;
    }
    for (uint32_t lv = first; lv < g.c[0].nb; lv++) {
        if (g.c[0].b[lv].hl == 1)
            continue;
        g.pk_line[n] = 1;
        g.pk_bs[n] = g.pk_be[n] = lv * nc;
        for (uint32_t i = lv * nc; i < (lv + 1) * nc; i++)
            if (g.gb[i] != NE && g.c[g.gc[i]].b[g.gb[i]].hl > 1)
                g.pk_be[n]++;
        n++;
    }
    g.npk = n;
    /* packing units: the band lines of a precinct, packet by packet */
    g.n_units = 0;
    for (uint32_t p = 0; p < g.npk; p++)
        for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
            if (g.n_units >= JXS_NU)
                return -1;
            g.u_p[g.n_units] = (uint8_t)p;
            g.u_gi[g.n_units++] = (uint8_t)gi;
        }
    memset(g.ex_gid, NE, sizeof(g.ex_gid));
    for (uint32_t gi = 0, e = 0; gi < g.nb_all; gi++)
        if (g.gb[gi] != NE)
            g.ex_gid[e++] = (uint8_t)gi;
    g.short_hdr = (W * nc) < 32752;
    g.pkt_hdr_bytes = g.short_hdr ? 5 : 7;
    g.prec_hdr_bytes = div_up(40 + 2 * g.nb_exist, 8);
    /* precinct rows (Pi.c:21-39, one precinct column) */
    g.rows = div_up(H, ph);
    uint32_t bl = 0;
    for (uint32_t c = 0; c < nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++) {
            JxsBand& bd = g.c[c].b[b];
            bd.lines[0] = bd.hl;
            bd.lines[1] = bd.h - (g.rows - 1) * bd.hl;
            bd.bl = bl;
            bl += bd.hl;
        }
    g.n_bl = bl;
    for (int t = 0; t < 2; t++) {
        uint32_t exist = 0;
        for (uint32_t p = 0; p < g.npk; p++) {
            uint32_t raw_bits = 0, any = 0;
            for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
                const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
                if (g.pk_line[p] < bd.lines[t]) {
                    any = 1;
                    raw_bits += bd.gw * 4;
                }
            }
            exist += any;
            g.raw_bytes[t][p] = div_up(raw_bits, 8);
        }
        g.pk_exist[t] = t == 0 ? g.npk : exist;
        g.rc_hdr_bytes[t] = g.prec_hdr_bytes + div_up(g.pkt_hdr_bytes * 8 * g.pk_exist[t], 8);
    }
    if (g.n_bl > JXS_NBL)
        return -1;
    return 0;
}

/* The rate control's unit and line words (once gains and priorities are
 * known): every band line of a precinct carries its band's parameters so a
 * wave can walk the lines in packet order from registers. */
static int unit_words(JxsGeom& g) {
    memset(g.u_info, 0, sizeof(g.u_info));
    memset(g.ln_info, 0, sizeof(g.ln_info));
    for (uint32_t u = 0; u < g.n_units; u++) {
        const uint32_t p = g.u_p[u], gi = g.u_gi[u], l = g.pk_line[p];
        if (g.gb[gi] == NE)
            return -1;
        const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
        const uint32_t line = bd.bl + (l < bd.hl ? l : 0);
        const uint32_t last = u + 1 == g.n_units || g.u_p[u + 1] != p ? 1u : 0u;
        if (bd.gain > 255 || bd.prio > 255 || line >= g.n_bl)
            return -1;
        g.u_info[u] = line | (bd.gain << 8) | (bd.prio << 16) | ((l < bd.lines[0] ? 1u : 0u) << UI_NORMAL) |
                      ((l < bd.lines[1] ? 1u : 0u) << (UI_NORMAL + 1)) | (last << UI_LAST) | (p << UI_PKT);
    }
    /* packing order: the units of the widest band lines of every precinct
     * first, so the longest packing waves start first */
    {
        uint32_t order[JXS_NU];
        for (uint32_t u = 0; u < g.n_units; u++)
            order[u] = u;
        auto width = [&](uint32_t u) {
            const uint32_t gi = g.u_gi[u];
            return g.c[g.gc[gi]].b[g.gb[gi]].gw;
        };
        for (uint32_t a = 1; a < g.n_units; a++) /* stable insertion sort, widest first */
            for (uint32_t b = a; b > 0 && width(order[b]) > width(order[b - 1]); b--) {
                const uint32_t t = order[b];
                order[b] = order[b - 1];
                order[b - 1] = t;
            }
        memset(g.u_rank, 0, sizeof(g.u_rank));
        for (uint32_t r = 0; r < g.n_units; r++)
            g.u_rank[order[r]] = (uint8_t)r;
    }
    memset(g.rq, 0, sizeof(g.rq));
    memset(g.rq_raw, 0, sizeof(g.rq_raw));
    for (uint32_t t = 0; t < 2; t++) {
        uint32_t n = 0, prev_p = ~0u;
        for (uint32_t u = 0; u < g.n_units; u++) {
            const uint32_t x = g.u_info[u];
            if (!((x >> (UI_NORMAL + t)) & 1u))
                continue;
            const uint32_t p = x >> UI_PKT, line = x & 0xFFu, gain = (x >> 8) & 0xFFu, prio = (x >> 16) & 0xFFu;
            if (n > 0 && prev_p == p)
                g.rq[t][n - 1] &= ~(1u << RQ_END);
            g.rq[t][n] = line * JXS_RT_STRIDE | (gain << RQ_GAIN) | (prio << RQ_PRIO) | (1u << RQ_END);
            g.rq_raw[t][n] = g.raw_bytes[t][p];
            prev_p = p;
            n++;
        }
        g.rq_n[t] = n;
        /* the walk cut into 1, 2 and 4 runs of whole packets whose longest
         * run is as short as it can be: several waves then walk one
         * precinct's lines in parallel */
        uint32_t runs[JXS_NU], nr = 0, len = 0;
        for (uint32_t k = 0; k < n; k++) {
            len++;
            if ((g.rq[t][k] >> RQ_END) & 1u) {
                runs[nr++] = len;
                len = 0;
            }
        }
        for (uint32_t lg = 0; lg < 3; lg++) {
            const uint32_t parts = 1u << lg;
            uint32_t cap = 0;
            for (uint32_t k = 0; k < nr; k++)
                cap = runs[k] > cap ? runs[k] : cap;
            for (;; cap++) { /* the smallest cap that needs at most `parts` runs */
                uint32_t used = 1, cur = 0;
                for (uint32_t k = 0; k < nr; k++) {
                    if (cur + runs[k] > cap) {
                        used++;
                        cur = 0;
                    }
                    cur += runs[k];
                }
                if (used <= parts || cap >= n)
                    break;
            }
            uint32_t part = 0, cur = 0, pos = 0;
            g.rq_seg[t][lg][0] = 0;
            for (uint32_t k = 0; k < nr; k++) {
                if (cur + runs[k] > cap) {
                    g.rq_seg[t][lg][++part] = (uint8_t)pos;
                    cur = 0;
                }
                cur += runs[k];
                pos += runs[k];
            }
            while (part < parts)
                g.rq_seg[t][lg][++part] = (uint8_t)n;
            for (uint32_t k = parts + 1; k < 8; k++)
                g.rq_seg[t][lg][k] = (uint8_t)n;
        }
    }
    for (uint32_t c = 0; c < g.nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++) {
            const JxsBand& bd = g.c[c].b[b];
            if (bd.gw >= 4096 || bd.sw >= 256 || bd.lines[0] > 2 || bd.lines[1] > 2)
                return -1;
            for (uint32_t k = 0; k < bd.hl; k++)
                g.ln_info[bd.bl + k] = bd.gw | (k << LI_SECOND) | (bd.lines[0] << LI_NL) | (bd.lines[1] << (LI_NL + 2)) |
                                       (bd.sw << LI_SW);
        }
    return 0;
}

/* ---------------- weight tables ----------------
 * ISO/IEC 21122 default gain/priority tables, as carried by
 * WeightTable.c:18-74 (values must be identical), indexed by global band. */
struct WTab {
    uint8_t g[JXS_NGB], p[JXS_NGB];
    uint32_t n;
};

#define X NE
static const uint8_t kG444[3][30] = {
    {3, 2, 2, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0},
    {4, 2, 2, 3, 2, 2, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0},
    {4, 3, 3, 3, 2, 2, 3, 2, 2, 2, 1, 1, 2, 1, 1, 2, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0}};
static const uint8_t kP444[3][30] = {
    {6, 8, 7, 1, 4, 5, 12, 14, 15, 0, 2, 3, 9, 11, 10, 13, 16, 17},
    {21, 1, 0, 15, 19, 18, 5, 9, 8, 14, 17, 16, 2, 4, 3, 7, 13, 11, 6, 12, 10, 20, 23, 22},
    {12, 15, 14, 3, 11, 10, 24, 26, 27, 0, 4, 5, 18, 21, 20, 19, 23, 22, 13, 16, 17, 2, 9, 6, 1, 7, 8, 25, 28, 29}};
static const uint8_t kG422[3][30] = {
    {2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {3, 3, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
static const uint8_t kP422[3][30] = {
    {8, 7, 6, 5, 3, 4, 13, 12, 14, 1, 0, 2, 9, 11, 10, 16, 15, 17},
    {0, 1, 2, 18, 19, 20, 6, 7, 8, 15, 16, 17, 3, 4, 5, 10, 12, 14, 9, 11, 13, 21, 22, 23},
    {14, 13, 12, 9, 11, 10, 25, 24, 26, 0, 1, 2, 19, 20, 18, 23, 22, 21, 17, 15, 16, 4, 8, 5, 3, 6, 7, 28, 27, 29}};
static const uint32_t kN44x[3] = {18, 24, 30};
/* 4:2:0, [V-1][H-1] */
static const uint8_t kG420[2][5][30] = {
    {{1, 1, 1, 1, 0, 0, 1, X, X, 0, X, X},
     {1, 1, 1, 1, 0, 0, 1, 0, 0, 1, X, X, 0, X, X},
     {2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, X, X, 0, X, X},
     {2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, X, X, 0, X, X},
     {3, 3, 3, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, X, X, 0, X, X}},
    {{0},
     {2, 1, 1, 1, 1, 1, 1, X, X, 0, X, X, 1, 1, 1, 1, 1, 1, 0, 0, 0},
     {2, 2, 2, 2, 1, 1, 1, 1, 1, 1, X, X, 0, X, X, 1, 1, 1, 1, 1, 1, 0, 0, 0},
     {3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, X, X, 0, X, X, 1, 1, 1, 1, 1, 1, 0, 0, 0},
     {3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, X, X, 0, X, X, 1, 1, 1, 1, 1, 1, 0, 0, 0}}};
static const uint8_t kP420[2][5][30] = {
    {{2, 5, 4, 6, 1, 0, 7, X, X, 3, X, X},
     {4, 7, 6, 8, 1, 0, 10, 3, 2, 9, X, X, 5, X, X},
     {11, 3, 2, 6, 9, 8, 10, 1, 0, 13, 5, 4, 12, X, X, 7, X, X},
     {7, 12, 11, 14, 3, 2, 6, 10, 9, 13, 1, 0, 16, 5, 4, 15, X, X, 8, X, X},
     {9, 16, 15, 7, 13, 12, 17, 3, 2, 6, 11, 10, 14, 1, 0, 19, 5, 4, 18, X, X, 8, X, X}},
    {{0},
     {10, 4, 3, 2, 9, 8, 1, X, X, 0, X, X, 16, 14, 12, 15, 13, 11, 7, 6, 5},
     {5, 12, 11, 13, 4, 3, 2, 10, 9, 1, X, X, 0, X, X, 19, 17, 15, 18, 16, 14, 8, 7, 6},
     {22, 6, 5, 7, 14, 13, 15, 4, 3, 2, 12, 11, 1, X, X, 0, X, X, 21, 19, 17, 20, 18, 16, 10, 9, 8},
     {0, 13, 12, 25, 7, 6, 8, 17, 16, 18, 5, 4, 3, 15, 14, 2, X, X, 1, X, X, 24, 22, 20, 23, 21, 19, 11, 10, 9}}};
static const uint32_t kN420[2][5] = {{12, 15, 18, 21, 24}, {0, 21, 24, 27, 30}};
#undef X

static int wtab_default(uint32_t fmt, uint32_t v, uint32_t h, WTab& t) {
    if (h == 5 && (fmt == F_YUV444 || fmt == F_PACKED444 || fmt == F_YUV422)) {
        const bool is422 = fmt == F_YUV422;
        t.n = kN44x[v];
        memcpy(t.g, is422 ? kG422[v] : kG444[v], t.n);
        memcpy(t.p, is422 ? kP422[v] : kP444[v], t.n);
        return 0;
    }
    if (fmt == F_YUV420 && v >= 1 && h >= 1 && h >= v && kN420[v - 1][h - 1]) {
        t.n = kN420[v - 1][h - 1];
        memcpy(t.g, kG420[v - 1][h - 1], t.n);
        memcpy(t.p, kP420[v - 1][h - 1], t.n);
        return 0;
    }
    return -1;
}

/* Tables for 4:2:2 / 4:4:4 at other levels are derived from the V2/H5 one
 * (WeightTable.c:128-334): remove the entries of the dropped levels, then
 * close the gaps of the priority order. */
static void wtab_drop(WTab& t, uint32_t grp) {
    for (uint32_t i = 3 * grp; i + 3 < t.n; i++) {
        t.g[i] = t.g[i + 3];
        t.p[i] = t.p[i + 3];
    }
    t.n -= 3;
}

static int wtab_derived(uint32_t fmt, uint32_t dv, uint32_t dh, WTab& t) {
    if (fmt != F_YUV422 && fmt != F_YUV444 && fmt != F_PACKED444)
        return -1;
    if (wtab_default(fmt, 2, 5, t) || dh < dv)
        return -1;
    for (int v = 2; v > (int)dv; v--) {
        wtab_drop(t, 4);
        wtab_drop(t, 0);
    }
    for (int h = 5; h > (int)dh;) {
        h--;
        if (h != 1)
            wtab_drop(t, 0);
        else
            t.n -= 3;
    }
    for (uint32_t i = 0; i < t.n; i++) {
        for (;;) {
            bool found = false;
            for (uint32_t j = 0; j < t.n && !found; j++)
                found = t.p[j] == i;
            if (found)
                break;
            for (uint32_t j = 0; j < t.n; j++)
                if (t.p[j] > i) {
                    t.p[j]--;
                    break;
                }
        }
    }
    return 0;
}

static inline uint32_t truncation(uint32_t gain, uint32_t prio, uint32_t q, uint32_t r) {
    const uint32_t pump = prio < r ? 1 : 0;
    if (q < gain + pump)
        return 0;
    const uint32_t t = q - gain - pump;
    return t > 15 ? 15 : t;
}

static bool all_truncated(const JxsGeom& g, uint32_t q, uint32_t r) {
    for (uint32_t c = 0; c < g.nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++)
            if (truncation(g.c[c].b[b].gain, g.c[c].b[b].prio, q, r) != 15)
                return false;
    return true;
}

/* ---------------- frame header ---------------- */
struct BitOut {
    uint8_t* m;
    uint32_t bit;
    void put(uint32_t v, uint32_t n) {
        for (int i = (int)n - 1; i >= 0; i--, bit++)
            if ((v >> i) & 1)
                m[bit >> 3] |= (uint8_t)(0x80u >> (bit & 7));
    }
    void align() { bit = (bit + 7) & ~7u; }
};

static uint32_t write_header(JxsGeom& g) {
    memset(g.hdr, 0, sizeof(g.hdr));
    BitOut o{g.hdr, 0};
    o.put(0xFF10, 16); /* SOC */
    o.put(0xFF50, 16); /* CAP */
    o.put(4, 16);
    for (int i = 0; i < 9; i++)
        o.put(i == 4 ? (g.format == F_YUV420) : (i == 8 ? 1u : 0u), 1);
    o.align();
    o.put(0xFF12, 16); /* PIH */
    o.put(26, 16);
    o.put(g.lcod, 32);
    o.put(0, 16);
    o.put(0, 16);
    o.put(g.w, 16);
    o.put(g.h, 16);
    o.put(0, 16); /* Cw */
    o.put(g.pps, 16);
    o.put(g.nc, 8);
    o.put(4, 8);
    o.put(8, 8);
    o.put(20, 8); /* Bw */
    o.put(8, 4);  /* Fq */
    o.put(4, 4);  /* Br */
    o.put(0, 8);
    o.put(g.dh, 4);
    o.put(g.dv, 4);
    o.put(!g.short_hdr, 1);
    o.put(1, 1);
    o.put(g.qpih, 2);
    o.put(g.fs, 2); /* Fs (EncHandle.c:274-278) */
    o.put(g.rm, 2); /* Rm (EncHandle.c:250-253) */
    o.put(0xFF13, 16); /* CDT */
    o.put(2 * g.nc + 2, 16);
    for (uint32_t c = 0; c < g.nc; c++) {
        o.put(g.depth, 8);
        o.put(g.c[c].sx, 4);
        o.put(g.c[c].sy, 4);
    }
    o.put(0xFF14, 16); /* WGT */
    o.put(2 * g.nb_exist + 2, 16);
    for (uint32_t gi = 0; gi < g.nb_all; gi++) {
        if (g.gc[gi] == NE)
            continue;
        const JxsBand& b = g.c[g.gc[gi]].b[g.gb[gi]];
        o.put(b.gain, 8);
        o.put(b.prio, 8);
    }
    o.align();
    return o.bit >> 3;
}

void assign_buffers(JxsGeom& g) {
    uint64_t co = 0, go = 0, so = 0, scr = 0;
    for (uint32_t c = 0; c < g.nc; c++) {
        JxsComp& cp = g.c[c];
        for (uint32_t b = 0; b < cp.nb; b++) {
            JxsBand& bd = cp.b[b];
            bd.coeff_off = co;
            bd.gcli_off = go;
            bd.smax_off = so;
            bd.cs = 4 * bd.gw;
            co += (uint64_t)bd.cs * bd.h;
            go += (uint64_t)bd.gw * bd.h;
            so += (uint64_t)bd.sw * bd.h;
        }
        if (cp.dv > 0)
            scr += (uint64_t)cp.w * cp.h + (uint64_t)div_up(cp.w, 2) * div_up(cp.h, 2);
        co = (co + 63) & ~(uint64_t)63;
        go = (go + 255) & ~(uint64_t)255;
        so = (so + 255) & ~(uint64_t)255;
    }
    g.n_coeff = co;
    g.n_gcli = go;
    g.n_smax = so;
    g.n_i32 = scr;
}

int32_t build_encoder_geometry(const EncParams& p, HostGeom& out, int verbose) {
    JxsGeom& g = out.g;
    memset(&g, 0, sizeof(g));
    uint32_t nc, sx[4], sy[4];
    auto bad = [&](const char* why) {
        if (verbose)
            fprintf(stderr, "Error: %s\n", why);
        return E_BAD_PARAM;
    };
    if (p.depth < 8 || p.depth > 14)
        return bad("bit depth must be 8..14");
    if (sampling(p.format, nc, sx, sy))
        return bad("unsupported colour format");
    if (p.decomp_v == 0 && p.format == F_YUV420)
        return bad("YUV420 requires vertical decomposition 1 or 2");
    uint32_t slice_h = p.slice_height;
    if (slice_h == 0)
        return bad("slice_height cannot be 0");
    if (slice_h >= p.height)
        slice_h = p.height;
    else if (slice_h % (1u << p.decomp_v))
        return bad("slice_height must be a multiple of 2^decomp_v");
    if (p.quantization > 1)
        return bad("quantization must be 0 or 1");
    if (p.bpp_den == 0 || p.bpp_num == 0)
        return bad("bpp numerator and denominator must be non-zero");
    const uint64_t bpf = ((uint64_t)p.width * p.height * p.bpp_num / p.bpp_den + 7) / 8;
    if (bpf >= ((uint64_t)1 << 32) || bpf == 0)
        return bad("impossible compression");
    if (p.width < 4)
        return bad("minimum width is 4");
    if ((p.format == F_YUV422 || p.format == F_YUV420) && (p.width % 2))
        return bad("format requires an even width");
    if (p.format == F_YUV420 && (p.height % 2))
        return bad("YUV420 requires an even height");
    if (p.decomp_v > 2 || p.decomp_h > 5 || p.decomp_h < p.decomp_v || (p.decomp_h == 0 && p.decomp_v == 0))
        return bad("decomposition levels out of range");
    uint32_t mw = p.width, mh = p.height, mv = p.decomp_v;
    if (p.format == F_YUV420) {
        mw >>= 1;
        mh >>= 1;
        mv -= 1;
    }
    if (p.format == F_YUV422)
        mw >>= 1;
    for (uint32_t i = 0; i < p.decomp_h && mw; i++)
        mw = mw > 1 ? mw - mw / 2 : 0;
    for (uint32_t i = 0; i < mv && mh; i++)
        mh = mh > 1 ? mh - mh / 2 : 0;
    if (mw == 0 || mh == 0)
        return bad("frame too small for the decomposition");
    if (build_layout(g, nc, sx, sy, p.width, p.height, p.decomp_h, p.decomp_v, slice_h))
        return bad("band layout");
    for (uint32_t c = 0; c < nc; c++)
        if (g.c[c].w > JXS_MAX_ROW)
            return bad("component wider than the row kernels support (8192 samples)");
    g.depth = p.depth;
    g.format = p.format;
    g.packed = p.format == F_PACKED444;
    g.lcod = (uint32_t)bpf;
    g.qpih = p.quantization;
    g.sig = p.significance ? 1 : 0;
    g.rc_mode = p.rc_mode;
    g.vpred = p.vpred;
    g.signs = p.signs;
    g.fs = p.signs ? 1u : 0u;
    g.rm = p.vpred == 2 ? 1u : 0u;
    WTab t;
    bool ok = nc == 3 && wtab_default(p.format, p.decomp_v, p.decomp_h, t) == 0;
    if (!ok && (nc != 3 || wtab_derived(p.format, p.decomp_v, p.decomp_h, t)))
        return bad("no weight table for this configuration");
    for (uint32_t c = 0; c < nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++) {
            JxsBand& bd = g.c[c].b[b];
            if (bd.gid >= t.n)
                return bad("weight table");
            bd.gain = t.g[bd.gid];
            bd.prio = t.p[bd.gid];
        }
    if (unit_words(g) != 0)
        return bad("band layout exceeds the rate control's unit words");
    uint32_t mr = 0;
    for (uint32_t c = 0; c < nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++)
            mr = g.c[c].b[b].prio > mr ? g.c[c].b[b].prio : mr;
    if (mr < 255)
        mr++;
    if (mr > g.nb_all - 1)
        mr = g.nb_all - 1;
    g.maxr = mr;
    uint32_t mq = 31;
    while (mq > 0 && all_truncated(g, mq, mr))
        mq--;
    g.maxq = mq;
    if (g.maxr >= JXS_NR)
        return bad("refinement range");
    for (uint32_t q = 0; q < JXS_NQ; q++) {
        g.empty_mask[q] = 0;
        for (uint32_t r = 0; r < JXS_NR; r++)
            if (all_truncated(g, q, r))
                g.empty_mask[q] |= 1u << r;
    }
    g.hdr_len = write_header(g);
    const uint32_t tags = g.hdr_len + 2 + 6 * g.slices;
    if (tags >= g.lcod)
        return bad("impossible compression, use a bigger bpp");
    /* slice sizes (EncHandle.c:760-809) */
    out.slice_size.assign(g.slices, 0);
    out.slice_off.assign(g.slices, 0);
    out.slice_bud.assign(g.slices, 0);
    {
        const uint32_t all = g.lcod - tags;
        const uint32_t last_prec = g.pps - (g.slices * g.pps - g.rows);
        const uint32_t per_prec = all / g.rows;
        uint32_t per_slice = per_prec * g.pps;
        uint32_t last_slice = per_prec * last_prec;
        uint32_t left = all - last_slice - per_slice * (g.slices - 1);
        if (g.slices > 1)
            per_slice += (left * (g.slices - 1) * g.pps / g.rows) / (g.slices - 1);
        last_slice += left * last_prec / g.rows;
        left = all - last_slice - per_slice * (g.slices - 1);
        uint32_t remain = g.lcod - g.hdr_len, off = g.hdr_len;
        for (uint32_t i = 0; i < g.slices; i++) {
            const uint32_t s = (i != g.slices - 1) ? per_slice + (i < left ? 1 : 0) + 6 : remain;
            out.slice_size[i] = s;
            out.slice_off[i] = off;
            out.slice_bud[i] = s - 6 - (i == g.slices - 1 ? 2 : 0);
            off += s;
            remain -= s;
        }
    }
    assign_buffers(g);
    return E_NONE;
}

/* ---------------- decoder header ---------------- */
int32_t geom_from_header(const CsHeader& h, uint32_t header_end, HostGeom& out, uint32_t* header_end_out) {
    JxsGeom& g = out.g;
    memset(&g, 0, sizeof(g));
    const CsHeader::Const& c = h.c;
    const CsHeader::Dyn& d = h.d;
    /* what this decoder does not decode: components without decomposition
     * (CWD), colour transforms (Cpih), non-linear transforms (NLT), and
     * coefficient precisions other than Bw 18 or 20 */
    if (c.sd != 0 || c.cpih != 0 || d.tnlt != 0 || (d.bw != 20 && d.bw != 18))
        return E_INVALID_BS;
    const uint32_t nc = c.nc;
    if (nc == 0 || nc > JXS_MAXC)
        return E_INVALID_BS;
    /* a precinct width Cw != 0 that still covers the whole picture width is
     * one precinct column, laid out as Cw = 0 (Pi.c:21-72) */
    if (c.cw != 0) {
        uint32_t sxm = 0;
        for (uint32_t k = 0; k < nc; k++)
            sxm = c.sx[k] > sxm ? c.sx[k] : sxm;
        const uint64_t cs = 8ull * c.cw * sxm * (1ull << c.dh);
        if ((uint64_t)(uint32_t)c.w > cs)
            return E_INVALID_BS;
    }
    uint32_t sx[4] = {1, 1, 1, 1}, sy[4] = {1, 1, 1, 1};
    for (uint32_t k = 0; k < nc; k++) {
        sx[k] = c.sx[k];
        sy[k] = c.sy[k];
        if (sx[k] < 1 || sx[k] > 2 || sy[k] < 1 || sy[k] > 2)
            return E_INVALID_BS;
    }
    if (build_layout(g, nc, sx, sy, (uint32_t)c.w, (uint32_t)c.h, (uint32_t)c.dh, (uint32_t)c.dv, c.hsl << c.dv))
        return E_INVALID_BS;
    g.lcod = d.lcod;
    g.depth = c.depth[0]; /* the output scaling uses the first component's depth (NltDec.c:196-200) */
    out.bw = d.bw;
    out.fq = d.fq;
    g.qpih = d.qpih;
    g.fs = d.fs;
    g.rm = d.rm;
    /* long packet headers when the frame asks for them (Lh) or the picture is
     * too wide for short ones (Packing.c:532) */
    if (d.lh) {
        g.short_hdr = 0;
        g.pkt_hdr_bytes = 7;
    }
    /* weights in band order; a table shorter than the band count leaves
     * zeros (Decoder.c:34-49) */
    uint32_t k = 0;
    for (uint32_t gi = 0; gi < g.nb_all; gi++) {
        if (g.gc[gi] == NE)
            continue;
        g.c[g.gc[gi]].b[g.gb[gi]].gain = k < CS_MAX_BANDS ? c.gain[k] : 0;
        g.c[g.gc[gi]].b[g.gb[gi]].prio = k < CS_MAX_BANDS ? c.prio[k] : 0;
        k++;
    }
    assign_buffers(g);
    if (header_end_out)
        *header_end_out = header_end;
    return E_NONE;
}

int32_t parse_codestream_header(const uint8_t* cs, size_t len, HostGeom& out, uint32_t* header_end) {
    CsHeader h;
    uint32_t used = 0;
    int32_t e = cs_get_header(cs, len, h, &used);
    if (!e)
        e = cs_pi_check(h, 0);
    if (!e)
        e = geom_from_header(h, used, out, header_end);
    return e;
}

/* ---------------- forward-transform jobs ---------------- */
static JxsBandOut band_out(const JxsGeom& g, uint32_t c, uint32_t b, uint16_t* coeff, uint8_t* gcli, uint8_t* smax,
                           uint16_t* hist) {
    const JxsBand& bd = g.c[c].b[b];
    JxsBandOut o;
    o.coeff = coeff + bd.coeff_off;
    o.gcli = gcli + bd.gcli_off;
    o.smax = smax + bd.smax_off;
    o.hist = hist;
    o.w = bd.w;
    o.gw = bd.gw;
    o.sw = bd.sw;
    o.cs = bd.cs;
    o.hl = bd.hl;
    o.bl = bd.bl;
    o.n_bl = g.n_bl;
    o.pad_ = 0;
    return o;
}

/* Launch shape of a transform job: columns per thread (8 for one horizontal
 * level, 16, or 32 for five horizontal levels or rows wider than 256 x 16),
 * threads per row (whole waves), rows per block (512 threads for 8 columns
 * per thread, else 256); rows are read 8 samples per load when
 * they are contiguous, every row starts 16-byte aligned and the width is a
 * multiple of 8. */
static std::mutex g_tile_mu;
static std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> g_tile;

TileKey tile_key(const JxsDwtJob& j) { return TileKey{j.kind, j.step, j.vert, j.lev_l, j.w}; }

bool tile_legal(uint32_t lev_l, uint32_t w, uint32_t chunk) {
    if (lev_l < 1 || (chunk != 8 && chunk != 16 && chunk != 32) || (1u << lev_l) > chunk)
        return false;
    const uint32_t block = chunk == 8 ? 512u : 256u;
    return ((w + chunk - 1) / chunk + 63) / 64 * 64 <= block;
}

void tile_set(const TileKey& k, uint32_t chunk) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const auto key = std::make_tuple(k.kind, k.step, k.vert, k.lev_l, k.w);
    if (chunk)
        g_tile[key] = chunk;
    else
        g_tile.erase(key);
}

uint32_t tile_get(const TileKey& k) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const auto it = g_tile.find(std::make_tuple(k.kind, k.step, k.vert, k.lev_l, k.w));
    return it == g_tile.end() ? 0u : it->second;
}

/* waves per precinct of the rate table kernel, per precinct shape (band
 * lines, precincts per slice, precinct rows) */
static std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> g_rate_waves;

uint32_t rate_waves_get(const JxsGeom& g) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const auto it = g_rate_waves.find(std::make_tuple(g.n_bl, g.pps, g.rows));
    return it == g_rate_waves.end() ? 4u : it->second;
}

void rate_waves_set(const JxsGeom& g, uint32_t waves) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const auto key = std::make_tuple(g.n_bl, g.pps, g.rows);
    if (waves == 1 || waves == 2 || waves == 4)
        g_rate_waves[key] = waves;
    else
        g_rate_waves.erase(key);
}

/* comp: the component of a packed (step 3) job, whose src is the row start + comp samples */
static JxsDwtJob finish_job(JxsDwtJob j, uint32_t comp = 0) {
    const uint32_t elem = j.kind == JXS_SRC_U8 ? 1u : j.kind == JXS_SRC_U16 ? 2u : 4u;
    /* vec: 1 = contiguous rows read 16 bytes per load; 2 + c = packed 16-bit
     * rows read 16 bytes per load, component c picked in registers */
    const bool rows16 = (j.stride * elem) % 16 == 0 && j.w % 8 == 0 && j.w >= 8;
    if (j.step == 1)
        j.vec = rows16 && ((uintptr_t)j.src & 15) == 0 ? 1u : 0u;
    else if (j.step == 3 && j.kind == JXS_SRC_U16 && comp < 3)
        j.vec = rows16 && (((uintptr_t)j.src - comp * elem) & 15) == 0 ? 2u + comp : 0u;
    else
        j.vec = 0;
    /* one horizontal level: 8 columns per thread in 512-thread blocks (half
     * the registers, twice the waves in flight) */
    if (j.lev_l == 1 && (j.w + 7) / 8 <= 512)
        j.chunk = 8;
    else
        j.chunk = (j.lev_l >= 5 || (j.w + 15) / 16 > 256) ? 32u : 16u;
    /* the autotuner's choice for this job shape, when it made one */
    const uint32_t tuned = tile_get(tile_key(j));
    if (tuned && tile_legal(j.lev_l, j.w, tuned))
        j.chunk = tuned;
    const uint32_t block = j.chunk == 8 ? 512u : 256u;
    j.tpr = ((j.w + j.chunk - 1) / j.chunk + 63) / 64 * 64;
    j.rpb = block / j.tpr;
    j.nblk = (j.rows + j.rpb - 1) / j.rpb;
    return j;
}

void fwd_jobs(const JxsGeom& g, const void* const planes[4], const uint32_t strides[4], uint16_t* coeff, uint8_t* gcli,
              uint8_t* smax, uint16_t* hist, int32_t* scratch, FwdJobs& jobs) {
    int32_t* scr = scratch;
    for (uint32_t c = 0; c < g.nc; c++) {
        const JxsComp& cp = g.c[c];
        const uint32_t nb = cp.nb;
        const uint32_t elem = g.depth <= 8 ? 1 : 2;
        JxsDwtJob j;
        memset(&j, 0, sizeof(j));
        j.src = g.packed ? (const void*)((const uint8_t*)planes[0] + (size_t)c * elem) : planes[c];
        j.kind = g.depth <= 8 ? JXS_SRC_U8 : JXS_SRC_U16;
        j.stride = g.packed ? strides[0] : strides[c];
        j.step = g.packed ? 3 : 1;
        j.w = cp.w;
        j.h = cp.h;
        j.shift = 20 - g.depth;
        j.mask = (1u << g.depth) - 1;
        j.offset = 1 << 19;
        auto band = [&](uint32_t b) { return band_out(g, c, b, coeff, gcli, smax, hist); };
        if (cp.dv == 0) {
            /* horizontal levels of every row */
            j.rows = cp.h;
            j.lev_l = cp.dh;
            for (uint32_t k = 0; k < cp.dh; k++)
                j.hf[k] = band(nb - 1 - k);
            j.lf_is_band = 1;
            j.lf = band(0);
            jobs.push(0, finish_job(j, g.packed ? c : 0));
            continue;
        }
        const uint32_t w1 = (cp.w + 1) / 2, h1 = (cp.h + 1) / 2;
        /* first vertical level: high rows -> LH1 / HH1, low rows -> HL1 and
         * (one vertical level) the remaining horizontal levels or (two) the
         * low-pass plane of the second level */
        j.vert = 1;
        j.rows = h1;
        j.hf[0] = band(nb - 3);
        j.lh = band(nb - 2);
        j.hh = band(nb - 1);
        if (cp.dv == 1) {
            j.lev_l = cp.dh;
            for (uint32_t k = 1; k < cp.dh; k++)
                j.hf[k] = band(cp.dh - k);
            j.lf_is_band = 1;
            j.lf = band(0);
            jobs.push(0, finish_job(j, g.packed ? c : 0));
            continue;
        }
        int32_t* B = scr;
        scr += ((size_t)w1 * h1 + 3) & ~(size_t)3; /* every plane 16-byte aligned */
        j.lev_l = 1;
        j.lf_is_band = 0;
        j.lf_i32 = B;
        j.lf_stride = w1;
        jobs.push(0, finish_job(j, g.packed ? c : 0));
        /* second vertical level on the w1 x h1 low-pass plane */
        JxsDwtJob k2;
        memset(&k2, 0, sizeof(k2));
        k2.src = B;
        k2.kind = JXS_SRC_I32;
        k2.stride = w1;
        k2.step = 1;
        k2.w = w1;
        k2.h = h1;
        k2.vert = 1;
        k2.rows = (h1 + 1) / 2;
        k2.lev_l = 1 + (cp.dh - 2);
        k2.hf[0] = band(nb - 6);
        for (uint32_t k = 1; k < k2.lev_l; k++)
            k2.hf[k] = band(cp.dh - 2 - (k - 1));
        k2.lf_is_band = 1;
        k2.lf = band(0);
        k2.lh = band(nb - 5);
        k2.hh = band(nb - 4);
        jobs.push(1, finish_job(k2));
    }
}

}  // namespace jxs
