/*
 * jxs_enc_ordered.hip -- gfx950 rate control and packing for the coding tools
 * whose cost depends on the precinct above: vertical prediction of the
 * bit-plane counts (coding_vertical_prediction_mode 1 / 2) and sign coding in
 * sub-packets of their own (coding_signs_handling 1 fast / 2 full).
 *
 * With vertical prediction a band line's bit-plane counts are coded against
 * the line above, and for the first line of a precinct that is the last line
 * of the precinct above at that precinct's truncation, so the rate control of
 * a precinct needs the result of the one above it (RateControl.c:587-676);
 * with fast sign coding the bytes the packer saves move to the next precinct
 * (PackStageProcess.c:293-333).  Both make the precincts of a slice a chain.
 * k_enc_slice runs one block per slice and walks its precincts in order, as
 * PackStageProcess.c:567-749 does, with the block's 256 threads sharing each
 * step: the transform's band-line histogram records give the zero-prediction
 * costs, one pass over the significance groups the vertical-prediction costs
 * of a truncation, one pass over the coefficients the signs of full sign
 * coding; the binary searches over Q and R (BinarySearch.c:9-78) run on
 * thread 0.  The precinct is then packed into the slice's staging words in
 * the frame's work area (every bit ORed in place after the block has cleared
 * them: each code's position is a block prefix sum), and finally the slice's
 * bytes are copied into the codestream.
 *
 * Frames coded with the default tools are left to k_rate / k_pack (their
 * blocks here return at once), and this kernel's frames to it (k_rate
 * returns for them and they have no packing units).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jxs_geom.h"

#define WAVE 64
#define EO_THREADS 256
#define EO_MAXP 32 /* precincts of a slice held in LDS (rate-control modes 2 and 3) */

__device__ __forceinline__ uint32_t eo_trunc(uint32_t gain, uint32_t prio, uint32_t q, uint32_t r) {
    const uint32_t pump = prio < r ? 1u : 0u;
    if (q < gain + pump)
        return 0;
    const uint32_t t = q - gain - pump;
    return t > 15u ? 15u : t;
}

/* quantization() of one coefficient of a group with bit-plane count gc
 * (Quant.c:9-45; bands at truncation 0 are not quantised,
 * QuantStageProcess.c:20-22) */
__device__ __forceinline__ uint32_t eo_quant(uint32_t v, uint32_t gc, uint32_t gt, uint32_t qpih) {
    if (gc <= gt)
        return 0;
    if (gt == 0)
        return v;
    const uint32_t sign = v & 0x8000u, m = v & 0x7FFFu;
    uint32_t o;
    if (qpih == 0) {
        o = (m >> gt) << gt;
    } else {
        const uint32_t sc = gc - gt + 1;
        const uint32_t d = ((((m << sc) - m) + (1u << gc)) >> (gc + 1)) & 0xFFFFu;
        o = (d << gt) & 0xFFFFu;
    }
    return o ? (o | sign) : 0u;
}

/* bits of a vertical-prediction code without its closing zero
 * (vlc_encode_get_bits, PackPrecinct.h:84-104) */
__device__ __forceinline__ uint32_t eo_vlc(int32_t x, int32_t r, int32_t t) {
    int32_t mx = r - t;
    if (mx < 0)
        mx = 0;
    if (x > mx)
        return (uint32_t)(x + mx);
    x *= 2;
    return (uint32_t)(x < 0 ? -x - 1 : x);
}

/* BinarySearch.c:9-78 */
struct EoSearch {
    int32_t beg, end, step, best, last;
    uint32_t below;
};
__device__ __forceinline__ void eo_bs_init(EoSearch& b, uint32_t beg, uint32_t end, uint32_t below, uint32_t step) {
    b.beg = (int32_t)beg;
    b.end = (int32_t)end;
    b.below = below;
    b.best = -1;
    b.last = -1;
    b.step = step ? (int32_t)step : (b.end - b.beg) / 2;
}
/* result: 0 begin, 1 too small, 2 too big, 3 out of range; returns 0 next, 1 found, 2 error */
__device__ __forceinline__ int32_t eo_bs_next(EoSearch& b, int32_t result, uint32_t* next) {
    if (result == 1) {
        if (b.below && (b.best == -1 || b.best < b.last))
            b.best = b.last;
        if (b.end >= b.last + 1)
            b.beg = b.last + 1;
        else if (b.best == -1)
            return 2;
        else {
            *next = (uint32_t)b.best;
            return 1;
        }
    } else if (result == 2 || result == 3) {
        if (result == 2 && !b.below && (b.best == -1 || b.best > b.last))
            b.best = b.last;
        if (b.beg <= b.last - 1)
            b.end = b.last - 1;
        else if (b.best == -1)
            return 2;
        else {
            *next = (uint32_t)b.best;
            return 1;
        }
    }
    if (b.step > b.end - b.beg)
        b.step = (b.end - b.beg + 1) / 2;
    b.last = b.step + b.beg;
    *next = (uint32_t)b.last;
    return 0;
}

/* a band line present in the precinct being coded, in packet order */
struct EoLine {
    uint32_t gi, c, b, l, pk, last_in_pk;
    uint32_t gw, sw, w, cs, y;
    uint32_t g0, s0, x0; /* flattened first group, significance group, coefficient */
    const uint8_t* gl;   /* bit-plane counts of the line */
    const uint8_t* top;  /* of the line above (vertical prediction), or null */
    const uint8_t* sl;   /* significance-group maxima */
    const uint16_t* cl;  /* coefficients */
};

struct EoCtx {
    JxsGeom g;
    EoLine ln[JXS_NU];
    uint32_t nl, ng, ns, nx; /* lines, and their groups, significance groups, coefficients in all */
    uint32_t row, type, has_top;
    uint32_t gt[JXS_NGB], top_gt[JXS_NGB];
    uint32_t method[JXS_NGB];
    uint32_t vsum[JXS_NU], vred[JXS_NU], nz[JXS_NU]; /* per line: vertical-prediction code bits (+ closing zeros),
                                                          of them in flagged significance groups, non-zero coefficients */
    uint32_t bd[JXS_NU], bg[JXS_NU], bs[JXS_NU], bn[JXS_NU]; /* per line: data, bit-plane-count, significance, sign bits */
    uint32_t chist[JXS_NU][16];                              /* per line: coefficients by magnitude bit length */
    uint32_t pk_raw[JXS_NPK], pk_d[JXS_NPK], pk_g[JXS_NPK], pk_s[JXS_NPK], pk_n[JXS_NPK];
    uint32_t total;
    /* the slice's precincts (rate-control modes 2 and 3) */
    uint32_t sbud[EO_MAXP];
    /* search state */
    uint32_t q, r, err, stop;
    /* packing */
    uint64_t pos;            /* bit position in the slice's staging words */
    uint32_t pk_pos[JXS_NPK]; /* byte offset of each packet inside the precinct */
    uint32_t ln_bit[JXS_NU][4];
    uint32_t scan[EO_THREADS / WAVE];
};

__device__ __forceinline__ uint32_t eo_block_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= o)
            x += y;
    }
    if (lane == WAVE - 1)
        tmp[wid] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t w = 0; w < EO_THREADS / WAVE; w++) {
        const uint32_t t = tmp[w];
        before += w < wid ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

/* ORs the low n (<= 64) bits of v, MSB first, at bit position bp of the
 * staging words (big-endian bit order inside each word). */
__device__ __forceinline__ void eo_put(uint32_t* st, uint64_t bp, uint64_t v, uint32_t n) {
    if (!n)
        return;
    v = n == 64 ? v : (v & ((1ull << n) - 1));
    uint64_t wi = bp >> 5;
    uint32_t off = (uint32_t)(bp & 31); /* bits already used in word wi */
    while (n) {
        const uint32_t room = 32 - off;
        const uint32_t take = n < room ? n : room;
        const uint32_t chunk = (uint32_t)((v >> (n - take)) & ((take == 32) ? 0xFFFFFFFFull : ((1ull << take) - 1)));
        const uint32_t word = chunk << (room - take);
        if (word)
            atomicOr(st + wi, word);
        n -= take;
        wi++;
        off = 0;
    }
}

/* the line of flattened index k (binary search over the lines' first index) */
template <int F>
__device__ __forceinline__ uint32_t eo_find(const EoCtx& C, uint32_t k) {
    uint32_t lo = 0, hi = C.nl - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t f = F == 0 ? C.ln[mid].g0 : F == 1 ? C.ln[mid].s0 : C.ln[mid].x0;
        if (f <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

/* The band lines of precinct row `row` (all threads; ends with a barrier). */
__device__ void eo_lines(EoCtx& C, const JxsFrameDesc& F, uint32_t row, uint32_t has_top) {
    const JxsGeom& g = C.g;
    if (threadIdx.x == 0) {
        C.row = row;
        C.type = row + 1 >= g.rows ? 1u : 0u;
        C.has_top = has_top;
        uint32_t n = 0, ng = 0, ns = 0, nx = 0;
        for (uint32_t u = 0; u < g.n_units; u++) {
            const uint32_t ui = g.u_info[u];
            if (!((ui >> (UI_NORMAL + C.type)) & 1u))
                continue;
            EoLine& L = C.ln[n++];
            L.gi = g.u_gi[u];
            L.c = g.gc[L.gi];
            L.b = g.gb[L.gi];
            const JxsBand& bd = g.c[L.c].b[L.b];
            L.l = (ui & 0xFFu) - bd.bl;
            L.pk = g.u_p[u];
            L.gw = bd.gw;
            L.sw = bd.sw;
            L.w = bd.w;
            L.cs = bd.cs;
            L.y = row * bd.hl + L.l;
            L.g0 = ng;
            L.s0 = ns;
            L.x0 = nx;
            ng += bd.gw;
            ns += bd.sw;
            nx += bd.w;
            L.gl = F.gcli + bd.gcli_off + (size_t)L.y * bd.gw;
            L.sl = F.smax + bd.smax_off + (size_t)L.y * bd.sw;
            L.cl = F.coeff + bd.coeff_off + (size_t)L.y * bd.cs;
            /* the line above: the band's previous line, or its last line in
             * the precinct above (RateControl.c:603-613) */
            L.top = nullptr;
            if (L.l > 0)
                L.top = L.gl - bd.gw;
            else if (has_top)
                L.top = F.gcli + bd.gcli_off + (size_t)((row - 1) * bd.hl + bd.lines[0] - 1) * bd.gw;
        }
        for (uint32_t k = 0; k < n; k++)
            C.ln[k].last_in_pk = k + 1 == n || C.ln[k + 1].pk != C.ln[k].pk;
        C.nl = n;
        C.ng = ng;
        C.ns = ns;
        C.nx = nx;
    }
    __syncthreads();
}

/* Coefficients of each line by magnitude bit length, for full sign coding
 * with deadzone quantisation (the coeff_lookup_table of RateControl.c:126-139). */
__device__ void eo_coeff_hist(EoCtx& C) {
    for (uint32_t k = threadIdx.x; k < C.nl * 16; k += EO_THREADS)
        C.chist[k / 16][k % 16] = 0;
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < C.nx; x += EO_THREADS) {
        const uint32_t li = eo_find<2>(C, x);
        const EoLine& L = C.ln[li];
        const uint32_t m = L.cl[x - L.x0] & 0x7FFFu;
        atomicAdd(&C.chist[li][m ? 32 - __clz(m) : 0], 1u);
    }
    __syncthreads();
}

/* The bytes of the precinct's packets at (q, r) without headers, with the
 * per-line and per-packet choices left in C: precinct_get_budget_bytes
 * (RateControl.c:704-781) over rate_control_calculate_band_best_method
 * (RateControl.c:457-689).  vp: vertical prediction in force, sg: sign coding
 * (0 off, 1 fast, 2 full).  All threads; the total is in C.total. */
__device__ void eo_probe(EoCtx& C, const JxsFrameDesc& F, uint32_t q, uint32_t r, uint32_t vp, uint32_t sg) {
    const JxsGeom& g = C.g;
    const uint32_t tid = threadIdx.x;
    for (uint32_t gi = tid; gi < g.nb_all; gi += EO_THREADS) {
        if (g.gb[gi] == 0xFF)
            continue;
        const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
        C.gt[gi] = eo_trunc(bd.gain, bd.prio, q, r);
    }
    for (uint32_t k = tid; k < C.nl; k += EO_THREADS) {
        C.vsum[k] = 0;
        C.vred[k] = 0;
        C.nz[k] = 0;
    }
    __syncthreads();
    const bool vpred = vp && C.has_top;
    if (vpred) {
        /* one thread per significance group of 8 bit-plane counts
         * (rate_control_calc_vpred_cost_nosigf_c / _sigf_c, RateControl.c:291-361) */
        for (uint32_t k = tid; k < C.ns; k += EO_THREADS) {
            const uint32_t li = eo_find<1>(C, k);
            const EoLine& L = C.ln[li];
            const uint32_t gt = C.gt[L.gi];
            const uint32_t gtop = L.l > 0 ? gt : C.top_gt[L.gi];
            const uint32_t gmax = gt > gtop ? gt : gtop;
            const uint32_t i0 = (k - L.s0) * 8, i1 = min(i0 + 8, L.gw);
            uint32_t sum = 0;
            bool flag = true;
            for (uint32_t i = i0; i < i1; i++) {
                const uint32_t gv = L.gl[i], tv = L.top[i];
                const int32_t mt = (int32_t)(tv > gmax ? tv : gmax);
                const int32_t gc = (int32_t)(gv > gt ? gv : gt);
                sum += eo_vlc(gc - mt, mt, (int32_t)gt) + 1u;
                if (g.rm ? gv > gt : mt != gc)
                    flag = false;
            }
            atomicAdd(&C.vsum[li], sum);
            if (flag)
                atomicAdd(&C.vred[li], sum);
        }
    }
    if (sg == 2 && g.qpih != 0) {
        /* full sign coding, uniform quantisation: the coefficients that stay
         * non-zero (RateControl.c:411-452) */
        for (uint32_t x = tid; x < C.nx; x += EO_THREADS) {
            const uint32_t li = eo_find<2>(C, x);
            const EoLine& L = C.ln[li];
            const uint32_t i = x - L.x0;
            if (eo_quant(L.cl[i], L.gl[i / 4], C.gt[L.gi], 1))
                atomicAdd(&C.nz[li], 1u);
        }
    }
    __syncthreads();
    /* per line: data, sign and zero-prediction bit-plane-count bits from the
     * band-line histogram record (RateControl.c:18-256) */
    const uint16_t* recs = F.hist + (size_t)C.row * F.n_bl * JXS_HREC;
    for (uint32_t k = tid; k < C.nl; k += EO_THREADS) {
        const EoLine& L = C.ln[k];
        const uint32_t gt = C.gt[L.gi];
        const uint16_t* rec = recs + (size_t)(g.c[L.c].b[L.b].bl + L.l) * JXS_HREC;
        uint32_t cnt = 0, data = 0, sig = 0;
        for (uint32_t v = 0; v < 16; v++) {
            const uint32_t h = rec[v];
            if (v <= gt)
                cnt += h;
            else
                data += (v - gt + 1) * h;
        }
        const uint32_t full = L.gw / 8;
        if (gt >= 15) {
            sig = full;
        } else {
            for (uint32_t v = 0; v <= gt; v++)
                sig += rec[16 + v];
        }
        const uint32_t tail = rec[HREC_TAIL];
        const uint32_t zeroed = 8 * sig + (tail != 0xFFFFu && tail <= gt ? L.gw - full * 8 : 0u);
        const uint32_t coded = L.gw - cnt;
        C.bd[k] = 4 * (sg ? data - coded : data);
        uint32_t nsign = 0;
        if (sg == 1) {
            nsign = 4 * coded;
        } else if (sg == 2) {
            if (g.qpih == 0) {
                for (uint32_t v = gt + 1; v < 16; v++)
                    nsign += C.chist[k][v];
            } else {
                nsign = C.nz[k];
            }
        }
        C.bn[k] = nsign;
        /* candidates, stashed for the band's choice: zero prediction without
         * and with significance, vertical prediction without and with */
        C.vsum[k] = vpred ? C.vsum[k] : 0u;
        C.bg[k] = data + cnt;           /* zero prediction, no significance */
        C.bs[k] = data + cnt - zeroed;  /* zero prediction, with significance (flags not counted) */
    }
    __syncthreads();
    /* the band's method (strict comparisons in the reference's order) and its
     * lines' bit counts */
    for (uint32_t gi = tid; gi < g.nb_all; gi += EO_THREADS) {
        if (g.gb[gi] == 0xFF)
            continue;
        uint32_t nosig = 0, withsig = 0, vnosig = 0, vsig = 0, n = 0, sw = 0;
        for (uint32_t k = 0; k < C.nl; k++) {
            if (C.ln[k].gi != gi)
                continue;
            n++;
            sw = C.ln[k].sw;
            nosig += C.bg[k];
            withsig += C.bs[k] + sw;
            vnosig += C.vsum[k];
            vsig += C.vsum[k] - C.vred[k] + sw;
        }
        uint32_t m = 0, best = nosig;
        if (n) {
            if (g.sig && best > withsig) {
                best = withsig;
                m = 1;
            }
            if (vpred) {
                if (best > vnosig) {
                    best = vnosig;
                    m = 2;
                }
                if (g.sig && best > vsig) {
                    best = vsig;
                    m = 3;
                }
            }
        }
        C.method[gi] = m;
    }
    __syncthreads();
    for (uint32_t k = tid; k < C.nl; k += EO_THREADS) {
        const EoLine& L = C.ln[k];
        const uint32_t m = C.method[L.gi];
        const uint32_t gz = C.bg[k], gzs = C.bs[k], v = C.vsum[k], vr = C.vred[k];
        C.bg[k] = m == 0 ? gz : m == 1 ? gzs : m == 2 ? v : v - vr;
        C.bs[k] = (m == 1 || m == 3) ? L.sw : 0u;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t total = 0;
        for (uint32_t p = 0; p < g.npk; p++) {
            C.pk_d[p] = C.pk_g[p] = C.pk_s[p] = C.pk_n[p] = C.pk_raw[p] = 0;
        }
        uint32_t d = 0, gg = 0, s = 0, n = 0;
        for (uint32_t k = 0; k < C.nl; k++) {
            d += C.bd[k];
            gg += C.bg[k];
            s += C.bs[k];
            n += C.bn[k];
            if (C.ln[k].last_in_pk) {
                const uint32_t p = C.ln[k].pk;
                const uint32_t db = (d + 7) / 8, gb = (gg + 7) / 8, sb = (s + 7) / 8, nb = (n + 7) / 8;
                C.pk_d[p] = db;
                C.pk_n[p] = nb;
                total += db + nb;
                const uint32_t raw = g.raw_bytes[C.type][p];
                if (sb + gb > raw) {
                    C.pk_raw[p] = 1;
                    C.pk_s[p] = 0;
                    C.pk_g[p] = raw;
                    total += raw;
                } else {
                    C.pk_raw[p] = 0;
                    C.pk_s[p] = sb;
                    C.pk_g[p] = gb;
                    total += sb + gb;
                }
                d = gg = s = n = 0;
            }
        }
        C.total = total;
    }
    __syncthreads();
}

/* rate_control_precinct (RateControl.c:1259-1333): the smallest acceptable Q,
 * then the largest acceptable R, by the reference's binary searches; C holds
 * the choices at (Q, R) afterwards.  Returns the data bytes, or sets C.err. */
__device__ uint32_t eo_rc(EoCtx& C, const JxsFrameDesc& F, uint32_t budget, uint32_t vp, uint32_t sg, uint32_t* hdr_out) {
    const JxsGeom& g = C.g;
    const uint32_t hdr = g.rc_hdr_bytes[C.type];
    *hdr_out = hdr;
    if (budget > 1048575u || budget <= hdr) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return 0;
    }
    const uint32_t avail = budget - hdr;
    __shared__ EoSearch S;
    __shared__ int32_t res, next;
    __shared__ uint32_t val, last_tested, last_bytes;
    /* Q: smallest acceptable */
    if (threadIdx.x == 0) {
        eo_bs_init(S, 0, g.maxq, 0, 6 > g.maxq ? 0u : 6u);
        next = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0)
            res = eo_bs_next(S, next, &val);
        __syncthreads();
        if (res != 0)
            break;
        const uint32_t q = val;
        if ((g.empty_mask[q] >> 0) & 1u) {
            if (threadIdx.x == 0)
                next = 3;
            __syncthreads();
            continue;
        }
        eo_probe(C, F, q, 0, vp, sg);
        if (threadIdx.x == 0)
            next = C.total > avail ? 1 : 2;
        __syncthreads();
    }
    if (res != 1) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return 0;
    }
    const uint32_t q = val;
    /* R: largest acceptable */
    __syncthreads();
    if (threadIdx.x == 0) {
        eo_bs_init(S, 0, g.maxr, 1, 6 >= g.maxr ? 0u : 6u);
        next = 0;
        last_tested = 0xFFFFFFFFu;
        last_bytes = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0)
            res = eo_bs_next(S, next, &val);
        __syncthreads();
        if (res != 0)
            break;
        const uint32_t r = val;
        if (threadIdx.x == 0)
            last_tested = 0xFFFFFFFFu;
        if ((g.empty_mask[q] >> r) & 1u) {
            if (threadIdx.x == 0)
                next = 3;
            __syncthreads();
            continue;
        }
        eo_probe(C, F, q, r, vp, sg);
        if (threadIdx.x == 0) {
            last_tested = r;
            last_bytes = C.total;
            next = C.total <= avail ? 1 : 2;
        }
        __syncthreads();
    }
    if (res != 1) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return 0;
    }
    const uint32_t r = val;
    if (r != last_tested)
        eo_probe(C, F, q, r, vp, sg); /* the choices at (Q, R) */
    if (threadIdx.x == 0) {
        C.q = q;
        C.r = r;
    }
    __syncthreads();
    return C.total;
}

/* The non-zero coefficients of each line at the chosen truncations: the
 * signs a sign sub-packet sends (pack_sign_c, PackPrecinct.c:200-233). */
__device__ void eo_count_signs(EoCtx& C) {
    for (uint32_t k = threadIdx.x; k < C.nl; k += EO_THREADS)
        C.nz[k] = 0;
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < C.nx; x += EO_THREADS) {
        const uint32_t li = eo_find<2>(C, x);
        const EoLine& L = C.ln[li];
        const uint32_t i = x - L.x0;
        if (eo_quant(L.cl[i], L.gl[i / 4], C.gt[L.gi], C.g.qpih))
            atomicAdd(&C.nz[li], 1u);
    }
    __syncthreads();
}

/* pack_precinct (PackPrecinct.c:235-572) into the staging words at bit
 * C.pos: total bytes with headers and zero padding; with fast sign coding
 * each packet header carries its sign sub-packet's true length and the bytes
 * saved shorten the precinct (cut) or pad it.  Returns the bytes saved (fast
 * sign coding), and advances C.pos past the precinct. */
__device__ uint32_t eo_pack(EoCtx& C, uint32_t* st, uint32_t total, bool cut) {
    const JxsGeom& g = C.g;
    const uint32_t tid = threadIdx.x;
    const uint32_t sgm = g.signs;
    if (sgm)
        eo_count_signs(C);
    __shared__ uint32_t retrieve;
    if (tid == 0) {
        /* packet layout with the sign sub-packets as sent */
        uint32_t ret = 0, n = 0, pos = g.prec_hdr_bytes;
        for (uint32_t k = 0; k < C.nl; k++) {
            n += C.nz[k];
            if (C.ln[k].last_in_pk) {
                const uint32_t p = C.ln[k].pk;
                if (sgm == 1) {
                    const uint32_t sent = (n + 7) / 8;
                    ret += C.pk_n[p] - sent;
                    C.pk_n[p] = sent;
                }
                n = 0;
            }
        }
        /* line bit offsets inside their sub-packets */
        uint32_t bd = 0, bg = 0, bs = 0, bn = 0, first = 1;
        for (uint32_t k = 0; k < C.nl; k++) {
            const uint32_t p = C.ln[k].pk;
            if (first) {
                C.pk_pos[p] = pos;
                pos += g.pkt_hdr_bytes + C.pk_s[p] + C.pk_g[p] + C.pk_d[p] + C.pk_n[p];
                bd = bg = bs = bn = 0;
                first = 0;
            }
            C.ln_bit[k][0] = bs;
            C.ln_bit[k][1] = bg;
            C.ln_bit[k][2] = bd;
            C.ln_bit[k][3] = bn;
            bs += C.bs[k];
            bg += C.pk_raw[p] ? 4 * C.ln[k].gw : C.bg[k];
            bd += C.bd[k];
            bn += C.nz[k];
            if (C.ln[k].last_in_pk)
                first = 1;
        }
        retrieve = sgm == 1 ? ret : 0u;
        /* precinct header (PackPrecinct.c:241-278) */
        const uint64_t p0 = C.pos;
        const uint32_t hlen = g.prec_hdr_bytes;
        const uint32_t lprc = total - hlen - (cut ? retrieve : 0u);
        eo_put(st, p0, lprc, 24);
        eo_put(st, p0 + 24, C.q, 8);
        eo_put(st, p0 + 32, C.r, 8);
        uint64_t b = p0 + 40;
        static const uint32_t dpb[4] = {0, 2, 1, 3}; /* D[p,b]: bit 0 vertical prediction, bit 1 significance */
        for (uint32_t gi = 0; gi < g.nb_all; gi++) {
            if (g.gb[gi] == 0xFF)
                continue;
            const JxsBand& band = g.c[g.gc[gi]].b[g.gb[gi]];
            if (band.lines[C.type] > 0)
                eo_put(st, b, dpb[C.method[gi]], 2);
            b += 2;
        }
        /* packet headers (PackHeaders.c:98-127) */
        for (uint32_t k = 0; k < C.nl; k++) {
            if (k > 0 && !C.ln[k - 1].last_in_pk)
                continue;
            const uint32_t p = C.ln[k].pk;
            uint64_t h = p0 + 8ull * C.pk_pos[p];
            eo_put(st, h, C.pk_raw[p], 1);
            if (g.short_hdr) {
                eo_put(st, h + 1, C.pk_d[p], 15);
                eo_put(st, h + 16, C.pk_g[p], 13);
                eo_put(st, h + 29, C.pk_n[p], 11);
            } else {
                eo_put(st, h + 1, C.pk_d[p], 20);
                eo_put(st, h + 21, C.pk_g[p], 20);
                eo_put(st, h + 41, C.pk_n[p], 15);
            }
        }
    }
    __syncthreads();
    const uint64_t p0 = C.pos;
    /* sub-packets, one band line at a time, its codes at block prefix sums */
    for (uint32_t k = 0; k < C.nl; k++) {
        const EoLine& L = C.ln[k];
        const uint32_t p = L.pk, gt = C.gt[L.gi], m = C.method[L.gi];
        const uint64_t pk0 = p0 + 8ull * (C.pk_pos[p] + g.pkt_hdr_bytes);
        const uint64_t sig0 = pk0, gcl0 = pk0 + 8ull * C.pk_s[p], dat0 = gcl0 + 8ull * C.pk_g[p];
        const uint64_t sgn0 = dat0 + 8ull * C.pk_d[p];
        const bool raw = C.pk_raw[p];
        const uint32_t gtop = L.l > 0 ? gt : C.top_gt[L.gi];
        const uint32_t gmax = gt > gtop ? gt : gtop;
        /* significance flags (one per group of 8) */
        if (!raw && (m == 1 || m == 3)) {
            for (uint32_t s = tid; s < L.sw; s += EO_THREADS) {
                uint32_t flag;
                if (m == 1) {
                    flag = L.sl[s] <= gt;
                } else {
                    flag = 1;
                    for (uint32_t i = 8 * s; i < min(8 * s + 8, L.gw); i++) {
                        const uint32_t gv = L.gl[i], tv = L.top[i];
                        const uint32_t mt = tv > gmax ? tv : gmax, gc = gv > gt ? gv : gt;
                        if (g.rm ? gv > gt : mt != gc)
                            flag = 0;
                    }
                }
                if (flag)
                    eo_put(st, sig0 + C.ln_bit[k][0] + s, 1, 1);
            }
        }
        /* bit-plane counts and data, GROUPS per thread per step */
        uint32_t gbase = 0, dbase = 0;
        for (uint32_t i0 = 0; i0 < L.gw; i0 += EO_THREADS) {
            const uint32_t i = i0 + tid;
            uint32_t glen = 0, dlen = 0, gcode = 0;
            uint64_t dcode = 0;
            if (i < L.gw) {
                const uint32_t gv = L.gl[i];
                if (raw) {
                    glen = 4;
                    gcode = gv;
                } else {
                    bool send = true;
                    uint32_t n;
                    if (m >= 2) {
                        const uint32_t tv = L.top[i];
                        const int32_t mt = (int32_t)(tv > gmax ? tv : gmax), gc = (int32_t)(gv > gt ? gv : gt);
                        n = eo_vlc(gc - mt, mt, (int32_t)gt);
                        if (m == 3) {
                            /* the group of 8 is skipped when flagged */
                            const uint32_t s = i / 8;
                            bool flag = true;
                            for (uint32_t j = 8 * s; j < min(8 * s + 8, L.gw); j++) {
                                const uint32_t gj = L.gl[j], tj = L.top[j];
                                const uint32_t mj = tj > gmax ? tj : gmax, cj = gj > gt ? gj : gt;
                                if (g.rm ? gj > gt : mj != cj)
                                    flag = false;
                            }
                            send = !flag;
                        }
                    } else {
                        n = gv > gt ? gv - gt : 0u;
                        if (m == 1)
                            send = L.sl[i / 8] > gt;
                    }
                    if (send) {
                        glen = n + 1;
                        gcode = ((1u << n) - 1) << 1; /* n ones, then a zero (n <= 31) */
                    }
                }
                if (gv > gt) {
                    uint32_t qv[4];
                    for (uint32_t t4 = 0; t4 < 4; t4++)
                        qv[t4] = 4 * i + t4 < L.w ? eo_quant(L.cl[4 * i + t4], gv, gt, g.qpih) : 0u;
                    if (!sgm) {
                        for (uint32_t t4 = 0; t4 < 4; t4++)
                            dcode = dcode << 1 | (qv[t4] >> 15);
                        dlen = 4;
                    }
                    for (int bit = (int)gv - 1; bit >= (int)gt; bit--) {
                        for (uint32_t t4 = 0; t4 < 4; t4++)
                            dcode = dcode << 1 | ((qv[t4] >> bit) & 1u);
                        dlen += 4;
                    }
                }
            }
            uint32_t gtot, dtot;
            const uint32_t go = eo_block_scan(glen, C.scan, &gtot);
            const uint32_t dof = eo_block_scan(dlen, C.scan, &dtot);
            if (glen && (gcode || raw))
                eo_put(st, gcl0 + C.ln_bit[k][1] + gbase + go, gcode, glen);
            if (dlen)
                eo_put(st, dat0 + C.ln_bit[k][2] + dbase + dof, dcode, dlen);
            gbase += gtot;
            dbase += dtot;
        }
        /* signs of the coefficients that stay non-zero */
        if (sgm) {
            uint32_t nbase = 0;
            for (uint32_t x0 = 0; x0 < L.w; x0 += EO_THREADS) {
                const uint32_t x = x0 + tid;
                uint32_t qv = 0;
                if (x < L.w)
                    qv = eo_quant(L.cl[x], L.gl[x / 4], gt, g.qpih);
                uint32_t tot;
                const uint32_t o = eo_block_scan(qv ? 1u : 0u, C.scan, &tot);
                if (qv && (qv >> 15))
                    eo_put(st, sgn0 + C.ln_bit[k][3] + nbase + o, 1, 1);
                nbase += tot;
            }
        }
    }
    __syncthreads();
    const uint32_t ret = retrieve;
    if (tid == 0) {
        /* the precinct with its zero padding (with fast sign coding also the
         * saved bytes, unless the precinct is cut) */
        C.pos = p0 + 8ull * ((uint64_t)total - (cut ? ret : 0u));
    }
    __syncthreads();
    return ret;
}

/* Precinct budgets of a slice for rate-control modes 2 and 3
 * (PackStageProcess.c:340-481): one Q and R for the slice by the binary
 * searches on the summed bytes, without vertical prediction and with fast
 * sign coding in place of full (RateControl.c:1086-1210, 1336-1425). */
__device__ void eo_slice_budgets(EoCtx& C, const JxsFrameDesc& F, uint32_t first_row, uint32_t nprec, uint32_t sbudget) {
    const JxsGeom& g = C.g;
    const uint32_t sg = g.signs ? 1u : 0u;
    __shared__ EoSearch S;
    __shared__ int32_t res, next;
    __shared__ uint32_t val;
    __shared__ uint64_t sum;
    uint32_t hsum = 0;
    for (uint32_t i = 0; i < nprec; i++) {
        const uint32_t type = first_row + i + 1 >= g.rows ? 1u : 0u;
        hsum += g.rc_hdr_bytes[type];
    }
    if ((sbudget + nprec - 1) / nprec > 1048575u || sbudget <= hsum || nprec > EO_MAXP) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return;
    }
    const uint64_t avail = sbudget - hsum;
    auto sum_at = [&](uint32_t q, uint32_t r) {
        if (threadIdx.x == 0)
            sum = 0;
        __syncthreads();
        for (uint32_t i = 0; i < nprec; i++) {
            eo_lines(C, F, first_row + i, 0);
            eo_probe(C, F, q, r, 0, sg);
            if (threadIdx.x == 0) {
                sum += C.total;
                C.sbud[i] = C.total;
            }
            __syncthreads();
        }
    };
    uint32_t q = 0, r = 0;
    if (threadIdx.x == 0) {
        eo_bs_init(S, 0, g.maxq, 0, 6 > g.maxq ? 0u : 6u);
        next = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0)
            res = eo_bs_next(S, next, &val);
        __syncthreads();
        if (res != 0)
            break;
        q = val;
        if (g.empty_mask[q] & 1u) {
            if (threadIdx.x == 0)
                next = 3;
            __syncthreads();
            continue;
        }
        sum_at(q, 0);
        if (threadIdx.x == 0)
            next = sum > avail ? 1 : 2;
        __syncthreads();
    }
    if (res != 1) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return;
    }
    q = val;
    __syncthreads();
    if (threadIdx.x == 0) {
        eo_bs_init(S, 0, g.maxr, 1, 6 >= g.maxr ? 0u : 6u);
        next = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0)
            res = eo_bs_next(S, next, &val);
        __syncthreads();
        if (res != 0)
            break;
        r = val;
        if ((g.empty_mask[q] >> r) & 1u) {
            if (threadIdx.x == 0)
                next = 3;
            __syncthreads();
            continue;
        }
        sum_at(q, r);
        if (threadIdx.x == 0)
            next = sum <= avail ? 1 : 2;
        __syncthreads();
    }
    if (res != 1) {
        if (threadIdx.x == 0)
            C.err = 1;
        __syncthreads();
        return;
    }
    r = val;
    sum_at(q, r); /* each precinct's data bytes at (Q, R) in C.sbud */
    if (threadIdx.x == 0) {
        uint32_t bud[EO_MAXP], pad[EO_MAXP];
        uint32_t left = (uint32_t)avail;
        for (uint32_t i = 0; i < nprec; i++) {
            const uint32_t type = first_row + i + 1 >= g.rows ? 1u : 0u;
            left -= C.sbud[i];
            bud[i] = g.rc_hdr_bytes[type] + C.sbud[i];
            pad[i] = 0;
        }
        const uint32_t last = nprec - 1;
        pad[last] = left;
        bud[last] += left;
        if (g.rc_mode == 3) {
            uint32_t mx = 0, get = 0;
            for (uint32_t i = 0; i < nprec; i++)
                mx = max(mx, bud[i] - pad[i]);
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
            uint32_t add_first = bud[0] * (g.signs == 1 ? 20u : 10u) / 100;
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
        for (uint32_t i = 0; i < nprec; i++)
            C.sbud[i] = bud[i];
    }
    __syncthreads();
}

/* The staging words of slice s: after the frame's codestream-sized shadow
 * area, slice s at word ((slice_off + 3) / 4 + s), so slices never share a
 * word. */
__device__ __forceinline__ uint32_t* eo_stage(const JxsFrameDesc& F, uint32_t s, const uint32_t* slice_off) {
    return F.shadow + ((slice_off[s] + 3) >> 2) + s;
}

__global__ void __launch_bounds__(EO_THREADS) k_enc_slice(const JxsFrameDesc* __restrict__ frames, int nfr,
                                                          uint32_t total_slices) {
    __shared__ EoCtx C;
    const uint32_t gs = blockIdx.x;
    if (gs >= total_slices)
        return;
    int fl = 0, fh = nfr - 1;
    while (fl < fh) {
        const int mid = (fl + fh + 1) >> 1;
        if (frames[mid].slice_base <= gs)
            fl = mid;
        else
            fh = mid - 1;
    }
    const JxsFrameDesc& F = frames[fl];
    if (!F.ordered)
        return;
    const uint32_t tid = threadIdx.x;
    {
        const uint4* src = (const uint4*)F.g;
        uint4* dst = (uint4*)&C.g;
        for (uint32_t k = tid; k < sizeof(JxsGeom) / 16; k += EO_THREADS)
            dst[k] = src[k];
    }
    if (tid == 0) {
        C.err = 0;
        C.pos = 0;
    }
    __syncthreads();
    const JxsGeom& g = C.g;
    const uint32_t s = gs - F.slice_base;
    const uint32_t first_row = s * g.pps;
    const uint32_t nprec = (s == g.slices - 1) ? g.rows - first_row : g.pps;
    const uint32_t soff = F.slice_off[s];
    const uint32_t ssize = (s + 1 < g.slices ? F.slice_off[s + 1] : g.lcod) - soff;
    const uint32_t sbudget = F.slice_bud[s];
    uint32_t* st = eo_stage(F, s, F.slice_off);
    const uint32_t nw = (ssize + 3) / 4;
    for (uint32_t k = tid; k < nw; k += EO_THREADS)
        st[k] = 0;
    __syncthreads();
    /* slice header (PackHeaders.c:129-138) */
    if (tid == 0) {
        eo_put(st, 0, 0xFF20, 16);
        eo_put(st, 16, 4, 16);
        eo_put(st, 32, s, 16);
        C.pos = 48;
    }
    __syncthreads();
    /* budgets (PackStageProcess.c:596-652) */
    uint32_t minb = sbudget / nprec, first = minb, left = sbudget - minb * nprec;
    if (g.rc_mode == 1) {
        first = minb * (g.signs == 1 ? 120u : 110u) / 100;
        if (first > sbudget)
            first = sbudget;
        if (nprec > 1) {
            minb = (sbudget - first) / (nprec - 1);
            left = sbudget - minb * (nprec - 1) - first;
        } else {
            left = 0;
        }
    }
    if (g.rc_mode >= 2)
        eo_slice_budgets(C, F, first_row, nprec, sbudget);
    const bool full_dz = g.signs == 2 && g.qpih == 0;
    uint32_t carry = 0;
    for (uint32_t i = 0; i < nprec && !C.err; i++) {
        const uint32_t row = first_row + i;
        const uint32_t budget = (g.rc_mode >= 2 ? C.sbud[i] : (i == 0 ? first : minb) + (i < left ? 1u : 0u)) + carry;
        /* the precinct above, for vertical prediction (PackStageProcess.c:655-660) */
        if (tid == 0 && i > 0)
            for (uint32_t gi = 0; gi < JXS_NGB; gi++)
                C.top_gt[gi] = C.gt[gi];
        __syncthreads();
        eo_lines(C, F, row, g.vpred && i > 0);
        if (full_dz)
            eo_coeff_hist(C);
        uint32_t hdr = 0;
        const uint32_t data = eo_rc(C, F, budget, g.vpred, g.signs, &hdr);
        if (C.err)
            break;
        uint32_t pad = budget - hdr - data, total = budget;
        bool cut = false;
        if (g.rc_mode != 0 && i + 1 < nprec) {
            total -= pad;
            carry = pad;
            pad = 0;
            cut = g.signs == 1;
        }
        if (tid == 0) {
            JxsPrec& P = F.prec[row];
            P.q = C.q;
            P.r = C.r;
            P.total = total;
            P.pad = pad;
            P.offset = soff + (uint32_t)(C.pos >> 3);
            P.err = 0;
        }
        const uint32_t saved = eo_pack(C, st, total, cut);
        if (cut)
            carry += saved;
    }
    __syncthreads();
    if (C.err) {
        if (tid == 0)
            atomicOr(F.status, 1u);
        return;
    }
    if (s == g.slices - 1 && tid == 0)
        eo_put(st, C.pos, 0xFF11, 16); /* EOC */
    __syncthreads();
    /* the slice's bytes into the codestream (the staging words read where the
     * atomics left them); the first slice also writes the frame header */
    for (uint32_t k = tid; k < ssize; k += EO_THREADS) {
        const uint32_t wv = __hip_atomic_load(st + (k >> 2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        F.out[soff + k] = (uint8_t)(wv >> (24 - 8 * (k & 3)));
    }
    if (s == 0)
        for (uint32_t k = tid; k < g.hdr_len; k += EO_THREADS)
            F.out[k] = g.hdr[k];
}

extern "C" hipError_t jxs_launch_enc_slice(const JxsFrameDesc* frames, int nfr, uint32_t total_slices, hipStream_t st) {
    if (!total_slices)
        return hipSuccess;
    hipLaunchKernelGGL(k_enc_slice, dim3(total_slices), dim3(EO_THREADS), 0, st, frames, nfr, total_slices);
    return hipGetLastError();
}
