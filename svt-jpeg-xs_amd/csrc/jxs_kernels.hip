/*
 * jxs_kernels.hip -- gfx950 kernels of the JPEG XS encode path.
 *
 *   k_vert     vertical 5/3 lifting of one decomposition level; the first
 *              level reads the input planes and applies the input scaling
 *              (NltEnc.c:33-71) on the fly.  One lane per column, a segment of
 *              rows per block, coalesced row loads.
 *   k_horz     one block per row: the row is staged in LDS and all horizontal
 *              levels of that row are lifted in place; every band produced is
 *              written as 16-bit sign-magnitude coefficients (image_shift,
 *              NltEnc.c:12-31) together with its GCLI per 4 coefficients and
 *              the maximum per 8 GCLIs (GcStageProcess.c:22-97), using wave
 *              shuffles for the maxima.
 *   k_rc_table one block per precinct: GCLI / significance histograms per band
 *              line (RateControl.c:18-154), then the packet-level byte count
 *              of the precinct for every (Q, R) (RateControl.c:457-781).
 *   k_select   one lane per slice: precinct budgets with the +10 % first
 *              precinct and the padding carry (PackStageProcess.c:600-684),
 *              the reference's binary searches replayed on the table
 *              (RateControl.c:889-1083, BinarySearch.c), precinct offsets.
 *   k_pack     one block per (packet, precinct): precinct / packet headers,
 *              significance, bit-plane-count and data sub-packets
 *              (PackPrecinct.c:235-572) with quantisation (Quant.c:9-45)
 *              fused; bit offsets from block prefix sums, bits assembled in an
 *              LDS ring with atomicOr and flushed as bytes.
 *
 * All arithmetic is integer.  Every kernel takes a frame-descriptor array so
 * one launch covers a batch of heterogeneous frames.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jxs_geom.h"

#define WAVE 64

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int32_t load_src(const void* src, uint32_t kind, size_t idx, uint32_t shift, uint32_t mask,
                                            int32_t offset) {
    if (kind == JXS_SRC_I32)
        return ((const int32_t*)src)[idx];
    uint32_t v = kind == JXS_SRC_U8 ? (uint32_t)((const uint8_t*)src)[idx] : (uint32_t)((const uint16_t*)src)[idx];
    return (int32_t)((v & mask) << shift) - offset;
}

__device__ __forceinline__ uint16_t to_sign_mag(int32_t v) {
    if (v >= 0)
        return (uint16_t)((v + 128) >> 8);
    const uint32_t m = (uint32_t)((-v + 128) >> 8);
    return (uint16_t)(m ? (m | 0x8000u) : 0u);
}

__device__ __forceinline__ uint32_t bitlen(uint32_t v) { return v ? 32u - __clz(v) : 0u; }

__device__ __forceinline__ uint32_t truncation(uint32_t gain, uint32_t prio, uint32_t q, uint32_t r) {
    const uint32_t pump = prio < r ? 1u : 0u;
    if (q < gain + pump)
        return 0;
    const uint32_t t = q - gain - pump;
    return t > 15u ? 15u : t;
}

/* Exclusive block scan of v over blockDim.x (multiple of 64) threads. */
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= o)
            x += y;
    }
    if (lane == WAVE - 1)
        tmp[wid] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t w = 0; w < nw; w++) {
        const uint32_t t = tmp[w];
        before += w < wid ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

/* ------------------------------------------------------------------ */
/* forward transform                                                   */
/* ------------------------------------------------------------------ */
#define VSEG 16 /* output row pairs per block row */

__device__ __forceinline__ int find_job_u32(const uint32_t* bases, int n, uint32_t blk) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bases[mid] <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_vert(const JxsVertJob* __restrict__ jobs, const uint32_t* __restrict__ bases,
                                              int njobs) {
    const int j = find_job_u32(bases, njobs, blockIdx.x);
    const JxsVertJob J = jobs[j];
    const uint32_t t = blockIdx.x - J.base;
    const uint32_t tx = t % J.tiles_x, ty = t / J.tiles_x;
    const uint32_t x = tx * blockDim.x + threadIdx.x;
    if (x >= J.w)
        return;
    const uint32_t n = J.h, nh = n / 2, nl = n - nh;
    const uint32_t i0 = ty * VSEG, i1 = min(i0 + VSEG, nl);
    const size_t col = (size_t)x * J.step;
    auto X = [&](uint32_t row) { return load_src(J.src, J.kind, (size_t)row * J.stride + col, J.shift, J.mask, J.offset); };
    int32_t* L = J.dst;
    int32_t* H = J.dst + (size_t)nl * J.w;
    /* hf(k) = x(2k+1) - ((x(2k) + x(2k+2)) >> 1), or x(2k+1) - x(2k) at an even end */
    int32_t hprev = 0;
    int32_t xe = X(2 * i0); /* x(2i) */
    if (i0 > 0) {
        const int32_t xp = X(2 * i0 - 2), xo = X(2 * i0 - 1);
        hprev = xo - ((xp + xe) >> 1);
    }
    for (uint32_t i = i0; i < i1; i++) {
        int32_t hcur = 0, xn = 0;
        if (i < nh) {
            const int32_t xo = X(2 * i + 1);
            if (2 * i + 2 < n) {
                xn = X(2 * i + 2);
                hcur = xo - ((xe + xn) >> 1);
            } else {
                hcur = xo - xe;
            }
        }
        int32_t lf;
        if (i == 0)
            lf = xe + ((hcur + 1) >> 1);
        else if (i < nh)
            lf = xe + ((hprev + hcur + 2) >> 2);
        else
            lf = xe + ((hprev + 1) >> 1);
        L[(size_t)i * J.w + x] = lf;
        if (i < nh)
            H[(size_t)i * J.w + x] = hcur;
        hprev = hcur;
        xe = xn;
    }
}

#define HMAX JXS_MAX_ROW

/* Writes `count` values buf[start + k * step] as one band row with GCLI and
 * significance maxima.  Called by all threads of the block. */
__device__ void emit_band_row(const JxsBandOut& o, uint32_t y, const int32_t* buf, uint32_t start, uint32_t step,
                              uint32_t count) {
    uint16_t* crow = o.coeff + (size_t)y * o.w;
    uint8_t* grow = o.gcli + (size_t)y * o.gw;
    uint8_t* srow = o.smax + (size_t)y * o.sw;
    for (uint32_t base = 0; base < o.gw; base += blockDim.x) {
        const uint32_t gidx = base + threadIdx.x;
        uint32_t orv = 0;
        if (gidx < o.gw) {
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t i = 4 * gidx + k;
                if (i < count) {
                    const uint16_t v = to_sign_mag(buf[start + i * step]);
                    crow[i] = v;
                    orv |= v & 0x7FFFu;
                }
            }
            grow[gidx] = (uint8_t)bitlen(orv);
        }
        uint32_t m = bitlen(orv);
        m = max(m, (uint32_t)__shfl_xor((int)m, 1, WAVE));
        m = max(m, (uint32_t)__shfl_xor((int)m, 2, WAVE));
        m = max(m, (uint32_t)__shfl_xor((int)m, 4, WAVE));
        if ((gidx & 7) == 0 && gidx / 8 < o.sw)
            srow[gidx / 8] = (uint8_t)m;
    }
}

__global__ void __launch_bounds__(256) k_horz(const JxsHorzJob* __restrict__ jobs, const uint32_t* __restrict__ bases,
                                              int njobs) {
    __shared__ int32_t buf[HMAX];
    const int j = find_job_u32(bases, njobs, blockIdx.x);
    const JxsHorzJob& J = jobs[j];
    const uint32_t y = blockIdx.x - J.base;
    const uint32_t w = J.w;
    const size_t rbase = (size_t)y * J.stride;
    for (uint32_t i = threadIdx.x; i < w; i += blockDim.x)
        buf[i] = load_src(J.src, J.kind, (rbase + i) * (size_t)1 + (size_t)i * (J.step - 1), J.shift, J.mask, J.offset);
    __syncthreads();
    uint32_t n = w, s = 1;
    for (uint32_t lvl = 0; lvl < J.levels; lvl++) {
        const uint32_t nh = n / 2, nl = n - nh;
        for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x) {
            const int32_t a = buf[2 * i * s];
            if (2 * i + 2 < n)
                buf[(2 * i + 1) * s] -= (a + buf[(2 * i + 2) * s]) >> 1;
            else
                buf[(2 * i + 1) * s] -= a;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
            if (i == 0)
                buf[0] += (buf[s] + 1) >> 1;
            else if (i < nh)
                buf[2 * i * s] += (buf[(2 * i - 1) * s] + buf[(2 * i + 1) * s] + 2) >> 2;
            else
                buf[2 * i * s] += (buf[(2 * i - 1) * s] + 1) >> 1;
        }
        __syncthreads();
        emit_band_row(J.hf[lvl], y, buf, s, 2 * s, nh);
        s *= 2;
        n = nl;
    }
    if (J.lf_is_band) {
        emit_band_row(J.lf, y, buf, 0, s, n);
    } else {
        int32_t* dst = J.lf_i32 + (size_t)y * J.lf_stride;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            dst[i] = buf[i * s];
    }
}

/* ------------------------------------------------------------------ */
/* rate control                                                        */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int find_frame(const JxsFrameDesc* fr, int nfr, uint32_t blk) {
    int lo = 0, hi = nfr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (fr[mid].prec_base <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

/* Packet-level bytes of one precinct for (q, r), reading band-line tables
 * from lut (n_bl x JXS_LUT).  Mirrors precinct_get_budget_bytes with the
 * per-band method choice of rate_control_calculate_band_best_method. */
__device__ uint32_t precinct_bytes(const JxsGeom& g, uint32_t type, const uint32_t* lut, uint32_t q, uint32_t r) {
    uint32_t total = 0;
    for (uint32_t p = 0; p < g.npk; p++) {
        uint32_t d = 0, gb = 0, sb = 0;
        const uint32_t l = g.pk_line[p];
        for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            const uint32_t nl = bd.lines[type];
            if (l >= nl)
                continue;
            const uint32_t gt = truncation(bd.gain, bd.prio, q, r);
            uint32_t nosig_sum = 0, sig_sum = bd.sw * nl, mine_nosig = 0, mine_sig = 0;
            for (uint32_t k = 0; k < nl; k++) {
                const uint32_t* L = lut + (size_t)(bd.bl + k) * JXS_LUT;
                const uint32_t ns = L[LUT_DATA + gt] + L[LUT_CNT + gt];
                uint32_t zeroed = L[LUT_SIG + gt] * 8;
                const int32_t tail = (int32_t)L[LUT_TAIL];
                if (tail >= 0 && (uint32_t)tail <= gt)
                    zeroed += bd.gw - (bd.gw / 8) * 8;
                nosig_sum += ns;
                sig_sum += ns - zeroed;
                if (k == l) {
                    mine_nosig = ns;
                    mine_sig = ns - zeroed;
                    d += 4 * L[LUT_DATA + gt];
                }
            }
            if (g.sig && nosig_sum > sig_sum) {
                gb += mine_sig;
                sb += bd.sw;
            } else {
                gb += mine_nosig;
            }
        }
        const uint32_t db = (d + 7) >> 3, gbb = (gb + 7) >> 3, sbb = (sb + 7) >> 3;
        total += db;
        total += (sbb + gbb > g.raw_bytes[type][p]) ? g.raw_bytes[type][p] : sbb + gbb;
    }
    return total;
}

__global__ void __launch_bounds__(256) k_rc_table(const JxsFrameDesc* __restrict__ frames, int nfr) {
    __shared__ uint32_t hist[JXS_NBL][16];
    __shared__ uint32_t shist[JXS_NBL][16];
    __shared__ int32_t tailv[JXS_NBL];
    __shared__ uint32_t lut[JXS_NBL * JXS_LUT];
    const int f = find_frame(frames, nfr, blockIdx.x);
    const JxsFrameDesc& F = frames[f];
    const JxsGeom& g = *F.g;
    const uint32_t row = blockIdx.x - F.prec_base;
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    for (uint32_t i = threadIdx.x; i < JXS_NBL * 16; i += blockDim.x) {
        (&hist[0][0])[i] = 0;
        (&shist[0][0])[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < JXS_NBL; i += blockDim.x)
        tailv[i] = -1;
    __syncthreads();
    for (uint32_t c = 0; c < g.nc; c++)
        for (uint32_t b = 0; b < g.c[c].nb; b++) {
            const JxsBand& bd = g.c[c].b[b];
            for (uint32_t l = 0; l < bd.lines[type]; l++) {
                const uint32_t y = row * bd.hl + l, bl = bd.bl + l;
                const uint8_t* gl = F.gcli + bd.gcli_off + (size_t)y * bd.gw;
                const uint8_t* sl = F.smax + bd.smax_off + (size_t)y * bd.sw;
                for (uint32_t i = threadIdx.x; i < bd.gw; i += blockDim.x)
                    atomicAdd(&hist[bl][gl[i] & 15], 1u);
                const uint32_t full = bd.gw / 8;
                for (uint32_t i = threadIdx.x; i < full; i += blockDim.x)
                    atomicAdd(&shist[bl][sl[i] & 15], 1u);
                if (threadIdx.x == 0 && full < bd.sw)
                    tailv[bl] = sl[full];
            }
        }
    __syncthreads();
    /* cumulative tables per band line */
    uint32_t* glut = F.lut + (size_t)row * g.n_bl * JXS_LUT;
    for (uint32_t bl = threadIdx.x; bl < g.n_bl; bl += blockDim.x) {
        uint32_t* L = lut + bl * JXS_LUT;
        uint32_t acc = 0, sacc = 0;
        for (int t = 0; t < 16; t++) {
            acc += hist[bl][t];
            sacc += shist[bl][t];
            L[LUT_CNT + t] = acc;
            L[LUT_SIG + t] = sacc;
        }
        /* data[t] = sum_{i > t} (i - t + 1) * hist[i], built from the top */
        uint32_t above = 0, d = 0;
        L[LUT_DATA + 15] = 0;
        for (int t = 14; t >= 0; t--) {
            above += hist[bl][t + 1];
            d += above + hist[bl][t + 1];
            L[LUT_DATA + t] = d;
        }
        L[LUT_TAIL] = (uint32_t)tailv[bl];
        for (int k = 0; k < JXS_LUT; k++)
            glut[bl * JXS_LUT + k] = L[k];
    }
    __syncthreads();
    uint32_t* tab = F.table + (size_t)row * JXS_NQ * JXS_NR;
    const uint32_t nq = g.maxq + 1, nr = g.maxr + 1;
    for (uint32_t idx = threadIdx.x; idx < nq * nr; idx += blockDim.x) {
        const uint32_t q = idx / nr, r = idx % nr;
        tab[q * JXS_NR + r] = precinct_bytes(g, type, lut, q, r);
    }
}

/* BinarySearch.c:9-78 */
struct BSearch {
    int32_t beg, end, step, best, last;
    int below;
    __device__ void init(int32_t b, int32_t e, int bl, int32_t st) {
        beg = b;
        end = e;
        below = bl;
        best = -1;
        last = -1;
        step = st ? st : (end - beg) / 2;
    }
    /* res: 0 begin, 1 too small, 2 too big, 3 out of range.
     * returns 0 continue, 1 found, 2 error */
    __device__ int next(int res, uint32_t* out) {
        if (res == 1) {
            if (below && (best == -1 || best < last))
                best = last;
            if (end >= last + 1)
                beg = last + 1;
            else if (best == -1)
                return 2;
            else {
                *out = (uint32_t)best;
                return 1;
            }
        } else if (res == 2 || res == 3) {
            if (res == 2 && !below && (best == -1 || best > last))
                best = last;
            if (beg <= last - 1)
                end = last - 1;
            else if (best == -1)
                return 2;
            else {
                *out = (uint32_t)best;
                return 1;
            }
        }
        if (step > end - beg)
            step = (end - beg + 1) / 2;
        last = step + beg;
        *out = (uint32_t)last;
        return 0;
    }
};

__global__ void __launch_bounds__(64) k_select(const JxsFrameDesc* __restrict__ frames, int nfr, uint32_t total_slices) {
    const uint32_t gs = blockIdx.x * blockDim.x + threadIdx.x;
    if (gs >= total_slices)
        return;
    int lo = 0, hi = nfr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frames[mid].slice_base <= gs)
            lo = mid;
        else
            hi = mid - 1;
    }
    const JxsFrameDesc& F = frames[lo];
    const JxsGeom& g = *F.g;
    const uint32_t s = gs - F.slice_base;
    const uint32_t first_row = s * g.pps;
    const uint32_t nprec = (s == g.slices - 1) ? g.rows - first_row : g.pps;
    const uint32_t sbud = F.slice_bud[s];
    uint32_t minb = sbud / nprec;
    uint32_t first = minb * 110 / 100;
    if (first > sbud)
        first = sbud;
    uint32_t left = 0;
    if (nprec > 1) {
        minb = (sbud - first) / (nprec - 1);
        left = sbud - minb * (nprec - 1) - first;
    }
    uint32_t carry = 0, off = F.slice_off[s] + 6;
    bool failed = false;
    for (uint32_t i = 0; i < nprec; i++) {
        const uint32_t row = first_row + i, type = row + 1 >= g.rows ? 1 : 0;
        JxsPrec P;
        P.q = P.r = P.total = P.pad = 0;
        P.offset = off;
        P.err = 1;
        if (failed) {
            F.prec[row] = P;
            continue;
        }
        const uint32_t budget = (i == 0 ? first : minb) + (i < left ? 1 : 0) + carry;
        const uint32_t hdr = g.rc_hdr_bytes[type];
        const uint32_t* tab = F.table + (size_t)row * JXS_NQ * JXS_NR;
        bool ok = budget <= 1048575u && budget > hdr;
        uint32_t q = 0, r = 0, data = 0;
        if (ok) {
            const uint32_t avail = budget - hdr;
            BSearch bs;
            bs.init(0, (int32_t)g.maxq, 0, 6 > g.maxq ? 0 : 6);
            int res, nx = 0;
            while ((res = bs.next(nx, &q)) == 0) {
                if ((g.empty_mask[q] & 1u) != 0) {
                    nx = 3;
                    continue;
                }
                nx = tab[q * JXS_NR] > avail ? 1 : 2;
            }
            ok = res == 1;
            if (ok) {
                bs.init(0, (int32_t)g.maxr, 1, 6 >= g.maxr ? 0 : 6);
                nx = 0;
                while ((res = bs.next(nx, &r)) == 0) {
                    if ((g.empty_mask[q] >> r) & 1u) {
                        nx = 3;
                        continue;
                    }
                    nx = tab[q * JXS_NR + r] <= avail ? 1 : 2;
                }
                ok = res == 1;
                if (ok) {
                    data = tab[q * JXS_NR + r];
                    uint32_t pad = avail - data, total = budget;
                    if (i + 1 < nprec) {
                        total -= pad;
                        carry = pad;
                        pad = 0;
                    }
                    P.q = q;
                    P.r = r;
                    P.total = total;
                    P.pad = pad;
                    P.err = 0;
                    off += total;
                }
            }
        }
        if (!ok) {
            failed = true;
            atomicOr(F.status, 1u);
        }
        F.prec[row] = P;
    }
}

/* ------------------------------------------------------------------ */
/* packing                                                             */
/* ------------------------------------------------------------------ */
#define RING 2048 /* 32-bit words of the LDS bit ring (8 KiB) */

__device__ __forceinline__ void ring_put(uint32_t* ring, uint64_t pos, uint64_t v, uint32_t n) {
    while (n) {
        const uint32_t sh = (uint32_t)(pos & 31), room = 32 - sh, k = n < room ? n : room;
        const uint32_t part = (uint32_t)((v >> (n - k)) & ((k == 32) ? 0xFFFFFFFFull : ((1ull << k) - 1)));
        atomicOr(&ring[(pos >> 5) % RING], part << (room - k));
        pos += k;
        n -= k;
    }
}

/* Writes words [from, to) of the ring as bytes of the sub-packet (only bytes
 * below nbytes) and clears them for reuse. */
__device__ void ring_flush(uint32_t* ring, uint8_t* out, uint32_t nbytes, uint32_t from, uint32_t to) {
    for (uint32_t w = from + threadIdx.x; w < to; w += blockDim.x) {
        const uint32_t v = ring[w % RING];
        ring[w % RING] = 0;
        const uint32_t b0 = 4 * w;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (b0 + k < nbytes)
                out[b0 + k] = (uint8_t)(v >> (24 - 8 * k));
    }
    __syncthreads();
}

struct PackCtx {
    uint8_t gtli[JXS_MAXC][JXS_NBC];
    uint8_t msig[JXS_MAXC][JXS_NBC];
    uint32_t bits_d[JXS_NBL], bits_g[JXS_NBL], bits_s[JXS_NBL];
    uint32_t pk_d[JXS_NPK], pk_g[JXS_NPK], pk_s[JXS_NPK], pk_raw[JXS_NPK], pk_off[JXS_NPK], pk_any[JXS_NPK];
    uint32_t scan[8];
};

/* kinds of item streams */
enum { ITEM_RAW = 0, ITEM_SIG = 1, ITEM_VLC = 2, ITEM_DATA = 3 };

/* Emits the items of one band line into the ring at bit `pos0` (relative to
 * the sub-packet start); returns the number of bits written. */
__device__ uint32_t emit_line(uint32_t* ring, PackCtx& C, uint8_t* out, uint32_t nbytes, uint32_t& flushed, uint32_t pos0,
                              int kind, const uint8_t* gl, const uint8_t* sl, const uint16_t* cl, uint32_t w, uint32_t gw,
                              uint32_t sw, uint32_t gt, uint32_t use_sig, uint32_t qpih) {
    const uint32_t nitems = kind == ITEM_SIG ? sw : gw;
    uint32_t pos = pos0;
    for (uint32_t base = 0; base < nitems; base += blockDim.x) {
        /* everything below pos is final: flush the completed words */
        const uint32_t upto = pos >> 5;
        if (upto > flushed) {
            ring_flush(ring, out, nbytes, flushed, upto);
            flushed = upto;
        }
        const uint32_t i = base + threadIdx.x;
        uint32_t nb = 0;
        uint64_t v = 0;
        if (i < nitems) {
            if (kind == ITEM_RAW) {
                nb = 4;
                v = gl[i];
            } else if (kind == ITEM_SIG) {
                nb = 1;
                v = sl[i] <= gt ? 1u : 0u;
            } else if (kind == ITEM_VLC) {
                if (!(use_sig && sl[i / 8] <= gt)) {
                    const uint32_t gc = gl[i], m = gc > gt ? gc - gt : 0;
                    nb = m + 1;
                    v = ((1ull << m) - 1) << 1;
                }
            } else {
                const uint32_t gc = gl[i];
                if (gc > gt) {
                    uint32_t qv[4];
#pragma unroll
                    for (uint32_t k = 0; k < 4; k++) {
                        const uint32_t ci = 4 * i + k;
                        uint32_t val = 0;
                        if (ci < w) {
                            const uint32_t c = cl[ci], sign = c & 0x8000u, m = c & 0x7FFFu;
                            uint32_t o;
                            if (gt == 0) {
                                o = m;
                            } else if (qpih == 0) {
                                o = (m >> gt) << gt;
                            } else {
                                const uint32_t sc = gc - gt + 1;
                                const uint32_t dq = (((m << sc) - m + (1u << gc)) >> (gc + 1)) & 0xFFFFu;
                                o = (dq << gt) & 0xFFFFu;
                            }
                            val = o ? (o | sign) : 0u;
                        }
                        qv[k] = val;
                    }
                    v = ((qv[0] >> 15) << 3) | ((qv[1] >> 15) << 2) | ((qv[2] >> 15) << 1) | (qv[3] >> 15);
                    for (int bit = (int)gc - 1; bit >= (int)gt; bit--) {
                        const uint32_t nib = (((qv[0] >> bit) & 1u) << 3) | (((qv[1] >> bit) & 1u) << 2) |
                            (((qv[2] >> bit) & 1u) << 1) | ((qv[3] >> bit) & 1u);
                        v = (v << 4) | nib;
                    }
                    nb = 4 * (gc - gt + 1);
                }
            }
        }
        uint32_t tot;
        const uint32_t off = block_scan(nb, C.scan, &tot);
        if (nb)
            ring_put(ring, (uint64_t)pos + off, v, nb);
        __syncthreads();
        pos += tot;
    }
    return pos - pos0;
}

__global__ void __launch_bounds__(256) k_pack(const JxsFrameDesc* __restrict__ frames, int nfr) {
    __shared__ uint32_t ring[RING];
    __shared__ PackCtx C;
    const int f = find_frame(frames, nfr, blockIdx.y);
    const JxsFrameDesc& F = frames[f];
    const JxsGeom& g = *F.g;
    const uint32_t row = blockIdx.y - F.prec_base;
    const uint32_t p = blockIdx.x;
    if (p >= g.npk)
        return;
    const JxsPrec P = F.prec[row];
    if (P.err)
        return;
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    const uint32_t* lut = F.lut + (size_t)row * g.n_bl * JXS_LUT;
    /* per band: truncation and coding method for (Q, R) */
    for (uint32_t gi = threadIdx.x; gi < g.nb_all; gi += blockDim.x) {
        if (g.gb[gi] == 0xFF)
            continue;
        const uint32_t c = g.gc[gi], b = g.gb[gi];
        const JxsBand& bd = g.c[c].b[b];
        const uint32_t gt = truncation(bd.gain, bd.prio, P.q, P.r);
        C.gtli[c][b] = (uint8_t)gt;
        const uint32_t nl = bd.lines[type];
        uint32_t nosig = 0, sig = bd.sw * nl;
        for (uint32_t k = 0; k < nl; k++) {
            const uint32_t* L = lut + (size_t)(bd.bl + k) * JXS_LUT;
            const uint32_t ns = L[LUT_DATA + gt] + L[LUT_CNT + gt];
            uint32_t zeroed = L[LUT_SIG + gt] * 8;
            const int32_t tail = (int32_t)L[LUT_TAIL];
            if (tail >= 0 && (uint32_t)tail <= gt)
                zeroed += bd.gw - (bd.gw / 8) * 8;
            nosig += ns;
            sig += ns - zeroed;
            C.bits_d[bd.bl + k] = 4 * L[LUT_DATA + gt];
            C.bits_g[bd.bl + k] = ns;
            C.bits_s[bd.bl + k] = zeroed; /* temporarily: bits removed by significance */
        }
        const uint32_t use = (g.sig && nl > 0 && nosig > sig) ? 1u : 0u;
        C.msig[c][b] = (uint8_t)use;
        for (uint32_t k = 0; k < nl; k++) {
            if (use) {
                C.bits_g[bd.bl + k] -= C.bits_s[bd.bl + k];
                C.bits_s[bd.bl + k] = bd.sw;
            } else {
                C.bits_s[bd.bl + k] = 0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t off = P.offset + g.prec_hdr_bytes;
        for (uint32_t pp = 0; pp < g.npk; pp++) {
            uint32_t d = 0, gb = 0, sb = 0, any = 0;
            const uint32_t l = g.pk_line[pp];
            for (uint32_t gi = g.pk_bs[pp]; gi < g.pk_be[pp]; gi++) {
                const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
                if (l >= bd.lines[type])
                    continue;
                any = 1;
                d += C.bits_d[bd.bl + l];
                gb += C.bits_g[bd.bl + l];
                sb += C.bits_s[bd.bl + l];
            }
            const uint32_t db = (d + 7) >> 3, gbb = (gb + 7) >> 3, sbb = (sb + 7) >> 3;
            C.pk_any[pp] = any;
            C.pk_d[pp] = db;
            if (sbb + gbb > g.raw_bytes[type][pp]) {
                C.pk_raw[pp] = 1;
                C.pk_s[pp] = 0;
                C.pk_g[pp] = g.raw_bytes[type][pp];
            } else {
                C.pk_raw[pp] = 0;
                C.pk_s[pp] = sbb;
                C.pk_g[pp] = gbb;
            }
            C.pk_off[pp] = off;
            if (any)
                off += g.pkt_hdr_bytes + C.pk_s[pp] + C.pk_g[pp] + C.pk_d[pp];
        }
        /* precinct-level bytes, written by the first packet's block */
        if (p == 0 && (uint64_t)P.offset + P.total + (row + 1 == g.rows ? 2u : 0u) <= g.lcod && P.offset >= 6) {
            uint8_t* o = F.out;
            const uint32_t s = row / g.pps;
            if (row % g.pps == 0) {
                const uint32_t so = P.offset - 6;
                o[so] = 0xFF;
                o[so + 1] = 0x20;
                o[so + 2] = 0x00;
                o[so + 3] = 0x04;
                o[so + 4] = (uint8_t)(s >> 8);
                o[so + 5] = (uint8_t)s;
            }
            const uint32_t lprc = P.total - g.prec_hdr_bytes;
            uint8_t* h = o + P.offset;
            for (uint32_t k = 0; k < g.prec_hdr_bytes; k++)
                h[k] = 0;
            h[0] = (uint8_t)(lprc >> 16);
            h[1] = (uint8_t)(lprc >> 8);
            h[2] = (uint8_t)lprc;
            h[3] = (uint8_t)P.q;
            h[4] = (uint8_t)P.r;
            uint32_t bit = 40;
            for (uint32_t gi = 0; gi < g.nb_all; gi++) {
                if (g.gb[gi] == 0xFF)
                    continue;
                const uint32_t c = g.gc[gi], b = g.gb[gi];
                const uint32_t dmode = (g.c[c].b[b].lines[type] > 0 && C.msig[c][b]) ? 2u : 0u;
                h[bit >> 3] |= (uint8_t)(dmode << (6 - (bit & 7)));
                bit += 2;
            }
            /* zero padding closing the precinct */
            for (uint32_t k = 0; k < P.pad; k++)
                o[P.offset + P.total - P.pad + k] = 0;
            if (row + 1 == g.rows) {
                const uint32_t e = P.offset + P.total;
                o[e] = 0xFF;
                o[e + 1] = 0x11;
            }
        }
    }
    if (p == 0 && row == 0) {
        for (uint32_t k = threadIdx.x; k < g.hdr_len; k += blockDim.x)
            F.out[k] = g.hdr[k];
    }
    for (uint32_t k = threadIdx.x; k < RING; k += blockDim.x)
        ring[k] = 0;
    __syncthreads();
    if (!C.pk_any[p])
        return;
    /* the sizes above reproduce the rate-control table exactly; a packet that
     * would not fit its precinct is an internal error, never written */
    if ((uint64_t)C.pk_off[p] + g.pkt_hdr_bytes + C.pk_s[p] + C.pk_g[p] + C.pk_d[p] > (uint64_t)P.offset + P.total ||
        (uint64_t)P.offset + P.total > g.lcod) {
        if (threadIdx.x == 0)
            atomicOr(F.status, 2u);
        return;
    }
    uint8_t* o = F.out + C.pk_off[p];
    const uint32_t l = g.pk_line[p];
    if (threadIdx.x == 0) {
        /* packet header (PackHeaders.c:98-127), sign size 0 */
        const uint32_t raw = C.pk_raw[p], d = C.pk_d[p], gb = C.pk_g[p];
        if (g.short_hdr) {
            o[0] = (uint8_t)((raw << 7) | ((d >> 8) & 0x7F));
            o[1] = (uint8_t)d;
            o[2] = (uint8_t)(gb >> 5);
            o[3] = (uint8_t)((gb & 0x1F) << 3);
            o[4] = 0;
        } else {
            o[0] = (uint8_t)((raw << 7) | ((d >> 13) & 0x7F));
            o[1] = (uint8_t)(d >> 5);
            o[2] = (uint8_t)(((d & 0x1F) << 3) | ((gb >> 17) & 7));
            o[3] = (uint8_t)(gb >> 9);
            o[4] = (uint8_t)(gb >> 1);
            o[5] = (uint8_t)((gb & 1) << 7);
            o[6] = 0;
        }
    }
    o += g.pkt_hdr_bytes;
    /* sub-packets in order: significance, bit-plane counts, data */
    for (int sub = 0; sub < 3; sub++) {
        uint32_t nbytes;
        if (sub == 0) {
            if (C.pk_raw[p])
                continue;
            nbytes = C.pk_s[p];
        } else if (sub == 1) {
            nbytes = C.pk_g[p];
        } else {
            nbytes = C.pk_d[p];
        }
        uint32_t flushed = 0, pos = 0;
        for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
            const uint32_t c = g.gc[gi], b = g.gb[gi];
            const JxsBand& bd = g.c[c].b[b];
            if (l >= bd.lines[type])
                continue;
            const uint32_t y = row * bd.hl + l;
            const uint8_t* gl = F.gcli + bd.gcli_off + (size_t)y * bd.gw;
            const uint8_t* sl = F.smax + bd.smax_off + (size_t)y * bd.sw;
            const uint16_t* cl = F.coeff + bd.coeff_off + (size_t)y * bd.w;
            const uint32_t gt = C.gtli[c][b], ms = C.msig[c][b];
            int kind;
            if (sub == 0) {
                if (!ms)
                    continue;
                kind = ITEM_SIG;
            } else if (sub == 1) {
                kind = C.pk_raw[p] ? ITEM_RAW : ITEM_VLC;
            } else {
                kind = ITEM_DATA;
            }
            pos += emit_line(ring, C, o, nbytes, flushed, pos, kind, gl, sl, cl, bd.w, bd.gw, bd.sw, gt, ms, g.qpih);
        }
        ring_flush(ring, o, nbytes, flushed, (nbytes + 3) / 4);
        o += nbytes;
    }
}

/* ------------------------------------------------------------------ */
/* launchers (C linkage for the host code)                             */
/* ------------------------------------------------------------------ */
extern "C" {

hipError_t jxs_launch_vert(const JxsVertJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    hipLaunchKernelGGL(k_vert, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    return hipGetLastError();
}

hipError_t jxs_launch_horz(const JxsHorzJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    hipLaunchKernelGGL(k_horz, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    return hipGetLastError();
}

hipError_t jxs_launch_rc_table(const JxsFrameDesc* frames, int nfr, uint32_t precincts, hipStream_t st) {
    hipLaunchKernelGGL(k_rc_table, dim3(precincts), dim3(256), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_select(const JxsFrameDesc* frames, int nfr, uint32_t slices, hipStream_t st) {
    hipLaunchKernelGGL(k_select, dim3((slices + 63) / 64), dim3(64), 0, st, frames, nfr, slices);
    return hipGetLastError();
}

hipError_t jxs_launch_pack(const JxsFrameDesc* frames, int nfr, uint32_t max_npk, uint32_t precincts, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3(max_npk, precincts), dim3(256), 0, st, frames, nfr);
    return hipGetLastError();
}

}  // extern "C"

/* Clears n words (graph-capturable replacement for a memset node). */
__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

extern "C" hipError_t jxs_launch_zero(uint32_t* p, uint32_t n, hipStream_t st) {
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}
