/*
 * jxs_dec_kernels.hip -- gfx950 kernels of the JPEG XS decode path.
 *
 *   k_dec_scan    one lane per codestream: walks slice headers and the
 *                 precinct length chain (Lprc) to find every precinct; only
 *                 used when the caller has no precinct table.
 *   k_dec_packets one lane per precinct: band D-modes and GTLIs, and the
 *                 packet chain walked once into a table of sub-packet
 *                 offsets.
 *   k_dec_unpack  one block per (packet, precinct): the packet staged in LDS,
 *                 its band lines decoded as one group sequence, 1,024 groups
 *                 per pass: the significance flags, the unary bit-plane
 *                 counts located all at once (code j ends at the j-th zero
 *                 bit: zero counts per 32-bit word + block prefix sum), then
 *                 each group's data bits at prefix-summed offsets, read as
 *                 one 64-bit window, dequantised and turned into two's
 *                 complement (Packing.c:513-868, Dequant.c:9-57,
 *                 Precinct.c:9-29).
 *   k_dec_slice   one wave per slice coded with vertical prediction or
 *                 separate signs: its precincts and packets in order.
 *   k_inv_rows    one block per row: every band sample of the row loaded at
 *                 once (one memory round trip), all inverse horizontal levels in
 *                 LDS (Idwt.c:11-53 semantics), into an int32 row or, for
 *                 components without vertical levels, straight to the output
 *                 plane with the linear output scaling (NltDec.c:242-265).
 *   k_inv_vert    inverse vertical 5/3 lifting of a second vertical level,
 *                 one lane per column, into int32 rows.
 *   k_inv_last    the last vertical level with the finest horizontal level
 *                 of its low and high rows, one lane per output column pair
 *                 down a strip of row pairs, the output scaling fused.
 *
 * Every codestream read is bounds-checked: a corrupt or truncated
 * codestream sets the frame's status word, and no read leaves the aligned
 * 32-bit words that hold the codestream's bytes.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "jxs_dec.h"
#include "jxs_wave.h"

#define WAVE 64
#define GAS __attribute__((address_space(1)))

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

/* a load through the global address space (not a flat load) */
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const GAS T*)p;
}

/* base[idx] for a wave-uniform row base and a 32-bit per-lane index: the
 * address is the scalar base plus a 32-bit byte offset (no 64-bit per-lane
 * address arithmetic). */
template <typename T>
__device__ __forceinline__ T gldr(const T* base, uint32_t idx) {
    return *(const GAS T*)((const GAS uint8_t*)base + idx * (uint32_t)sizeof(T));
}

template <typename T>
__device__ __forceinline__ void gstr(void* base, uint32_t byte_off, T v) {
    *(GAS T*)((GAS uint8_t*)base + byte_off) = v;
}

__device__ __forceinline__ uint32_t dtrunc(uint32_t gain, uint32_t prio, uint32_t q, uint32_t r) {
    const uint32_t pump = prio < r ? 1u : 0u;
    if (q < gain + pump)
        return 0;
    const uint32_t t = q - gain - pump;
    return t > 15u ? 15u : t;
}

__device__ __forceinline__ uint32_t cs_byte(const JxsDecFrame& F, uint64_t pos) { return pos < F.cs_len ? F.cs[pos] : 0u; }

/* n (1..32) bits starting at absolute bit position bp, MSB first. */
__device__ __forceinline__ uint32_t cs_bits(const JxsDecFrame& F, uint64_t bp, uint32_t n) {
    const uint64_t b = bp >> 3;
    const uint32_t sh = (uint32_t)(bp & 7);
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 5; k++)
        acc = (acc << 8) | cs_byte(F, b + k);
    return (uint32_t)((acc >> (40 - sh - n)) & (n == 32 ? 0xFFFFFFFFull : ((1ull << n) - 1)));
}

/* The codestream as aligned big-endian 32-bit words.  Bit positions given to
 * cs_word/cs_bits64 count from the aligned word below F.cs (sh = 8 x the
 * misalignment).  A word is loaded only if it holds at least one codestream
 * byte, so the load never leaves the pages of the codestream; bytes of such a
 * word outside the codestream are never among the bits a caller keeps (every
 * caller bounds its bits by a sub-packet end inside the codestream). */
typedef const __attribute__((address_space(1))) uint32_t gu32; /* a word known to be in device memory */

struct CsView {
    gu32* w;
    uint32_t nw, sh;
};

__device__ __forceinline__ CsView cs_view(const JxsDecFrame& F) {
    const uintptr_t a = (uintptr_t)F.cs;
    CsView v;
    v.w = (gu32*)(a & ~(uintptr_t)3);
    v.sh = (uint32_t)(a & 3) * 8;
    v.nw = (uint32_t)(((uint64_t)(a & 3) + F.cs_len + 3) >> 2);
    return v;
}

__device__ __forceinline__ uint32_t cs_word(const CsView& v, uint32_t wi) { return wi < v.nw ? __builtin_bswap32(v.w[wi]) : 0u; }

/* the 64 bits from aligned bit position abp on, MSB first */
__device__ __forceinline__ uint64_t cs_bits64(const CsView& v, uint64_t abp) {
    const uint32_t wi = (uint32_t)(abp >> 5), sh = (uint32_t)(abp & 31);
    const uint64_t a = ((uint64_t)cs_word(v, wi) << 32) | cs_word(v, wi + 1);
    const uint32_t c = cs_word(v, wi + 2);
    return sh ? (a << sh) | (uint64_t)(c >> (32 - sh)) : a;
}

/* The 64 bits from bit position bp of the codestream (counted from F.cs),
 * with bits past its end read as 0, as cs_bits reads them. */
__device__ __forceinline__ uint64_t cs_bits64z(const JxsDecFrame& F, const CsView& V, uint64_t bp) {
    const uint64_t end = (uint64_t)F.cs_len * 8;
    if (bp >= end)
        return 0;
    uint64_t w = cs_bits64(V, bp + V.sh);
    if (end - bp < 64)
        w &= ~0ull << (64 - (end - bp));
    return w;
}

/* Every fourth bit of x, from bit 63 down (bits 63, 59, ..., 3), packed into a
 * 16-bit value with bit 63 as its MSB. */
__device__ __forceinline__ uint32_t gather4(uint64_t x) {
    x = (x >> 3) & 0x1111111111111111ull;
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    return (uint32_t)((x | (x >> 24)) & 0xFFFF);
}

/* Exclusive block scan of v (every thread of the block, NW waves); *total
 * the block's sum.  One barrier: tmp (NW words) is read after it, so two
 * scans in a row use different tmp words (a scan's barrier orders the
 * previous scan's reads before the next scan's writes). */
template <uint32_t NW>
__device__ __forceinline__ uint32_t dblock_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const uint32_t x = wave_scan_dpp(v);
    if (lane == WAVE - 1)
        tmp[wid] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) {
        const uint32_t t = tmp[w];
        before += w < wid ? t : 0u;
        tot += t;
    }
    *total = tot;
    return before + x - v;
}

__device__ __forceinline__ int find_dec_frame(const JxsDecFrame* fr, int nfr, uint32_t blk) {
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

/* The frame of block blk, one memory round trip for up to 64 frames: each
 * lane of the wave reads one frame's first block (increasing with the frame)
 * and the wave counts those at or below blk. */
__device__ __forceinline__ int find_dec_frame_wave(const JxsDecFrame* fr, int nfr, uint32_t blk) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    int f = 0;
    for (int c0 = 0; c0 < nfr; c0 += WAVE) {
        const int i = c0 + (int)lane;
        const uint64_t m = __ballot(i < nfr && fr[i].prec_base <= blk);
        f += __popcll(m);
        if (m != ~0ull)
            break;
    }
    return f - 1;
}

/* The codestream offset of precinct `row`: from the caller's precinct table
 * when the frame has one, else as k_dec_scan found it. */
__device__ __forceinline__ uint32_t prec_offset(const JxsDecFrame& F, uint32_t row) {
    return F.prec_in ? F.prec_in[(size_t)row * F.prec_in_stride] : F.prec_off[row];
}

/* One block per codestream: copies a known precinct table, or walks the
 * slice headers and the precinct length chain (lane 0). */
__global__ void k_dec_scan(const JxsDecFrame* __restrict__ frames, int nfr) {
    const JxsDecFrame& F = frames[blockIdx.x];
    const JxsGeom& g = *F.g;
    if (F.prec_in) {
        for (uint32_t r = threadIdx.x; r < g.rows; r += blockDim.x)
            F.prec_off[r] = F.prec_in[(size_t)r * F.prec_in_stride];
        return;
    }
    if (threadIdx.x != 0)
        return;
    uint64_t pos = F.hdr_end;
    for (uint32_t s = 0; s < g.slices; s++) {
        if (cs_byte(F, pos) != 0xFF || cs_byte(F, pos + 1) != 0x20 || cs_byte(F, pos + 2) != 0 || cs_byte(F, pos + 3) != 4 ||
            ((cs_byte(F, pos + 4) << 8) | cs_byte(F, pos + 5)) != s || pos + 6 > F.cs_len) {
            atomicOr(F.status, 1u);
            for (uint32_t r = s * g.pps; r < g.rows; r++)
                F.prec_off[r] = F.cs_len;
            return;
        }
        pos += 6;
        const uint32_t first = s * g.pps;
        const uint32_t n = (s == g.slices - 1) ? g.rows - first : g.pps;
        for (uint32_t i = 0; i < n; i++) {
            F.prec_off[first + i] = (uint32_t)pos;
            const uint32_t lprc = (cs_byte(F, pos) << 16) | (cs_byte(F, pos + 1) << 8) | cs_byte(F, pos + 2);
            pos += g.prec_hdr_bytes + (uint64_t)lprc;
            if (pos > F.cs_len) {
                atomicOr(F.status, 1u);
                for (uint32_t r = first + i + 1; r < g.rows; r++)
                    F.prec_off[r] = F.cs_len;
                return;
            }
        }
    }
}

#define PMAXG 1024            /* groups of a packet decoded per pass */
#define PK_THREADS 256        /* threads of an unpack block */
#define PGPT (PMAXG / PK_THREADS) /* consecutive groups of a thread in a pass */
#define PNB (JXS_MAXC * JXS_NBC) /* band lines of a packet, at most */
#define PK_STAGE 3072         /* codestream words of a packet staged in LDS (12 KiB) */

/* one band line of the packet being decoded, in codestream order */
struct PkBand {
    uint32_t g0;  /* first group of the line in the packet's group sequence */
    uint32_t gt;  /* GTLI */
    uint32_t sig; /* significance flags coded */
    uint32_t s0;  /* first flag of the line in the significance sub-packet */
    uint32_t w;   /* coefficients of the line */
    uint32_t pad_;
    int16_t* row; /* coefficient row */
};

struct UnpackCtx {
    PkBand band[PNB + 1]; /* + a sentinel whose g0 is the packet's group count */
    uint32_t nband, s_bit, g_bit, d_bit, g_end, d_end, raw, ok;
    uint32_t zpos[PMAXG]; /* aligned bit position of the zero ending code j of the pass */
    uint32_t scan[2][PK_THREADS / WAVE]; /* alternating scan words */
    uint32_t stg[PK_STAGE + 1]; /* the packet's codestream words, byte-swapped, and a spare */
};

/* A block's packet words: s[0 .. n - 1] is the LDS copy of codestream words
 * lo .. lo + n - 1 (n = 0 when the packet did not fit, and V reads HBM). */
struct PkSrc {
    CsView V;
    const uint32_t* s;
    uint32_t lo, n;
};

/* One lane per precinct of the launch: parses the precinct header's band
 * D-modes, derives every band's GTLI from the precinct's quantisation and
 * refinement (Precinct.c:9-29 restated), and walks the packet chain once
 * (Packing.c:513-868: packet header, then significance, GCLI and data
 * sub-packets), recording where each packet's sub-packets start and end. */
__global__ void __launch_bounds__(64) k_dec_packets(const JxsDecFrame* __restrict__ frames, int nfr, uint32_t precincts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= precincts)
        return;
    const int f = find_dec_frame(frames, nfr, t);
    const JxsDecFrame& F = frames[f];
    const JxsGeom& g = *F.g;
    const uint32_t row = t - F.prec_base;
    if (row >= g.rows)
        return;
    uint32_t* T = F.pkt + (size_t)row * F.pkt_stride;
    if (row >= F.rows_valid) {
        /* a precinct of a slice the host found broken: skipped, the frame's
         * error comes from the host */
        T[JXS_DPKT_BAD] = 1u;
        T[JXS_DPKT_ORDERED] = 0u;
        return;
    }
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    const uint64_t off = prec_offset(F, row);
    const CsView V = cs_view(F);
    /* the precinct header and its band modes, read as three 64-bit windows */
    const uint64_t hw = cs_bits64z(F, V, off * 8), m0 = cs_bits64z(F, V, off * 8 + 40), m1 = cs_bits64z(F, V, off * 8 + 104);
    const uint32_t lprc = (uint32_t)(hw >> 40);
    const uint32_t q = (uint32_t)(hw >> 32) & 0xFFu, r = (uint32_t)(hw >> 24) & 0xFFu;
    uint16_t mode[JXS_MAXC * JXS_NBC];
    for (uint32_t k = 0; k < JXS_MAXC * JXS_NBC; k++)
        mode[k] = 0;
    uint32_t nmode = 0; /* band modes read so far: 2 bits each from bit 40 of the header */
    bool bad = false;
    bool ordered = g.fs != 0; /* separate sign sub-packets */
    for (uint32_t gi = 0; gi < g.nb_all; gi++) {
        if (g.gb[gi] == 0xFF)
            continue;
        const uint32_t c = g.gc[gi], b = g.gb[gi];
        const uint32_t d = (uint32_t)((nmode < 32 ? m0 >> (62 - 2 * nmode) : m1 >> (62 - 2 * (nmode - 32))) & 3u);
        nmode++;
        mode[c * JXS_NBC + b] = (uint16_t)(d | dtrunc(g.c[c].b[b].gain, g.c[c].b[b].prio, q, r) << 8);
        if (d & 1)
            ordered = true; /* vertical prediction: needs the line above */
    }
    for (uint32_t k = 0; k < JXS_MAXC * JXS_NBC / 2; k++)
        T[k] = mode[2 * k] | (uint32_t)mode[2 * k + 1] << 16;
    const uint64_t pend = off + g.prec_hdr_bytes + (uint64_t)lprc;
    uint64_t pos = off + g.prec_hdr_bytes;
    for (uint32_t pp = 0; pp < g.npk; pp++) {
        uint32_t* E = T + JXS_DPKT_HDR + JXS_DPKT_WORDS * pp;
        bool any = false;
        uint32_t sig_bits = 0, groups = 0;
        for (uint32_t gi = g.pk_bs[pp]; gi < g.pk_be[pp]; gi++) {
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            if (g.pk_line[pp] < bd.lines[type]) {
                any = true;
                groups += bd.gw;
                if (mode[g.gc[gi] * JXS_NBC + g.gb[gi]] & 2)
                    sig_bits += bd.sw;
            }
        }
        if (!any || bad) {
            E[4] = 0;
            continue;
        }
        uint32_t raw, ld, lg, ls;
        const uint64_t ph = cs_bits64z(F, V, pos * 8); /* the packet header */
        raw = (uint32_t)(ph >> 63);
        if (g.short_hdr) {
            ld = (uint32_t)(ph >> 48) & 0x7FFFu;
            lg = (uint32_t)(ph >> 35) & 0x1FFFu;
            ls = (uint32_t)(ph >> 24) & 0x7FFu;
        } else {
            ld = (uint32_t)(ph >> 43) & 0xFFFFFu;
            lg = (uint32_t)(ph >> 23) & 0xFFFFFu;
            ls = (uint32_t)(ph >> 8) & 0x7FFFu;
        }
        /* the three lengths must fit what is left of the precinct after the
         * header, the sign length counted even without a sign sub-packet
         * (Packing.c:584-586) */
        if (pos + g.pkt_hdr_bytes + (uint64_t)ld + lg + ls > pend)
            bad = true;
        /* raw bit-plane counts (4 bits a group) must fit their sub-packet
         * (Packing.c:735-740) */
        if (raw && (4ull * groups + 7) / 8 > lg)
            bad = true;
        if (!g.fs)
            ls = 0; /* the field is only a sub-packet length with Fs = 1 */
        const uint32_t sb = raw ? 0u : (sig_bits + 7) / 8;
        const uint64_t s_off = pos + g.pkt_hdr_bytes, g_off = s_off + sb, d_off = g_off + lg, nxt = d_off + ld + ls;
        E[0] = (uint32_t)(s_off * 8);
        E[1] = (uint32_t)(g_off * 8);
        E[2] = (uint32_t)(d_off * 8);
        E[3] = (uint32_t)(nxt * 8);
        E[4] = bad ? 0u : (1u | raw << 1);
        pos = nxt;
        if (pos > pend)
            bad = true; /* later packets are not decoded */
    }
    if (pend > F.cs_len || pos > pend || (uint64_t)pend * 8 >= 0xFFFFFFFFull)
        bad = true;
    if (bad)
        atomicOr(F.status, 1u);
    T[JXS_DPKT_BAD] = bad ? 1u : 0u;
    T[JXS_DPKT_ORDERED] = ordered ? 1u : 0u;
}

/* A coefficient as the reference decoder finishes it: its 16-bit word
 * after unpacking (the magnitude bit-planes shifted left by the GTLI within
 * 16 bits, the sign at bit 15: Packing.c:114-131, 221-247), dequantised when
 * its group's bit-plane count, compared as a signed byte, exceeds the GTLI
 * (Dequant.c:9-45), then turned into two's complement (Precinct.c:9-14).  A
 * valid codestream never carries magnitude bits into bit 15; a damaged one
 * decodes here to what the reference makes of it. */
__device__ __forceinline__ int16_t coef_final(uint32_t u, uint32_t gc, uint32_t gt, uint32_t qpih) {
    if (gt > 0 && (int32_t)(int8_t)gc > (int32_t)gt && (u & 0x7FFFu)) {
        if (qpih == 0) {
            u |= 1u << (gt - 1);
        } else {
            const uint32_t sign = u & 0x8000u, sc = (gc - gt + 1) & 0xFFu;
            uint32_t val = u & 0x7FFFu, acc = val;
            if (sc < 16)
                for (val >>= sc; val > 0; val >>= sc)
                    acc += val;
            u = (acc & 0xFFFFu) | sign;
        }
    }
    return (u & 0x8000u) ? (int16_t)(-(int32_t)(u & 0x7FFFu)) : (int16_t)u;
}

/* The 16-bit words (before the sign) of a group with np >= 1 magnitude
 * bit-planes and truncation gt whose planes start at aligned bit position
 * abp: only the last min(np, 16 - gt) planes reach the 16 bits kept. */
__device__ __forceinline__ void group_words(const CsView& V, uint64_t abp, uint32_t np, uint32_t gt, uint32_t w[4]) {
    const uint32_t q = min(np, 16u - gt);
    const uint64_t v = cs_bits64(V, abp + 4ull * (np - q));
#pragma unroll
    for (int t4 = 0; t4 < 4; t4++)
        w[t4] = ((gather4(v << t4) >> (16 - q)) << gt) & 0xFFFFu;
}

/* The packet-wide values the passes start from (all block-uniform). */
struct PkPass {
    uint32_t nband, total, raw;
    uint32_t s_bit; /* aligned bit position of the significance sub-packet */
    uint32_t gend;  /* aligned end of the GCLI sub-packet */
    uint32_t d_end; /* end of the data sub-packet (unaligned bit position) */
    uint32_t gp;    /* aligned bit position of the first GCLI code */
    uint32_t dp;    /* bit position of the first data bit */
    uint32_t qpih;
};

/* The four coefficients of a group with np <= 8 magnitude bit-planes,
 * truncation gt with np + gt <= 15, deadzone reconstruction (Qpih = 0): the
 * result of coef_final on each, as two words of two int16 each.  v holds the
 * group's data from its sign nibble on.  Each plane nibble is spread to one
 * bit per byte (n * 0x204081 puts nibble bit b at bit 8b, no carries), so the
 * planes accumulate into four byte-wide magnitudes at once; the halves of a
 * word then hold two coefficients through reconstruction and sign. */
__device__ __forceinline__ uint2 group_fast(uint64_t v, uint32_t np, uint32_t gt) {
    const uint32_t x = (uint32_t)(v >> 28); /* planes 0..7, one nibble each, first plane highest */
    uint32_t A = 0;                         /* byte 3 - t: coefficient t's magnitude */
    for (uint32_t q = 0; q < np; q++)
        A = (A << 1) | ((((x >> (28 - 4 * q)) & 15u) * 0x204081u) & 0x01010101u);
    const uint32_t sg = ((((uint32_t)(v >> 60)) * 0x204081u) & 0x01010101u); /* byte 3 - t: coefficient t's sign */
    uint32_t r[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        /* coefficients 2h (low half) and 2h + 1 (high half) */
        const uint32_t m = ((A >> (24 - 16 * h)) & 0xFFu) | (((A >> (16 - 16 * h)) & 0xFFu) << 16);
        const uint32_t s = ((sg >> (24 - 16 * h)) & 1u) | (((sg >> (16 - 16 * h)) & 1u) << 16);
        const uint32_t nz = ((m + 0x00FF00FFu) >> 8) & 0x00010001u;        /* 1 where m != 0 */
        const uint32_t mag = ((((m << 1) | nz) << gt) >> 1) & 0x7FFF7FFFu; /* (2m + 1) << gt >> 1, 0 for m = 0 */
        const uint32_t neg = s & nz;                                       /* -0 is 0 */
        r[h] = (mag ^ (neg * 0xFFFFu)) + neg;                              /* two's complement per half */
    }
    return make_uint2(r[0], r[1]);
}

template <bool STAGED>
__device__ __forceinline__ uint32_t pkx_word(const PkSrc& S, uint32_t wi) {
    if (STAGED)
        return S.s[min(wi - S.lo, (uint32_t)PK_STAGE)]; /* every read of a staged packet lies in its words */
    return cs_word(S.V, wi);
}

template <bool STAGED>
__device__ __forceinline__ uint64_t pkx_bits64(const PkSrc& S, uint32_t abp) {
    const uint32_t wi = abp >> 5, sh = abp & 31;
    const uint64_t a = ((uint64_t)pkx_word<STAGED>(S, wi) << 32) | pkx_word<STAGED>(S, wi + 1);
    const uint32_t c = pkx_word<STAGED>(S, wi + 2);
    return sh ? (a << sh) | (uint64_t)(c >> (32 - sh)) : a;
}

/* The passes of one packet (see k_dec_unpack); returns whether the GCLI
 * sub-packet ran out before every code was found.  UNIFORM: the frame's
 * quantisation is uniform (Qpih 1), else deadzone; a template parameter so
 * that its tests inside the data phase fold away instead of being kept as
 * lane masks across the passes. */
template <bool STAGED, bool UNIFORM>
__device__ __forceinline__ bool unpack_passes(UnpackCtx& C, const PkSrc& S, const PkPass& P0, uint32_t tid) {
    const CsView& V = S.V;
    const uint32_t nband = P0.nband, total = P0.total, raw = P0.raw, s_bit = P0.s_bit, gend = P0.gend,
                   d_end = P0.d_end, qpih = UNIFORM ? 1u : 0u;
    uint32_t gp = P0.gp, dp = P0.dp;
    uint32_t status = 0;
    bool err = false;
    for (uint32_t c0 = 0; c0 < total; c0 += PMAXG) {
        const uint32_t cn = min((uint32_t)PMAXG, total - c0);
        const uint32_t i0 = c0 + tid * PGPT; /* this thread's first group */
        const uint32_t nv = tid * PGPT < cn ? min((uint32_t)PGPT, cn - tid * PGPT) : 0u; /* its groups in the pass */
        /* band line of the first group (binary search over the lines' first groups) */
        uint32_t bl = 0;
        {
            uint32_t lo = 0, hi = nband - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (C.band[mid].g0 <= i0)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            bl = lo;
        }
        /* uni: the thread's PGPT groups lie in one band line (all but the
         * threads at a line's end and the pass's last ones), whose fields are
         * then read once */
        const uint32_t bl0 = bl, g00 = C.band[bl0].g0;
        const bool uni = nv == PGPT && i0 + PGPT <= C.band[bl0 + 1].g0;
        /* significance: a coded group is one whose group of 8 is significant */
        uint32_t bidx[PGPT], gts[PGPT], coded = 0, ncoded = 0;
        if (uni) {
            const uint32_t gt0 = C.band[bl0].gt;
#pragma unroll
            for (uint32_t k = 0; k < PGPT; k++) {
                bidx[k] = bl0;
                gts[k] = gt0;
            }
            coded = (1u << PGPT) - 1;
            if (!raw && C.band[bl0].sig) {
                /* the flags of the groups of 8 holding groups i0 and i0 + 3 */
                const uint32_t rel = i0 - g00, ab = s_bit + C.band[bl0].s0 + rel / 8;
                const uint32_t sh = ab & 31, a0 = pkx_word<STAGED>(S, ab >> 5), a1 = pkx_word<STAGED>(S, (ab >> 5) + 1);
                const uint32_t two = (sh ? (a0 << sh) | (a1 >> (32 - sh)) : a0) >> 30;
                const uint32_t r = rel & 7;
                const uint32_t m0 = r + PGPT <= 8 ? (1u << PGPT) - 1 : (1u << (8 - r)) - 1; /* groups under the first flag */
                const uint32_t insig = ((two & 2u) ? m0 : 0u) | ((two & 1u) ? ((1u << PGPT) - 1) & ~m0 : 0u);
                coded &= ~insig;
            }
            ncoded = (uint32_t)__popc(coded);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < PGPT; k++) {
                const uint32_t i = i0 + k;
                bidx[k] = 0;
                gts[k] = 0;
                if (k >= nv)
                    continue;
                while (C.band[bl + 1].g0 <= i)
                    bl++;
                bidx[k] = bl;
                gts[k] = C.band[bl].gt;
                bool cd = true;
                if (!raw && C.band[bl].sig) {
                    const uint32_t ab = s_bit + C.band[bl].s0 + (i - C.band[bl].g0) / 8;
                    cd = ((pkx_word<STAGED>(S, ab >> 5) >> (31 - (ab & 31))) & 1u) == 0; /* flag 1: insignificant */
                }
                coded |= (cd ? 1u : 0u) << k;
                ncoded += cd ? 1u : 0u;
            }
        }
        /* the zero bits (code ends) of one block-wide stretch of GCLI words
         * from word w on: this thread's word wi = w + tid, bits outside
         * [gp, gend) set */
        auto code_word = [&](uint32_t w) -> uint32_t {
            const uint32_t wi = w + tid;
            uint32_t word = 0xFFFFFFFFu;
            if (!raw && (uint64_t)wi * 32 < gend) {
                word = pkx_word<STAGED>(S, wi);
                const uint32_t lo = wi * 32;
                if (lo < gp)
                    word |= ~(0xFFFFFFFFu >> (gp - lo)); /* bits before the pass's first code */
                if (lo + 32 > gend)
                    word |= (0xFFFFFFFFu >> (gend - lo)); /* bits after the sub-packet end */
            }
            return ~word;
        };
        /* one scan numbers the coded groups (low half) and the first
         * stretch's zero bits (high half): both totals stay below 2^16 */
        uint32_t z0 = code_word(gp >> 5);
        uint32_t both, slot = 0; /* scan words: alternate between consecutive scans */
        const uint32_t ex0 = dblock_scan<PK_THREADS / WAVE>(ncoded | (uint32_t)__popc(z0) << 16, C.scan[slot], &both);
        slot ^= 1;
        const uint32_t cbase = ex0 & 0xFFFFu, ncodes = __builtin_amdgcn_readfirstlane(both & 0xFFFFu);
        /* GCLI codes */
        if (!raw) {
            uint32_t found = 0;
            for (uint32_t w = gp >> 5; found < ncodes; w += blockDim.x) {
                if ((uint64_t)w * 32 >= gend) {
                    err = true;
                    break;
                }
                const uint32_t wi = w + tid;
                uint32_t z, t, ex;
                if (w == (gp >> 5)) {
                    z = z0;
                    ex = ex0 >> 16;
                    t = both >> 16;
                } else {
                    z = code_word(w);
                    ex = dblock_scan<PK_THREADS / WAVE>((uint32_t)__popc(z), C.scan[slot], &t);
                    slot ^= 1;
                }
                uint32_t k = found + ex;
                while (z) {
                    const uint32_t lead = __clz(z);
                    if (k < ncodes)
                        C.zpos[k] = wi * 32 + lead;
                    k++;
                    z &= ~(0x80000000u >> lead);
                }
                found += __builtin_amdgcn_readfirstlane(t);
            }
            __syncthreads();
        }
        /* GCLI of every group, data bit counts */
        uint32_t gcs[PGPT], nbits = 0;
        uint32_t j = cbase;
#pragma unroll
        for (uint32_t k = 0; k < PGPT; k++) {
            gcs[k] = 0;
            if (k >= nv || err)
                continue;
            const uint32_t gt = gts[k];
            uint32_t gc = gt;
            if (raw) {
                const uint32_t ab = gp + 4 * (tid * PGPT + k);
                gc = (uint32_t)(pkx_bits64<STAGED>(S, ab) >> 60);
            } else if ((coded >> k) & 1u) {
                const uint32_t start = j == 0 ? gp : C.zpos[j - 1] + 1;
                const uint32_t run = C.zpos[j] - start;
                j++;
                if (run >= 32) { /* no code of 32 bits or more (Packing.c:100-107) */
                    status = 1;
                    gc = gt;
                } else {
                    gc = gt + run;
                }
            }
            gcs[k] = gc;
            nbits += gc > gt ? 4 * (gc - gt + 1) : 0u;
        }
        uint32_t dtotal;
        const uint32_t dbase = dblock_scan<PK_THREADS / WAVE>(nbits, C.scan[slot], &dtotal);
        /* each group's data */
        uint32_t gbp = dp + dbase; /* bit position of the group's data */
        uint2 outs[PGPT];
#pragma unroll
        for (uint32_t k = 0; k < PGPT; k++) {
            outs[k] = make_uint2(0, 0);
            if (k >= nv || err)
                continue;
            const uint32_t gt = gts[k];
            uint32_t gc = gcs[k];
            uint32_t np = gc > gt ? gc - gt : 0u;
            const uint32_t nb = np ? 4 * (np + 1) : 0u;
            const uint32_t gbp_k = gbp;
            gbp += nb;
            uint64_t v = 0;
            if (nb) {
                if ((uint64_t)gbp_k + nb > d_end) {
                    status = 1;
                    gc = gt; /* decoded as zeros */
                    np = 0;
                } else {
                    v = pkx_bits64<STAGED>(S, gbp_k + V.sh);
                }
            }
            if (np == 0) {
                /* no data: four zeros */
            } else if (qpih == 0 && np <= 8 && gt <= 15 && np + gt <= 15) {
                outs[k] = group_fast(v, np, gt);
            } else {
                uint32_t mag[4] = {0, 0, 0, 0};
                const uint32_t sg = (uint32_t)(v >> 60);
                if (np <= 15) {
#pragma unroll
                    for (int t4 = 0; t4 < 4; t4++)
                        mag[t4] = ((gather4(v << (4 + t4)) >> (16 - np)) << gt) & 0xFFFFu;
                } else { /* only in a damaged codestream: planes past the first window */
                    group_words(V, (uint64_t)gbp_k + 4 + V.sh, np, gt, mag);
                }
                uint32_t val[4];
#pragma unroll
                for (uint32_t t4 = 0; t4 < 4; t4++)
                    val[t4] = (uint16_t)coef_final(mag[t4] | ((sg >> (3 - t4)) & 1u) << 15, gc, gt, qpih);
                outs[k] = make_uint2(val[0] | val[1] << 16, val[2] | val[3] << 16);
            }
        }
        /* stores: the thread's groups in one line and inside its width go out
         * as 8-byte words from one base */
        int16_t* cp0 = C.band[bl0].row + 4 * (i0 - g00);
        if (uni && !err && 4 * (i0 - g00) + 4 * PGPT <= C.band[bl0].w && ((uintptr_t)cp0 & 7) == 0) {
#pragma unroll
            for (uint32_t k = 0; k < PGPT; k++)
                *(uint2*)(cp0 + 4 * k) = outs[k];
        } else {
#pragma unroll
            for (uint32_t k = 0; k < PGPT; k++) {
                if (k >= nv || err)
                    continue;
                const PkBand& B = C.band[bidx[k]];
                const uint32_t i = i0 + k - B.g0; /* group within its line */
                int16_t* cp = B.row + 4 * i;
                if (4 * i + 4 <= B.w && ((uintptr_t)cp & 7) == 0) {
                    *(uint2*)cp = outs[k];
                } else {
                    const uint32_t val[4] = {outs[k].x & 0xFFFFu, outs[k].x >> 16, outs[k].y & 0xFFFFu, outs[k].y >> 16};
                    for (uint32_t t4 = 0; t4 < 4 && 4 * i + t4 < B.w; t4++)
                        cp[t4] = (int16_t)val[t4];
                }
            }
        }
        if (err)
            break;
        /* the next pass starts after this pass's codes and data */
        if (raw)
            gp += 4 * cn;
        else if (ncodes)
            gp = __builtin_amdgcn_readfirstlane(C.zpos[ncodes - 1] + 1);
        dp += __builtin_amdgcn_readfirstlane(dtotal);
        __syncthreads(); /* zpos and the scan words are reused */
    }
    return err || __syncthreads_or(status);
}

/* The passes of a packet too long to stage (read from HBM): a call, so
 * that its registers do not weigh on the staged path. */
__device__ __noinline__ bool unpack_passes_hbm(UnpackCtx& C, const PkSrc& S, const PkPass& P0, uint32_t tid) {
    return P0.qpih ? unpack_passes<false, true>(C, S, P0, tid) : unpack_passes<false, false>(C, S, P0, tid);
}

/* One block per (packet, precinct).  Inside a packet the significance,
 * GCLI and data sub-packets each concatenate the packet's band lines at bit
 * granularity (Packing.c:513-868), so the packet is decoded as one sequence
 * of groups, PMAXG at a time: each of the block's PK_THREADS threads takes
 * PGPT consecutive groups,
 * reads their significance flags, a block scan numbers the coded groups,
 * the unary GCLI codes are located all at once (code j ends at the j-th zero
 * bit: zero counts per 32-bit word + block prefix sum), a second scan places
 * every group's data bits, and each group's data (the sign nibble, then
 * gcli - gtli bit planes of 4 bits, at most 64 bits) is read as one 64-bit
 * window, dequantised (Dequant.c:9-57) and stored in two's complement. */
__global__ void __launch_bounds__(PK_THREADS) __attribute__((amdgpu_waves_per_eu(8))) k_dec_unpack(const JxsDecFrame* __restrict__ frames, int nfr) {
    __shared__ UnpackCtx C;
    const int f = find_dec_frame_wave(frames, nfr, blockIdx.y);
    const JxsDecFrame& F = frames[f];
    const JxsGeom& g = *F.g;
    const uint32_t row = blockIdx.y - F.prec_base;
    const uint32_t p = blockIdx.x;
    if (p >= g.npk || p >= F.npk_dec || row >= g.rows)
        return;
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    const uint32_t l = g.pk_line[p];
    const CsView V = cs_view(F);
    const uint32_t tid = threadIdx.x;
    gu32* T = (gu32*)(uintptr_t)(F.pkt + (size_t)row * F.pkt_stride);
    gu32* E = T + JXS_DPKT_HDR + JXS_DPKT_WORDS * p;
    const uint32_t fl = E[4];
    /* the packet's words, from its significance sub-packet to two words past
     * its data (every word a pass reads: the significance flags, the GCLI
     * codes and each group's 64-bit data window), staged in LDS by the whole
     * block with independent loads; words past the codestream stage as 0, as
     * cs_word reads them */
    uint32_t st_lo = 0, st_n = 0;
    if ((fl & 1u) && V.nw > 0) {
        const uint64_t lo = (E[0] + (uint64_t)V.sh) >> 5;
        const uint64_t hi = (((uint64_t)E[3] + V.sh + 31) >> 5) + 2u;
        if (hi > lo && hi - lo <= PK_STAGE) {
            st_lo = (uint32_t)lo;
            st_n = (uint32_t)(hi - lo);
        }
    }
    st_lo = __builtin_amdgcn_readfirstlane(st_lo);
    st_n = __builtin_amdgcn_readfirstlane(st_n);
    for (uint32_t k0 = 0; k0 < st_n; k0 += 4 * PK_THREADS) {
        uint32_t sw[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t wi = st_lo + min(k0 + j * PK_THREADS + tid, st_n - 1);
            const uint32_t wv = V.w[min(wi, V.nw - 1)];
            sw[j] = wi < V.nw ? __builtin_bswap32(wv) : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) { /* past the packet: into the spare word */
            const uint32_t k = k0 + j * PK_THREADS + tid;
            C.stg[k < st_n ? k : PK_STAGE] = sw[j];
        }
    }
    if (tid >= WAVE && tid < 2 * WAVE) {
        /* a slice with an ordered precinct is decoded by k_dec_slice: one
         * lane per precinct of the slice */
        const uint32_t lane = tid - WAVE;
        const uint32_t e0 = E[0], e1 = E[1], e2 = E[2], e3 = E[3], bad = T[JXS_DPKT_BAD];
        const uint32_t r0 = row - row % g.pps, r1 = min(r0 + g.pps, g.rows);
        uint32_t ordered = 0;
        for (uint32_t rr = r0 + lane; rr < r1; rr += WAVE)
            ordered |= ((gu32*)(uintptr_t)F.pkt)[(size_t)rr * F.pkt_stride + JXS_DPKT_ORDERED];
        const bool any_ordered = __ballot(ordered != 0) != 0ull;
        if (lane == 0) {
            C.ok = (fl & 1u) && bad == 0 && !any_ordered ? 1u : 0u;
            C.raw = (fl >> 1) & 1u;
            C.s_bit = e0;
            C.g_bit = e1;
            C.g_end = e2;
            C.d_bit = e2;
            C.d_end = e3;
        }
    } else if (tid < WAVE) {
        /* the packet's band lines, one lane per candidate band (at most PNB < 64),
         * compacted in band order with running group and flag counts */
        const uint32_t gb0 = g.pk_bs[p], gb1 = g.pk_be[p];
        const uint32_t gi = tid; /* every band's indices are read with the packet's range */
        const uint32_t c = gi < g.nb_all ? g.gc[gi] : 0u, b = gi < g.nb_all ? g.gb[gi] : 0u;
        bool ex = false;
        uint32_t gw = 0, sws = 0, mode = 0, w = 0;
        int16_t* rp = nullptr;
        if (gi >= gb0 && gi < gb1) {
            const JxsBand& bd = g.c[c].b[b];
            if (l < bd.lines[type]) {
                ex = true;
                const uint32_t k = c * JXS_NBC + b;
                mode = (T[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
                gw = bd.gw;
                sws = (mode & 2) ? bd.sw : 0u;
                w = bd.w;
                rp = F.coef + bd.coeff_off + (size_t)(row * bd.hl + l) * bd.cs;
            }
        }
        const uint64_t m = __ballot(ex);
        const uint32_t idx = (uint32_t)__popcll(m & ((1ull << tid) - 1));
        uint32_t G = gw, S = sws; /* inclusive scans over the lanes */
#pragma unroll
        for (uint32_t o = 1; o < WAVE; o <<= 1) {
            const uint32_t yg = __shfl_up(G, o, WAVE), ys = __shfl_up(S, o, WAVE);
            if (tid >= o) {
                G += yg;
                S += ys;
            }
        }
        if (ex) {
            PkBand& B = C.band[idx];
            B.g0 = G - gw;
            B.gt = mode >> 8;
            B.sig = (mode & 2) ? 1u : 0u;
            B.s0 = S - sws;
            B.w = w;
            B.row = rp;
        }
        if (tid == WAVE - 1) {
            const uint32_t n = (uint32_t)__popcll(m);
            C.band[n].g0 = G;
            C.nband = n;
        }
    }
    __syncthreads();
    if (!C.ok)
        return;
    /* block-uniform values, kept in scalar registers */
    PkPass P0;
    P0.nband = __builtin_amdgcn_readfirstlane(C.nband);
    P0.total = __builtin_amdgcn_readfirstlane(C.band[C.nband].g0);
    P0.raw = __builtin_amdgcn_readfirstlane(C.raw);
    P0.s_bit = __builtin_amdgcn_readfirstlane(C.s_bit + V.sh);
    P0.gend = __builtin_amdgcn_readfirstlane(C.g_end + V.sh);
    P0.d_end = __builtin_amdgcn_readfirstlane(C.d_end);
    P0.gp = __builtin_amdgcn_readfirstlane(C.g_bit + V.sh);
    P0.dp = __builtin_amdgcn_readfirstlane(C.d_bit);
    P0.qpih = __builtin_amdgcn_readfirstlane(g.qpih);
    const PkSrc S = {V, C.stg, st_lo, st_n};
    const bool err = !st_n ? unpack_passes_hbm(C, S, P0, tid)
        : P0.qpih         ? unpack_passes<true, true>(C, S, P0, tid)
                          : unpack_passes<true, false>(C, S, P0, tid);
    if (err && tid == 0)
        atomicOr(F.status, 1u);
}

/* ------------------------------------------------------------------ */
/* in-order slice decoding: vertical prediction, separate sign sub-packets */
/* ------------------------------------------------------------------ */
#define SL_MAXG 8192 /* groups of one packet's band lines */


__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (uint32_t o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= o)
            x += y;
    }
    return x;
}

/* One wave per slice whose precincts use vertical prediction or whose
 * codestream sends signs apart (Packing.c:513-868): precincts and packets in
 * order, because a vertically predicted bit-plane count is coded against the
 * line above (the band's previous line, or its last line in the precinct
 * above: Packing.c:249-385) and the sign sub-packet follows the data of its
 * packet (Packing.c:221-247, 829-866).  Lane 0 decodes the significance flags
 * and bit-plane counts of a packet in order; the wave then decodes the data of
 * each band line 64 groups at a time at prefix-summed bit offsets, and with
 * Fs = 1 applies the signs, one bit per non-zero coefficient. */
__global__ void __launch_bounds__(WAVE) k_dec_slice(const JxsDecFrame* __restrict__ frames, int nfr) {
    __shared__ uint8_t gsh[SL_MAXG];
    __shared__ uint32_t sh_err;
    int lo = 0, hi = nfr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frames[mid].slice_base <= blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const JxsDecFrame& F = frames[lo];
    const JxsGeom& g = *F.g;
    const uint32_t s = blockIdx.x - F.slice_base;
    if (s >= g.slices)
        return;
    const uint32_t r0 = s * g.pps, r1 = min(r0 + g.pps, g.rows);
    uint32_t ordered = 0;
    for (uint32_t rr = r0; rr < r1; rr++)
        ordered |= F.pkt[(size_t)rr * F.pkt_stride + JXS_DPKT_ORDERED];
    if (!ordered)
        return;
    const uint32_t lane = threadIdx.x;
    const CsView V = cs_view(F);
    if (lane == 0)
        sh_err = 0;
    __syncthreads();
    for (uint32_t row = r0; row < r1; row++) {
        const uint32_t* T = F.pkt + (size_t)row * F.pkt_stride;
        if (T[JXS_DPKT_BAD])
            break; /* already reported */
        const uint32_t type = row + 1 >= g.rows ? 1 : 0;
        const uint64_t off = prec_offset(F, row);
        const uint64_t pend = off + g.prec_hdr_bytes + (uint64_t)cs_bits(F, off * 8, 24);
        uint64_t pos = off + g.prec_hdr_bytes;
        bool bad = false;
        if (row == r0) /* the first precinct of a slice has nothing above it (Packing.c:536-542) */
            for (uint32_t k = 0; k < JXS_MAXC * JXS_NBC; k++)
                bad |= ((T[k >> 1] >> (16 * (k & 1))) & 1u) != 0;
        const uint32_t npk = min(g.npk, F.npk_dec);
        for (uint32_t p = 0; p < npk && !bad; p++) {
            const uint32_t l = g.pk_line[p];
            uint32_t sig_bits = 0, any = 0;
            for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
                const uint32_t c = g.gc[gi], b = g.gb[gi];
                const JxsBand& bd = g.c[c].b[b];
                if (l < bd.lines[type]) {
                    any = 1;
                    const uint32_t k = c * JXS_NBC + b;
                    if ((T[k >> 1] >> (16 * (k & 1))) & 2u)
                        sig_bits += bd.sw;
                }
            }
            if (!any)
                continue;
            uint32_t raw, ld, lg, ls;
            if (g.short_hdr) {
                raw = cs_bits(F, pos * 8, 1);
                ld = cs_bits(F, pos * 8 + 1, 15);
                lg = cs_bits(F, pos * 8 + 16, 13);
                ls = cs_bits(F, pos * 8 + 29, 11);
            } else {
                raw = cs_bits(F, pos * 8, 1);
                ld = cs_bits(F, pos * 8 + 1, 20);
                lg = cs_bits(F, pos * 8 + 21, 20);
                ls = cs_bits(F, pos * 8 + 41, 15);
            }
            if (pos + g.pkt_hdr_bytes + (uint64_t)ld + lg + ls > pend) { /* Packing.c:584-586 */
                bad = true;
                break;
            }
            if (!g.fs)
                ls = 0;
            const uint64_t s_off = pos + g.pkt_hdr_bytes, g_off = s_off + (raw ? 0u : (sig_bits + 7) / 8);
            const uint64_t d_off = g_off + lg, n_off = d_off + ld, nxt = n_off + ls;
            if (nxt > pend || nxt > F.cs_len) {
                bad = true;
                break;
            }
            /* bit-plane counts of the packet's band lines, into LDS (and the
             * frame's GCLI planes, for the lines below) */
            if (raw) {
                uint32_t base = 0;
                for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
                    const uint32_t c = g.gc[gi], b = g.gb[gi];
                    const JxsBand& bd = g.c[c].b[b];
                    if (l >= bd.lines[type])
                        continue;
                    uint8_t* gl = F.gcli + bd.gcli_off + (size_t)(row * bd.hl + l) * bd.gw;
                    for (uint32_t k = lane; k < bd.gw; k += WAVE) {
                        const uint32_t v = (uint32_t)(cs_bits64(V, g_off * 8 + 4ull * (base + k) + V.sh) >> 60);
                        gl[k] = (uint8_t)v;
                        if (base + k < SL_MAXG)
                            gsh[base + k] = (uint8_t)v;
                    }
                    base += bd.gw;
                }
                if (base > SL_MAXG || (4ull * base + 7) / 8 > lg) /* raw counts fit their sub-packet */
                    bad = true;
            } else if (lane == 0) {
                uint64_t cur = g_off * 8, sbit = s_off * 8;
                const uint64_t gend = d_off * 8;
                uint32_t base = 0;
                for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p] && !bad; gi++) {
                    const uint32_t c = g.gc[gi], b = g.gb[gi];
                    const JxsBand& bd = g.c[c].b[b];
                    if (l >= bd.lines[type])
                        continue;
                    const uint32_t k = c * JXS_NBC + b;
                    const uint32_t md = (T[k >> 1] >> (16 * (k & 1))) & 0xFFFFu, gt = md >> 8;
                    const bool sg = md & 2u, vp = md & 1u;
                    uint8_t* gl = F.gcli + bd.gcli_off + (size_t)(row * bd.hl + l) * bd.gw;
                    const uint8_t* top = nullptr;
                    uint32_t t = gt;
                    if (vp) {
                        uint32_t gtop = gt;
                        if (l > 0) {
                            top = gl - bd.gw;
                        } else {
                            top = gl - (size_t)(bd.hl - bd.lines[0] + 1) * bd.gw; /* last line of the precinct above */
                            const uint32_t* Tt = T - F.pkt_stride;
                            gtop = (Tt[k >> 1] >> (16 * (k & 1) + 8)) & 0xFFu;
                        }
                        t = gtop > gt ? gtop : gt;
                    }
                    if (base + bd.gw > SL_MAXG) {
                        bad = true;
                        break;
                    }
                    uint32_t flag = 0;
                    for (uint32_t i = 0; i < bd.gw; i++) {
                        if (sg && (i & 7) == 0)
                            flag = cs_bits(F, sbit++, 1); /* 1: the group of 8 is insignificant */
                        const uint32_t mt = vp ? max((uint32_t)top[i], t) : 0u;
                        uint32_t v;
                        if (sg && flag) {
                            v = vp && !g.rm ? mt : gt;
                        } else {
                            uint32_t n = 0;
                            for (;;) {
                                const uint32_t x = cs_bits(F, cur, 32);
                                const uint32_t lead = __clz(~x);
                                if (lead < 32) {
                                    n += lead;
                                    cur += lead + 1;
                                    break;
                                }
                                n += 32;
                                cur += 32;
                                if (n >= 32)
                                    break;
                            }
                            /* no code of 32 bits or more (Packing.c:100-107), and
                             * the codes end inside the sub-packet */
                            if (cur > gend || n >= 32) {
                                bad = true;
                                break;
                            }
                            if (!vp) {
                                v = gt + n;
                            } else {
                                const int32_t th = (int32_t)mt - (int32_t)gt > 0 ? (int32_t)mt - (int32_t)gt : 0;
                                int32_t dm = 0;
                                if ((int32_t)n > 2 * th)
                                    dm = (int32_t)n - th;
                                else if (n > 0)
                                    dm = (n & 1) ? -(int32_t)((n + 1) / 2) : (int32_t)(n / 2);
                                v = (uint32_t)((int32_t)mt + dm) & 0xFFu; /* kept as a byte (Packing.c:305) */
                            }
                        }
                        gl[i] = (uint8_t)v;
                        gsh[base + i] = (uint8_t)v;
                    }
                    base += bd.gw;
                }
            }
            if (lane == 0 && bad)
                sh_err = 1;
            __syncthreads();
            bad = sh_err != 0;
            if (bad)
                break;
            /* data sub-packet, band line by band line */
            uint64_t dbit = d_off * 8, nbit = n_off * 8;
            uint32_t base = 0;
            for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
                const uint32_t c = g.gc[gi], b = g.gb[gi];
                const JxsBand& bd = g.c[c].b[b];
                if (l >= bd.lines[type])
                    continue;
                const uint32_t k = c * JXS_NBC + b;
                const uint32_t gt = ((T[k >> 1] >> (16 * (k & 1))) & 0xFFFFu) >> 8;
                int16_t* rowp = F.coef + bd.coeff_off + (size_t)(row * bd.hl + l) * bd.cs;
                uint32_t pad_nz = 0;
                for (uint32_t i0 = 0; i0 < bd.gw; i0 += WAVE) {
                    const uint32_t i = i0 + lane;
                    const uint32_t gc = i < bd.gw ? gsh[base + i] : 0u;
                    const uint32_t np = gc > gt ? gc - gt : 0u;
                    const uint32_t nb = np ? 4 * (np + (g.fs ? 0u : 1u)) : 0u;
                    const uint32_t incl = wave_incl_scan(nb);
                    const uint32_t tot = __shfl(incl, WAVE - 1, WAVE);
                    const uint64_t bp = dbit + incl - nb;
                    uint32_t mag[4] = {0, 0, 0, 0}, sgn = 0;
                    if (nb) {
                        if (bp + nb > n_off * 8) {
                            bad = true;
                        } else {
                            const uint32_t skip = g.fs ? 0u : 4u;
                            const uint64_t v = cs_bits64(V, bp + V.sh);
                            sgn = g.fs ? 0u : (uint32_t)(v >> 60);
                            if (np <= 15) {
#pragma unroll
                                for (int t4 = 0; t4 < 4; t4++)
                                    mag[t4] = ((gather4(v << (skip + t4)) >> (16 - np)) << gt) & 0xFFFFu;
                            } else { /* only in a damaged codestream */
                                group_words(V, bp + skip + V.sh, np, gt, mag);
                            }
                        }
                    }
                    if (i < bd.gw) {
                        for (uint32_t t4 = 0; t4 < 4; t4++) {
                            if (4 * i + t4 >= bd.w) {
                                pad_nz += mag[t4] != 0; /* signs of padding coefficients are skipped */
                                continue;
                            }
                            /* with signs apart, the word waits for its sign (below) */
                            rowp[4 * i + t4] = g.fs ? (int16_t)mag[t4]
                                                    : coef_final(mag[t4] | ((sgn >> (3 - t4)) & 1u) << 15, gc, gt, g.qpih);
                        }
                    }
                    dbit += tot;
                }
                base += bd.gw;
                if (g.fs) {
                    /* the line's signs: one bit per non-zero coefficient */
                    pad_nz = wave_incl_scan(pad_nz);
                    pad_nz = __shfl(pad_nz, WAVE - 1, WAVE);
                    __syncthreads(); /* the coefficients above were written by other lanes */
                    for (uint32_t x0 = 0; x0 < bd.w; x0 += WAVE) {
                        const uint32_t x = x0 + lane;
                        const uint32_t u = x < bd.w ? (uint16_t)rowp[x] : 0u;
                        const uint32_t nz = u != 0 ? 1u : 0u;
                        const uint32_t incl = wave_incl_scan(nz);
                        const uint32_t tot = __shfl(incl, WAVE - 1, WAVE);
                        if (nz) {
                            const uint64_t bp = nbit + incl - 1;
                            uint32_t s1 = 0;
                            if (bp >= nxt * 8)
                                bad = true;
                            else
                                s1 = cs_bits(F, bp, 1);
                            rowp[x] = coef_final(u | s1 << 15, gsh[base - bd.gw + x / 4], gt, g.qpih);
                        }
                        nbit += tot;
                    }
                    nbit += pad_nz;
                }
            }
            if (dbit > n_off * 8 || nbit > nxt * 8)
                bad = true;
            pos = nxt;
            bad = __any(bad);
            __syncthreads();
        }
        if (bad || (uint64_t)pos > pend) {
            if (lane == 0)
                atomicOr(F.status, 1u);
            break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* inverse transform                                                   */
/* ------------------------------------------------------------------ */
/* Linear output scaling (NltDec.c:22-60): Bw-bit coefficients to depth-bit
 * samples, rounded and clamped. */
__device__ __forceinline__ uint32_t nlt_out(int32_t v, uint32_t depth, uint32_t bw) {
    const int32_t dz = (int32_t)bw - (int32_t)depth;
    if (__builtin_expect(v > -(1 << 28) && v < (1 << 28) && bw <= 24, 1)) {
        /* the same in 32 bits: no sum below can leave the int32 range */
        const int32_t y = (v + ((1 << bw) >> 1) + ((1 << dz) >> 1)) >> dz;
        const int32_t m = (1 << depth) - 1;
        return (uint32_t)(y < 0 ? 0 : (y > m ? m : y));
    }
    int64_t x = (int64_t)v + ((1ll << bw) >> 1);
    x = (x + ((1ll << dz) >> 1)) >> dz;
    const int64_t m = (1ll << depth) - 1;
    return (uint32_t)(x < 0 ? 0 : (x > m ? m : x));
}

__device__ __forceinline__ void store_out(void* out, size_t idx, int32_t v, uint32_t depth, uint32_t bw) {
    const uint32_t o = nlt_out(v, depth, bw);
    if (depth <= 8)
        ((uint8_t*)out)[idx] = (uint8_t)o;
    else
        ((uint16_t*)out)[idx] = (uint16_t)o;
}

__device__ __forceinline__ int find_base(const uint32_t* bases, int n, uint32_t blk) {
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

/* MAXW: the widest row of the launch (LDS and the per-thread sample
 * registers are sized for it). */
template <int MAXW>
__global__ void __launch_bounds__(256) k_inv_rows(const JxsInvRowJob* __restrict__ jobs, const uint32_t* __restrict__ bases,
                                                  int njobs) {
    __shared__ __align__(16) int32_t buf[MAXW];
    const JxsInvRowJob& J = jobs[find_base(bases, njobs, blockIdx.x)];
    const uint32_t y = blockIdx.x - J.base;
    uint32_t n[6];
    n[0] = J.w;
    for (uint32_t k = 0; k < J.levels; k++)
        n[k + 1] = (n[k] + 1) / 2;
    const uint32_t L = J.levels;
    /* Every thread issues the loads of all its samples of the row (low band,
     * then each level's high band; at most MAXW / 256 of a band)
     * before it writes any into LDS, so the row costs one memory round trip. */
    const uint32_t tid = threadIdx.x;
    const size_t lfo = (size_t)y * J.lf_stride;
    const int32_t* lf32 = J.lf32 ? J.lf32 + lfo : nullptr;
    const int16_t* lf16 = J.lf16 + (J.lf32 ? 0 : lfo);
    const uint32_t nL = n[L];
    constexpr int NL = (MAXW + 255) / 256;
    int32_t vl[NL];
#pragma unroll
    for (int j = 0; j < NL; j++)
        vl[j] = 0;
    if (lf32) {
#pragma unroll
        for (int j = 0; j < NL; j++)
            if (tid + 256u * j < nL)
                vl[j] = gldr(lf32, tid + 256u * j);
    } else {
#pragma unroll
        for (int j = 0; j < NL; j++)
            if (tid + 256u * j < nL)
                vl[j] = gldr(lf16, tid + 256u * j); /* scaled by 2^Fq when stored */
    }
    /* high band of level k: floor(n[k] / 2) <= MAXW >> (k + 1) samples */
    constexpr int NH0 = ((MAXW >> 1) + 255) / 256, NH1 = ((MAXW >> 2) + 255) / 256,
                  NH2 = ((MAXW >> 3) + 255) / 256, NH3 = ((MAXW >> 4) + 255) / 256,
                  NH4 = ((MAXW >> 5) + 255) / 256;
    int32_t h0[NH0], h1[NH1], h2[NH2], h3[NH3], h4[NH4]; /* full registers: no 16-bit packing waits */
    auto ldh = [&](auto& arr, auto cnt, uint32_t k) {
        constexpr int N = decltype(cnt)::value;
        const uint32_t nh = k < L ? n[k] / 2 : 0u;
        const int16_t* hp = k < L ? J.hf[k] + (size_t)y * J.hf_stride[k] : nullptr;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint32_t i = tid + 256u * j;
            arr[j] = 0;
            if (i < nh)
                arr[j] = (int32_t)gldr(hp, i);
        }
    };
    const uint32_t fq = J.fq; /* LSHIFT32(coefficient, Fq), Decoder.c:789 */
    ldh(h0, std::integral_constant<int, NH0>{}, 0);
    ldh(h1, std::integral_constant<int, NH1>{}, 1);
    ldh(h2, std::integral_constant<int, NH2>{}, 2);
    ldh(h3, std::integral_constant<int, NH3>{}, 3);
    ldh(h4, std::integral_constant<int, NH4>{}, 4);
#pragma unroll
    for (int j = 0; j < NL; j++)
        if (tid + 256u * j < nL)
            buf[(tid + 256u * j) << L] = lf32 ? vl[j] : (int32_t)((uint32_t)vl[j] << fq);
    auto sth = [&](const auto& arr, auto cnt, uint32_t k) {
        constexpr int N = decltype(cnt)::value;
        const uint32_t nh = k < L ? n[k] / 2 : 0u;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint32_t i = tid + 256u * j;
            if (i < nh)
                buf[(2 * i + 1) << k] = (int32_t)((uint32_t)arr[j] << fq);
        }
    };
    sth(h0, std::integral_constant<int, NH0>{}, 0);
    sth(h1, std::integral_constant<int, NH1>{}, 1);
    sth(h2, std::integral_constant<int, NH2>{}, 2);
    sth(h3, std::integral_constant<int, NH3>{}, 3);
    sth(h4, std::integral_constant<int, NH4>{}, 4);
    __syncthreads();
    /* The edge rules are the interior steps with the missing neighbour
     * mirrored: (2x + 2) >> 2 == (x + 1) >> 1 and (2x) >> 1 == x. */
    for (int k = (int)L - 1; k >= 0; k--) {
        const uint32_t nk = n[k], nh = nk / 2, nl = nk - nh;
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
            const uint32_t il = i == 0 ? 1u : 2 * i - 1, ir = i < nh ? 2 * i + 1 : 2 * i - 1;
            buf[(2 * i) << k] -= (buf[il << k] + buf[ir << k] + 2) >> 2;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x) {
            const uint32_t ir = 2 * i + 2 < nk ? 2 * i + 2 : 2 * i;
            buf[(2 * i + 1) << k] += (buf[(2 * i) << k] + buf[ir << k]) >> 1;
        }
        __syncthreads();
    }
    /* four consecutive samples per thread and store where the row allows it */
    const uint32_t w = J.w, n4 = w / 4;
    if (J.out32) {
        int32_t* o = J.out32 + (size_t)y * J.out_stride;
        uint32_t done = 0;
        if (((uintptr_t)o & 15) == 0) {
            for (uint32_t q = threadIdx.x; q < n4; q += blockDim.x)
                *(GAS i32x4*)(o + 4 * q) = *(const i32x4*)&buf[4 * q];
            done = 4 * n4;
        }
        for (uint32_t i = done + threadIdx.x; i < w; i += blockDim.x)
            o[i] = buf[i];
    } else {
        const size_t r0 = (size_t)y * J.out_stride;
        const bool b8 = J.depth <= 8;
        uint32_t done = 0;
        if (b8 ? (((uintptr_t)((uint8_t*)J.out + r0) & 3) == 0) : (((uintptr_t)((uint16_t*)J.out + r0) & 7) == 0)) {
            for (uint32_t q = threadIdx.x; q < n4; q += blockDim.x) {
                const i32x4 v = *(const i32x4*)&buf[4 * q];
                const uint32_t a = nlt_out(v.x, J.depth, J.bw), b = nlt_out(v.y, J.depth, J.bw),
                               c = nlt_out(v.z, J.depth, J.bw), d = nlt_out(v.w, J.depth, J.bw);
                if (b8)
                    *(GAS uint32_t*)((uint8_t*)J.out + r0 + 4 * q) = a | b << 8 | c << 16 | d << 24;
                else
                    *(GAS u32x2*)((uint16_t*)J.out + r0 + 4 * q) = u32x2{a | b << 16, c | d << 16};
            }
            done = 4 * n4;
        }
        for (uint32_t i = done + threadIdx.x; i < w; i += blockDim.x)
            store_out(J.out, r0 + i, buf[i], J.depth, J.bw);
    }
}

/* ------------------------------------------------------------------ */
/* final vertical level with the finest horizontal level                */
/* ------------------------------------------------------------------ */
#define LV_NT 256 /* column pairs of a block */

/* Level-1 inverse horizontal 5/3 at columns 2j and 2j + 1 of a row of width
 * w (Idwt.c semantics, the same edges as the row kernel's finest level):
 * lo(j), lo(j+1) are the low band's samples, hm/h0/hp its high band's at
 * j - 1, j, j + 1 (an index outside the band's range is never used). */
__device__ __forceinline__ void lv_h1(uint32_t j, uint32_t w, int32_t l0, int32_t l1, int32_t hm, int32_t h0, int32_t hp,
                                      int32_t& ev, int32_t& od) {
    const uint32_t nh = w / 2;
    /* even sample 2j, and 2j + 2 for the odd one between them */
    ev = j == 0 ? l0 - ((h0 + 1) >> 1) : (j < nh ? l0 - ((hm + h0 + 2) >> 2) : l0 - ((hm + 1) >> 1));
    od = 0;
    if (j < nh) {
        if (2 * j + 2 < w) {
            const int32_t e1 = j + 1 < nh ? l1 - ((h0 + hp + 2) >> 2) : l1 - ((h0 + 1) >> 1);
            od = h0 + ((ev + e1) >> 1);
        } else {
            od = h0 + ev;
        }
    }
}

/* The inputs of one row pair's horizontal step for column pair j: the low
 * and high bands of low row i and of high row i around j. */
struct LvIn {
    int32_t l0, l1, hm, h0, hp; /* low row */
    int32_t m0, m1, gm, g0, gp; /* high row */
};

/* The band rows of one row pair step: low row i's low samples (int32 rows
 * or a band) and HL band, high row i's LH and HH bands; advanced a row at a
 * time with adds (the row bases stay in scalar registers). */
struct LvRows {
    const int32_t* l32;
    const int16_t* l16;
    const int16_t* hl;
    const int16_t* lh;
    const int16_t* hh;
};

__device__ __forceinline__ LvRows lv_rows(const JxsInvLastJob& J, uint32_t i) {
    LvRows r;
    r.l32 = J.low32 ? J.low32 + (size_t)i * J.low_stride : nullptr;
    r.l16 = J.low32 ? nullptr : J.low16 + (size_t)i * J.low_stride;
    r.hl = J.hl + (size_t)i * J.hl_stride;
    r.lh = J.lh + (size_t)i * J.lh_stride;
    r.hh = J.hh + (size_t)i * J.hh_stride;
    return r;
}

__device__ __forceinline__ void lv_next(LvRows& r, const JxsInvLastJob& J, bool l32) {
    if (l32)
        r.l32 += J.low_stride;
    else
        r.l16 += J.low_stride;
    r.hl += J.hl_stride;
    r.lh += J.lh_stride;
    r.hh += J.hh_stride;
}

/* A thread's columns: low samples j and j1 = j + 1, high samples jm, jh, jp
 * around j, each clamped to its band as lv_h1 expects. */
struct LvCols {
    uint32_t j, j1, jm, jh, jp, fq;
};

__device__ __forceinline__ LvIn lv_load(const LvRows& r, const LvCols& c, bool l32, bool low_row, bool high_row) {
    LvIn v = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto sc = [&](int16_t x) { return (int32_t)((uint32_t)(int32_t)x << c.fq); }; /* LSHIFT32(coefficient, Fq), Decoder.c:789 */
    if (low_row) {
        if (l32) {
            v.l0 = gldr(r.l32, c.j);
            v.l1 = gldr(r.l32, c.j1);
        } else {
            v.l0 = sc(gldr(r.l16, c.j));
            v.l1 = sc(gldr(r.l16, c.j1));
        }
        v.hm = sc(gldr(r.hl, c.jm));
        v.h0 = sc(gldr(r.hl, c.jh));
        v.hp = sc(gldr(r.hl, c.jp));
    }
    if (high_row) {
        v.m0 = sc(gldr(r.lh, c.j));
        v.m1 = sc(gldr(r.lh, c.j1));
        v.gm = sc(gldr(r.hh, c.jm));
        v.g0 = sc(gldr(r.hh, c.jh));
        v.gp = sc(gldr(r.hh, c.jp));
    }
    return v;
}

/* lv_h1 at an interior column pair (1 <= j, j + 1 < w / 2): no edge rule
 * applies. */
__device__ __forceinline__ void lv_h1i(int32_t l0, int32_t l1, int32_t hm, int32_t h0, int32_t hp, int32_t& ev,
                                       int32_t& od) {
    ev = l0 - ((hm + h0 + 2) >> 2);
    od = h0 + ((ev + (l1 - ((h0 + hp + 2) >> 2))) >> 1);
}

/* The output side of a thread: the scaling constants (nlt_out's 32-bit
 * path), the sample size, whether both of its columns exist and whether a
 * pair can go out as one store. */
struct LvOut {
    int32_t add, m;
    uint32_t dz, depth, bw, x, row_bytes;
    bool b8, two, pair, fast;
};

/* Two output samples at columns x, x + 1 of the row at `row` (bytes). */
__device__ __forceinline__ void lv_store2(uint8_t* row, const LvOut& o, int32_t a, int32_t b) {
    uint32_t oa, ob;
    if (__builtin_expect(o.fast && a > -(1 << 28) && a < (1 << 28) && b > -(1 << 28) && b < (1 << 28), 1)) {
        oa = (uint32_t)min(max((a + o.add) >> o.dz, 0), o.m);
        ob = (uint32_t)min(max((b + o.add) >> o.dz, 0), o.m);
    } else {
        oa = nlt_out(a, o.depth, o.bw);
        ob = nlt_out(b, o.depth, o.bw);
    }
    if (o.b8) {
        if (o.pair) {
            gstr(row, o.x, (uint16_t)(oa | ob << 8));
        } else {
            gstr(row, o.x, (uint8_t)oa);
            if (o.two)
                gstr(row, o.x + 1, (uint8_t)ob);
        }
    } else {
        if (o.pair) {
            gstr(row, 2 * o.x, oa | ob << 16);
        } else {
            gstr(row, 2 * o.x, (uint16_t)oa);
            if (o.two)
                gstr(row, 2 * o.x + 2, (uint16_t)ob);
        }
    }
}

/* One thread per output column pair j and a strip of J.seg row pairs of a
 * component's final vertical level: low row L(i) and high row H(i) at
 * columns 2j, 2j + 1 come from their finest horizontal level (lv_h1), and
 * the vertical 5/3 lifting (the steps and edges of k_inv_vert) runs down the
 * strip in registers: the step for output rows 2i, 2i + 1 carries x(2i) and
 * H(i) from the step before and makes L(i + 1), H(i + 1); the next step's
 * loads are issued before this step's arithmetic.  No barrier: neighbours'
 * samples are read again (from cache) instead of exchanged.  Row bases
 * advance by adds and every per-thread constant is set once, so a step is
 * mostly vector work. */
__global__ void __launch_bounds__(LV_NT) k_inv_last(const JxsInvLastJob* __restrict__ jobs,
                                                   const uint32_t* __restrict__ bases, int njobs) {
    const JxsInvLastJob& J = jobs[find_base(bases, njobs, blockIdx.x)];
    const uint32_t t = blockIdx.x - J.base;
    const uint32_t tx = t % J.tiles_x, ty = t / J.tiles_x;
    const uint32_t j = tx * LV_NT + threadIdx.x;
    const uint32_t w = J.w, n1 = (w + 1) / 2, nhw = w / 2;
    if (j >= n1)
        return;
    const uint32_t n = J.h, nh = n / 2, nl = n - nh;
    const uint32_t i0 = ty * J.seg, i1 = min(i0 + J.seg, nl);
    if (i0 >= i1)
        return;
    const bool interior = j >= 1 && j + 1 < nhw; /* no horizontal edge rule at j */
    const bool l32 = J.low32 != nullptr;
    LvCols c;
    c.j = j;
    c.j1 = min(j + 1, n1 - 1);
    c.jm = j > 0 ? j - 1 : 0;
    c.jh = min(j, nhw - 1);
    c.jp = min(j + 1, nhw - 1);
    c.fq = J.fq;
    LvOut o;
    o.depth = J.depth;
    o.bw = J.bw;
    o.dz = J.bw - J.depth;
    o.add = ((1 << J.bw) >> 1) + ((1 << o.dz) >> 1);
    o.m = (1 << J.depth) - 1;
    o.fast = J.bw <= 24;
    o.b8 = J.depth <= 8;
    o.x = 2 * j;
    o.two = 2 * j + 1 < w;
    o.row_bytes = J.out_stride * (o.b8 ? 1u : 2u);
    o.pair = o.two && (((uintptr_t)J.out | o.row_bytes) & (o.b8 ? 1u : 3u)) == 0;
    auto hstep = [&](int32_t l0, int32_t l1, int32_t hm, int32_t h0, int32_t hp, int32_t& ev, int32_t& od) {
        if (interior)
            lv_h1i(l0, l1, hm, h0, hp, ev, od);
        else
            lv_h1(j, w, l0, l1, hm, h0, hp, ev, od);
    };
    /* x(2 i0) and H(i0) */
    int32_t ea, eb, ha, hb; /* even row and high row at columns 2j, 2j + 1 */
    {
        const bool has_h0 = i0 < nh, has_hm = i0 > 0;
        const LvIn a = lv_load(lv_rows(J, i0), c, l32, true, has_h0);
        const LvIn m = has_hm ? lv_load(lv_rows(J, i0 - 1), c, l32, false, true) : LvIn{};
        int32_t la, lb, pa = 0, pb = 0;
        hstep(a.l0, a.l1, a.hm, a.h0, a.hp, la, lb);
        ha = hb = 0;
        if (has_h0)
            hstep(a.m0, a.m1, a.gm, a.g0, a.gp, ha, hb);
        if (has_hm)
            hstep(m.m0, m.m1, m.gm, m.g0, m.gp, pa, pb);
        if (i0 == 0) {
            ea = la - ((ha + 1) >> 1);
            eb = lb - ((hb + 1) >> 1);
        } else if (i0 < nh) {
            ea = la - ((pa + ha + 2) >> 2);
            eb = lb - ((pb + hb + 2) >> 2);
        } else {
            ea = la - ((pa + 1) >> 1);
            eb = lb - ((pb + 1) >> 1);
        }
    }
    LvRows rn = lv_rows(J, i0 + 1); /* the rows the next loads read */
    LvIn nx = {};
    if (i0 < nh && 2 * i0 + 2 < n)
        nx = lv_load(rn, c, l32, true, i0 + 1 < nh);
    lv_next(rn, J, l32);
    uint8_t* orow = (uint8_t*)J.out + (size_t)(2 * i0) * o.row_bytes; /* output row 2i */
    for (uint32_t i = i0; i < i1; i++) {
        const bool has_odd = i < nh;
        const bool need_next = has_odd && 2 * i + 2 < n;
        const bool has_hn = i + 1 < nh;
        const LvIn cur = nx;
        /* the loads of the step after this one (rows i + 2), issued before its arithmetic */
        if (i + 1 < i1 && i + 1 < nh && 2 * i + 4 < n)
            nx = lv_load(rn, c, l32, true, i + 2 < nh);
        lv_next(rn, J, l32);
        int32_t oa = 0, ob = 0, na = 0, nb = 0;
        if (need_next) {
            int32_t la, lb, ga = 0, gb = 0;
            hstep(cur.l0, cur.l1, cur.hm, cur.h0, cur.hp, la, lb);
            if (has_hn)
                hstep(cur.m0, cur.m1, cur.gm, cur.g0, cur.gp, ga, gb);
            na = has_hn ? la - ((ha + ga + 2) >> 2) : la - ((ha + 1) >> 1);
            nb = has_hn ? lb - ((hb + gb + 2) >> 2) : lb - ((hb + 1) >> 1);
            oa = ha + ((ea + na) >> 1);
            ob = hb + ((eb + nb) >> 1);
            ha = ga;
            hb = gb;
        } else if (has_odd) {
            oa = ha + ea;
            ob = hb + eb;
        }
        lv_store2(orow, o, ea, eb);
        if (has_odd)
            lv_store2(orow + o.row_bytes, o, oa, ob);
        orow += 2 * (size_t)o.row_bytes;
        ea = na;
        eb = nb;
    }
}

#define IVSEG 16

__global__ void __launch_bounds__(256) k_inv_vert(const JxsInvVertJob* __restrict__ jobs, const uint32_t* __restrict__ bases,
                                                  int njobs) {
    const JxsInvVertJob J = jobs[find_base(bases, njobs, blockIdx.x)];
    const uint32_t t = blockIdx.x - J.base;
    const uint32_t tx = t % J.tiles_x, ty = t / J.tiles_x;
    const uint32_t x = tx * blockDim.x + threadIdx.x;
    if (x >= J.w)
        return;
    const uint32_t n = J.h, nh = n / 2, nl = n - nh;
    const int32_t* Lp = J.src;
    const int32_t* Hp = J.src + (size_t)nl * J.w;
    auto Lv = [&](uint32_t i) { return Lp[(size_t)i * J.w + x]; };
    auto Hv = [&](uint32_t i) { return Hp[(size_t)i * J.w + x]; };
    auto even = [&](uint32_t i) -> int32_t {
        if (i == 0)
            return Lv(0) - ((Hv(0) + 1) >> 1);
        if (i < nh)
            return Lv(i) - ((Hv(i - 1) + Hv(i) + 2) >> 2);
        return Lv(i) - ((Hv(i - 1) + 1) >> 1);
    };
    const uint32_t i0 = ty * IVSEG, i1 = min(i0 + IVSEG, nl);
    int32_t e = even(i0);
    for (uint32_t i = i0; i < i1; i++) {
        int32_t e_next = 0, odd = 0;
        const bool has_odd = i < nh;
        if (has_odd) {
            if (2 * i + 2 < n) {
                e_next = even(i + 1);
                odd = Hv(i) + ((e + e_next) >> 1);
            } else {
                odd = Hv(i) + e;
            }
        }
        if (J.out32) {
            J.out32[(size_t)(2 * i) * J.w + x] = e;
            if (has_odd)
                J.out32[(size_t)(2 * i + 1) * J.w + x] = odd;
        } else {
            store_out(J.out, (size_t)(2 * i) * J.out_stride + x, e, J.depth, J.bw);
            if (has_odd)
                store_out(J.out, (size_t)(2 * i + 1) * J.out_stride + x, odd, J.depth, J.bw);
        }
        e = e_next;
    }
}

extern "C" {

hipError_t jxs_launch_dec_scan(const JxsDecFrame* frames, int nfr, hipStream_t st) {
    hipLaunchKernelGGL(k_dec_scan, dim3(nfr), dim3(256), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_dec_packets(const JxsDecFrame* frames, int nfr, uint32_t precincts, hipStream_t st) {
    if (!precincts)
        return hipSuccess;
    hipLaunchKernelGGL(k_dec_packets, dim3((precincts + 63) / 64), dim3(64), 0, st, frames, nfr, precincts);
    return hipGetLastError();
}

hipError_t jxs_launch_dec_unpack(const JxsDecFrame* frames, int nfr, uint32_t max_npk, uint32_t precincts, hipStream_t st) {
    if (!precincts)
        return hipSuccess;
    hipLaunchKernelGGL(k_dec_unpack, dim3(max_npk, precincts), dim3(PK_THREADS), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_dec_slice(const JxsDecFrame* frames, int nfr, uint32_t slices, hipStream_t st) {
    if (!slices)
        return hipSuccess;
    hipLaunchKernelGGL(k_dec_slice, dim3(slices), dim3(WAVE), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_inv_rows(const JxsInvRowJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, uint32_t maxw,
                               hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    if (maxw <= 1024)
        hipLaunchKernelGGL(k_inv_rows<1024>, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    else if (maxw <= 2048)
        hipLaunchKernelGGL(k_inv_rows<2048>, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    else if (maxw <= 4096)
        hipLaunchKernelGGL(k_inv_rows<4096>, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    else if (maxw <= 7680)
        hipLaunchKernelGGL(k_inv_rows<7680>, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    else if (maxw <= JXS_MAX_ROW)
        hipLaunchKernelGGL(k_inv_rows<JXS_MAX_ROW>, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t jxs_launch_inv_last(const JxsInvLastJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks,
                              hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    hipLaunchKernelGGL(k_inv_last, dim3(blocks), dim3(LV_NT), 0, st, jobs, bases, njobs);
    return hipGetLastError();
}

hipError_t jxs_launch_inv_vert(const JxsInvVertJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    hipLaunchKernelGGL(k_inv_vert, dim3(blocks), dim3(256), 0, st, jobs, bases, njobs);
    return hipGetLastError();
}

}  // extern "C"
