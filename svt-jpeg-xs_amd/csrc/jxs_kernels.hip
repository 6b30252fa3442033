/*
 * jxs_kernels.hip -- gfx950 kernels of the JPEG XS encode path.
 *
 *   k_dwt      one block per output row pair of one vertical 5/3 level: the
 *              low and the high row are lifted vertically from the input
 *              (with the input scaling of NltEnc.c:33-71 on the fly for the
 *              first level) straight into LDS, then through their horizontal
 *              levels in place; every band produced is
 *              written as 16-bit sign-magnitude coefficients (image_shift,
 *              NltEnc.c:12-31) together with its GCLI per 4 coefficients and
 *              the maximum per 8 GCLIs (GcStageProcess.c:22-97), using wave
 *              shuffles for the maxima, and the histogram record of the band
 *              line (GCLI and significance-maximum counts, from wave ballots)
 *              the rate control needs (RateControl.c:18-154).
 *   k_rc_table one block per precinct: from the histogram records, the
 *              packet-level byte count of the precinct for every (Q, R)
 *              (RateControl.c:457-781).
 *   k_select   one wave per slice: precinct budgets with the +10 % first
 *              precinct and the padding carry (PackStageProcess.c:600-684),
 *              the reference's binary searches replayed on the LDS-staged
 *              tables (RateControl.c:889-1083, BinarySearch.c), precinct
 *              offsets.
 *   k_plan     one wave per precinct: per-band truncation and coding method,
 *              packet sizes and offsets, slice / precinct headers, padding,
 *              end marker and frame header.
 *   k_pack     one wave per band line of a precinct: its significance,
 *              bit-plane-count and data segments (PackPrecinct.c:235-572)
 *              with quantisation (Quant.c:9-45) fused, each at the bit
 *              offset the plan gives it, plus the packet header for a
 *              packet's first band line; bit offsets inside a segment from
 *              wave prefix sums, bits assembled in per-wave LDS rings and
 *              written out as dwords.
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
/* Global-memory access through address-space-1 pointers: the buffers are
 * reached through pointers held in descriptors, which the compiler would
 * otherwise address with flat instructions (slower, and counted against the
 * LDS wait counter). */
#define GAS __attribute__((address_space(1)))
/* Read-only descriptors (geometry, plans) written before the kernel starts:
 * constant address space, so the compiler keeps them in scalar loads even
 * after the kernel's own global stores. */
#define CAS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const GAS T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(GAS T*)p = v;
}
/* 8-byte aligned pairs of words */
__device__ __forceinline__ uint2 gld2(const void* p) {
    const uint64_t v = *(const GAS uint64_t*)p;
    return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}
__device__ __forceinline__ void gst2(void* p, uint2 v) { *(GAS uint64_t*)p = ((uint64_t)v.y << 32) | v.x; }

__device__ __forceinline__ int32_t load_src(const void* src, uint32_t kind, size_t idx, uint32_t shift, uint32_t mask,
                                            int32_t offset) {
    if (kind == JXS_SRC_I32)
        return gld((const int32_t*)src + idx);
    uint32_t v = kind == JXS_SRC_U8 ? (uint32_t)gld((const uint8_t*)src + idx) : (uint32_t)gld((const uint16_t*)src + idx);
    return (int32_t)((v & mask) << shift) - offset;
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

/* Workgroups are dealt to the 8 XCDs round robin.  This maps the dispatch
 * index to a logical block so that each XCD runs one contiguous range of
 * rows: neighbouring row pairs, which read overlapping input rows, then
 * share that XCD's L2. */
#define NXCD 8
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t q = n / NXCD, r = n % NXCD, x = b % NXCD;
    return x * q + min(x, r) + b / NXCD;
}

#define HMAX JXS_MAX_ROW

/* Writes `count` values buf[start + k * step] as one band row (row stride
 * o.cs, the padding of the last group zero) with its GCLIs and significance
 * maxima, and the row's histogram record when o.hist is set.  hcnt is a
 * zeroed LDS array of 32 words; it is zero again on return.  Called by all
 * threads of the block.  Straight-line per group: the four LDS reads use
 * clamped indexes and are issued together, the sign-magnitude conversion
 * is branch-free, GCLI counts come from wave ballots (lanes 0..15 count one
 * value each) and significance-maximum counts from LDS atomics (one lane in
 * eight holds a maximum). */
__device__ __forceinline__ void emit_band_row(const JxsBandOut o, uint32_t y, const int32_t* buf, uint32_t start,
                                              uint32_t step, uint32_t count, uint32_t* hcnt) {
    uint16_t* crow = o.coeff + (size_t)y * o.cs;
    uint8_t* grow = o.gcli + (size_t)y * o.gw;
    uint8_t* srow = o.smax + (size_t)y * o.sw;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t full = o.gw / 8;
    const bool hist = o.hist != nullptr;
    /* per-lane masks: ~(ballot of GCLI bit k XOR mk) has a one for every
     * lane whose GCLI bit k equals bit k of `lane` */
    const uint64_t m0 = (lane & 1u) ? ~0ull : 0ull, m1 = (lane & 2u) ? ~0ull : 0ull;
    const uint64_t m2 = (lane & 4u) ? ~0ull : 0ull, m3 = (lane & 8u) ? ~0ull : 0ull;
    uint32_t mycnt = 0;
    for (uint32_t base = 0; base < o.gw; base += blockDim.x) {
        const uint32_t gidx = base + threadIdx.x;
        const bool valid = gidx < o.gw;
        const uint32_t i0 = 4 * min(gidx, o.gw - 1);
        int32_t x[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            x[k] = buf[start + min(i0 + k, count - 1) * step];
        uint32_t v[4], orv = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t a = (uint32_t)(x[k] < 0 ? -x[k] : x[k]);
            const uint32_t mg = (a + 128) >> 8;
            v[k] = (i0 + k < count && mg) ? (mg | (((uint32_t)x[k] >> 31) << 15)) : 0u;
            orv |= v[k];
        }
        const uint32_t gc = bitlen(orv & 0x7FFFu);
        if (valid) {
            gst2(crow + 4 * gidx, make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16)));
            gst(grow + gidx, (uint8_t)gc);
        }
        uint32_t m = valid ? gc : 0u;
        m = max(m, (uint32_t)__shfl_xor((int)m, 1, WAVE));
        m = max(m, (uint32_t)__shfl_xor((int)m, 2, WAVE));
        m = max(m, (uint32_t)__shfl_xor((int)m, 4, WAVE));
        const bool head = valid && (gidx & 7) == 0;
        if (head)
            gst(srow + gidx / 8, (uint8_t)m);
        if (hist) {
            const uint64_t vg = __ballot(valid);
            const uint64_t g0 = __ballot(gc & 1u), g1 = __ballot(gc & 2u), g2 = __ballot(gc & 4u),
                           g3 = __ballot(gc & 8u);
            mycnt += (uint32_t)__popcll(vg & ~(g0 ^ m0) & ~(g1 ^ m1) & ~(g2 ^ m2) & ~(g3 ^ m3));
            if (head) {
                if (gidx / 8 >= full)
                    hcnt[HREC_TAIL] = m + 1; /* the partial significance group */
                else if (m < 15)
                    atomicAdd(&hcnt[16 + m], 1u); /* maximum 15: implied by the group count */
            }
        }
    }
    if (!hist)
        return;
    if (lane < 16 && mycnt)
        atomicAdd(&hcnt[lane], mycnt);
    __syncthreads();
    if (threadIdx.x < JXS_HREC) {
        const uint32_t t = threadIdx.x, c = hcnt[t];
        uint16_t* rec = o.hist + ((size_t)(y / o.hl) * o.n_bl + o.bl + y % o.hl) * JXS_HREC;
        gst(rec + t, (uint16_t)(t == HREC_TAIL ? (c ? c - 1 : 0xFFFFu) : c));
        hcnt[t] = 0;
    }
    __syncthreads();
}

/* Field-wise copy of a band output descriptor out of constant memory. */
__device__ __forceinline__ JxsBandOut band_out(const CAS JxsBandOut& b) {
    JxsBandOut o;
    o.coeff = b.coeff;
    o.gcli = b.gcli;
    o.smax = b.smax;
    o.hist = b.hist;
    o.w = b.w;
    o.gw = b.gw;
    o.sw = b.sw;
    o.cs = b.cs;
    o.hl = b.hl;
    o.bl = b.bl;
    o.n_bl = b.n_bl;
    o.pad_ = 0;
    return o;
}

/* Lifts the row of w samples in buf through `levels` horizontal 5/3 levels
 * in place (the low-pass samples of level k at stride 2^k) and writes the
 * high band of every level and the final low-pass output (a band, or int32
 * samples at lf_i32).  buf is loaded and synchronised by the caller. */
__device__ __forceinline__ void horz_emit(int32_t* buf, uint32_t w, uint32_t levels, const CAS JxsBandOut* hf,
                                          uint32_t lf_is_band, const CAS JxsBandOut& lf, int32_t* lf_i32, uint32_t y,
                                          uint32_t* hcnt) {
    uint32_t n = w, s = 1;
    for (uint32_t lvl = 0; lvl < levels; lvl++) {
        const uint32_t nh = n / 2, nl = n - nh;
        /* predict: hf(i) = x(2i+1) - ((x(2i) + x(2i+2)) >> 1); at an even
         * end x(2i+2) is replaced by x(2i), which gives x(2i+1) - x(2i) */
        for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x) {
            const uint32_t e = 2 * i + 2 < n ? 2 * i + 2 : 2 * i;
            buf[(2 * i + 1) * s] -= (buf[2 * i * s] + buf[e * s]) >> 1;
        }
        __syncthreads();
        /* update: lf(i) = x(2i) + ((hf(i-1) + hf(i) + 2) >> 2); a missing
         * neighbour is replaced by the other one, which gives the one-sided
         * (hf + 1) >> 1 at both ends (n >= 2 at every level) */
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
            const uint32_t l = i > 0 ? 2 * i - 1 : 2 * i + 1, r = i < nh ? 2 * i + 1 : 2 * i - 1;
            buf[2 * i * s] += (buf[l * s] + buf[r * s] + 2) >> 2;
        }
        __syncthreads();
        emit_band_row(band_out(hf[lvl]), y, buf, s, 2 * s, nh, hcnt);
        s *= 2;
        n = nl;
    }
    if (lf_is_band) {
        emit_band_row(band_out(lf), y, buf, 0, s, n, hcnt);
    } else {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            gst(lf_i32 + i, buf[i * s]);
    }
}

/* Sample access for the transform, specialised per input kind so that the
 * batched loads of a block are straight-line code.  Input samples are
 * scaled as NltEnc.c:33-71 does: x = (v & mask) << shift - offset; the
 * int32 low-pass planes of the second level are read as they are. */
struct Scale {
    uint32_t shift, mask;
    int32_t offset;
    __device__ __forceinline__ int32_t operator()(uint32_t v) const { return (int32_t)((v & mask) << shift) - offset; }
};

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int KIND>
__device__ __forceinline__ int32_t ld1(const void* row, uint32_t x, uint32_t step, const Scale& sc) {
    if (KIND == JXS_SRC_U16)
        return sc(gld((const uint16_t*)row + (size_t)x * step));
    if (KIND == JXS_SRC_U8)
        return sc(gld((const uint8_t*)row + (size_t)x * step));
    return gld((const int32_t*)row + (size_t)x * step);
}

/* columns x .. x+3 of a row with one load (contiguous rows, x a multiple
 * of 4, 16-byte aligned plane and stride: checked by the host) */
template <int KIND>
__device__ __forceinline__ int4 ld4(const void* row, uint32_t x, const Scale& sc) {
    if (KIND == JXS_SRC_U16) {
        const u32x2 v = *(const GAS u32x2*)((const uint16_t*)row + x);
        return make_int4(sc(v.x & 0xFFFFu), sc(v.x >> 16), sc(v.y & 0xFFFFu), sc(v.y >> 16));
    }
    if (KIND == JXS_SRC_U8) {
        const uint32_t v = *(const GAS uint32_t*)((const uint8_t*)row + x);
        return make_int4(sc(v & 0xFFu), sc((v >> 8) & 0xFFu), sc((v >> 16) & 0xFFu), sc(v >> 24));
    }
    const u32x4 v = *(const GAS u32x4*)((const int32_t*)row + x);
    return make_int4((int32_t)v.x, (int32_t)v.y, (int32_t)v.z, (int32_t)v.w);
}

/* The vertical 5/3 step of one column from its input samples x(2i-2) ..
 * x(2i+2) (DWT of CodeDeprecated.cc:84-127 / Dwt.c:337-398 semantics):
 *   hf(k) = x(2k+1) - ((x(2k) + x(2k+2)) >> 1), or x(2k+1) - x(2k) at an even end
 *   lf(k) = x(2k) + ((hf(k-1) + hf(k) + 2) >> 2), one-sided at the ends */
struct VertFlags {
    bool prev, h, next; /* i > 0, i < n/2, 2i+2 < n */
};
__device__ __forceinline__ void vert53(const VertFlags& f, int32_t xm2, int32_t xm1, int32_t xc, int32_t xp1, int32_t xp2,
                                       int32_t& lf, int32_t& hf) {
    const int32_t hcur = f.h ? (f.next ? xp1 - ((xc + xp2) >> 1) : xp1 - xc) : 0;
    const int32_t hprev = f.prev ? xm1 - ((xm2 + xc) >> 1) : 0;
    if (!f.prev)
        lf = xc + ((hcur + 1) >> 1);
    else if (f.h)
        lf = xc + ((hprev + hcur + 2) >> 2);
    else
        lf = xc + ((hprev + 1) >> 1);
    hf = hcur;
}

#define QUAD_BATCH 4 /* column quads per thread whose loads are issued together */
#define FILL_BATCH 4 /* columns per thread whose loads are issued together */

/* Low row i into lo[] and, when it exists, high row i into hi[] for the w
 * columns, from the five input rows r[0..4] = 2i-2 .. 2i+2 (clamped into
 * the plane, so every load is in bounds; the flags drop what lies
 * outside).  The loads of a whole batch are issued before any is used. */
template <int KIND, bool VEC>
__device__ __forceinline__ void vert_fill(const void* const r[5], const VertFlags& f, uint32_t w, uint32_t step,
                                          const Scale& sc, int32_t* lo, int32_t* hi) {
    if (VEC) {
        const uint32_t nq = w / 4;
        for (uint32_t q0 = 0; q0 < nq; q0 += QUAD_BATCH * blockDim.x) {
            int4 v[5][QUAD_BATCH];
#pragma unroll
            for (uint32_t k = 0; k < QUAD_BATCH; k++) {
                const uint32_t x = 4 * min(q0 + k * blockDim.x + threadIdx.x, nq - 1);
#pragma unroll
                for (uint32_t j = 0; j < 5; j++)
                    v[j][k] = ld4<KIND>(r[j], x, sc);
            }
#pragma unroll
            for (uint32_t k = 0; k < QUAD_BATCH; k++) {
                const uint32_t q = q0 + k * blockDim.x + threadIdx.x;
                if (q >= nq)
                    break;
                int4 l, h;
                vert53(f, v[0][k].x, v[1][k].x, v[2][k].x, v[3][k].x, v[4][k].x, l.x, h.x);
                vert53(f, v[0][k].y, v[1][k].y, v[2][k].y, v[3][k].y, v[4][k].y, l.y, h.y);
                vert53(f, v[0][k].z, v[1][k].z, v[2][k].z, v[3][k].z, v[4][k].z, l.z, h.z);
                vert53(f, v[0][k].w, v[1][k].w, v[2][k].w, v[3][k].w, v[4][k].w, l.w, h.w);
                *(int4*)(lo + 4 * q) = l;
                if (f.h)
                    *(int4*)(hi + 4 * q) = h;
            }
        }
    } else {
        for (uint32_t x0 = 0; x0 < w; x0 += FILL_BATCH * blockDim.x) {
            int32_t v[5][FILL_BATCH];
#pragma unroll
            for (uint32_t k = 0; k < FILL_BATCH; k++) {
                const uint32_t x = min(x0 + k * blockDim.x + threadIdx.x, w - 1);
#pragma unroll
                for (uint32_t j = 0; j < 5; j++)
                    v[j][k] = ld1<KIND>(r[j], x, step, sc);
            }
#pragma unroll
            for (uint32_t k = 0; k < FILL_BATCH; k++) {
                const uint32_t x = x0 + k * blockDim.x + threadIdx.x;
                if (x >= w)
                    break;
                int32_t l, h;
                vert53(f, v[0][k], v[1][k], v[2][k], v[3][k], v[4][k], l, h);
                lo[x] = l;
                if (f.h)
                    hi[x] = h;
            }
        }
    }
}

/* Input row i into lo[] (components without vertical levels). */
template <int KIND, bool VEC>
__device__ __forceinline__ void row_fill(const void* row, uint32_t w, uint32_t step, const Scale& sc, int32_t* lo) {
    if (VEC) {
        const uint32_t nq = w / 4;
        for (uint32_t q0 = 0; q0 < nq; q0 += QUAD_BATCH * blockDim.x) {
            int4 v[QUAD_BATCH];
#pragma unroll
            for (uint32_t k = 0; k < QUAD_BATCH; k++)
                v[k] = ld4<KIND>(row, 4 * min(q0 + k * blockDim.x + threadIdx.x, nq - 1), sc);
#pragma unroll
            for (uint32_t k = 0; k < QUAD_BATCH; k++) {
                const uint32_t q = q0 + k * blockDim.x + threadIdx.x;
                if (q < nq)
                    *(int4*)(lo + 4 * q) = v[k];
            }
        }
    } else {
        for (uint32_t x0 = 0; x0 < w; x0 += FILL_BATCH * blockDim.x) {
            int32_t v[FILL_BATCH];
#pragma unroll
            for (uint32_t k = 0; k < FILL_BATCH; k++)
                v[k] = ld1<KIND>(row, min(x0 + k * blockDim.x + threadIdx.x, w - 1), step, sc);
#pragma unroll
            for (uint32_t k = 0; k < FILL_BATCH; k++) {
                const uint32_t x = x0 + k * blockDim.x + threadIdx.x;
                if (x < w)
                    lo[x] = v[k];
            }
        }
    }
}

/* Element size of an input kind. */
__device__ __forceinline__ uint32_t kind_bytes(uint32_t kind) {
    return kind == JXS_SRC_U8 ? 1u : kind == JXS_SRC_U16 ? 2u : 4u;
}

/* The work of one block of k_dwt for one input kind. */
template <int KIND, bool VEC>
__device__ __forceinline__ void dwt_block(const CAS JxsDwtJob& J, uint32_t i, int32_t* lo, int32_t* hi, uint32_t* hcnt) {
    const uint32_t w = J.w, n = J.h, step = J.step;
    const Scale sc = {J.shift, J.mask, J.offset};
    const uint8_t* base = (const uint8_t*)J.src;
    const size_t pitch = (size_t)J.stride * kind_bytes(KIND);
    if (J.vert) {
        const VertFlags f = {i > 0, i < n / 2, 2 * i + 2 < n};
        const void* r[5];
#pragma unroll
        for (int j = 0; j < 5; j++) {
            const int row = min(max((int)(2 * i) + j - 2, 0), (int)n - 1);
            r[j] = base + (size_t)row * pitch;
        }
        vert_fill<KIND, VEC>(r, f, w, step, sc, lo, hi);
    } else {
        row_fill<KIND, VEC>(base + (size_t)i * pitch, w, step, sc, lo);
    }
    __syncthreads();
    horz_emit(lo, w, J.lev_l, J.hf, J.lf_is_band, J.lf, J.lf_is_band ? nullptr : J.lf_i32 + (size_t)i * J.lf_stride,
              i, hcnt);
    if (J.vert && i < n / 2) {
        __syncthreads();
        horz_emit(hi, w, 1, &J.hh, 1, J.lh, nullptr, i, hcnt);
    }
}

/* One block per output row (pair).  With a vertical level, block i computes
 * low row i and high row i from input rows 2i-2 .. 2i+2 straight into LDS,
 * lifts the low row through the horizontal levels of the low rows and
 * writes its bands (or the int32 low-pass plane of the next vertical level),
 * then the high row through one horizontal level into the LH and HH bands.
 * The intermediate rows never leave LDS.  Without a vertical level, block i
 * lifts input row i.  Dynamic LDS: two rows of the widest job. */
__global__ void __launch_bounds__(256) k_dwt(const JxsDwtJob* __restrict__ jobs, const uint32_t* __restrict__ bases,
                                             int njobs, uint32_t nblocks, uint32_t wmax) {
    extern __shared__ __align__(16) int32_t rows_lds[];
    __shared__ uint32_t hcnt[JXS_HREC];
    int32_t* lo = rows_lds;
    int32_t* hi = rows_lds + ((wmax + 3) & ~3u); /* 16-byte aligned */
    const uint32_t blk = xcd_remap(blockIdx.x, nblocks);
    const int j = find_job_u32(bases, njobs, blk);
    const CAS JxsDwtJob& J = ((const CAS JxsDwtJob*)jobs)[j];
    const uint32_t i = blk - J.base;
    if (threadIdx.x < JXS_HREC)
        hcnt[threadIdx.x] = 0;
    const uint32_t kind = J.kind;
    const bool vec = J.vec != 0;
    if (kind == JXS_SRC_U16) {
        if (vec)
            dwt_block<JXS_SRC_U16, true>(J, i, lo, hi, hcnt);
        else
            dwt_block<JXS_SRC_U16, false>(J, i, lo, hi, hcnt);
    } else if (kind == JXS_SRC_U8) {
        if (vec)
            dwt_block<JXS_SRC_U8, true>(J, i, lo, hi, hcnt);
        else
            dwt_block<JXS_SRC_U8, false>(J, i, lo, hi, hcnt);
    } else {
        if (vec)
            dwt_block<JXS_SRC_I32, true>(J, i, lo, hi, hcnt);
        else
            dwt_block<JXS_SRC_I32, false>(J, i, lo, hi, hcnt);
    }
}

/* Copies a frame's geometry into LDS (all threads of the block take part;
 * the caller synchronises).  The rate-control kernels index it with
 * data-dependent band and packet numbers. */
__device__ __forceinline__ void stage_geom(JxsGeom* dst, const JxsGeom* src) {
    const uint4* s = (const uint4*)src;
    uint4* d = (uint4*)dst;
    constexpr uint32_t n = sizeof(JxsGeom) / 16;
    /* blocks of 64 threads or more: at most 5 copies per thread, loaded
     * together */
    const uint32_t t = threadIdx.x, bd = blockDim.x;
    /* indexes clamped into the struct: every load is in bounds, only the
     * stores are predicated */
    const uint4 v0 = s[min(t, n - 1)];
    const uint4 v1 = s[min(t + bd, n - 1)];
    const uint4 v2 = s[min(t + 2 * bd, n - 1)];
    const uint4 v3 = s[min(t + 3 * bd, n - 1)];
    const uint4 v4 = s[min(t + 4 * bd, n - 1)];
    if (t < n)
        d[t] = v0;
    if (t + bd < n)
        d[t + bd] = v1;
    if (t + 2 * bd < n)
        d[t + 2 * bd] = v2;
    if (t + 3 * bd < n)
        d[t + 3 * bd] = v3;
    if (t + 4 * bd < n)
        d[t + 4 * bd] = v4;
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

/* Bit counts of one band line at truncation t from its histogram record:
 * the GCLI sub-packet without significance (one unary code per group,
 * RateControl.c:18-63), the groups a significance flag removes
 * (RateControl.c:65-154) and the data bit-planes with signs
 * (RateControl.c:457-520). */
struct LineBits {
    uint32_t ns, zeroed, data;
};
__device__ __forceinline__ LineBits line_bits(const uint16_t* rec, uint32_t t, uint32_t gw) {
    uint32_t cnt = 0, data = 0, sig = 0;
#pragma unroll
    for (uint32_t v = 0; v < 16; v++) {
        const uint32_t h = rec[v];
        cnt += v <= t ? h : 0u;
        data += v > t ? (v - t + 1) * h : 0u;
    }
    const uint32_t full = gw / 8;
    if (t >= 15) {
        sig = full;
    } else {
#pragma unroll
        for (uint32_t v = 0; v < 15; v++)
            sig += v <= t ? (uint32_t)rec[16 + v] : 0u;
    }
    const uint32_t tail = rec[HREC_TAIL];
    LineBits b;
    b.ns = data + cnt;
    b.zeroed = sig * 8 + ((tail != 0xFFFFu && tail <= t) ? gw - full * 8 : 0u);
    b.data = 4 * data;
    return b;
}

/* One block per precinct: the packet-level byte count of the precinct for
 * every (Q, R) (precinct_get_budget_bytes with the per-band method choice of
 * rate_control_calculate_band_best_method, RateControl.c:457-781), from the
 * band-line histograms the transform wrote.  Per band line and truncation the
 * bit counts are tabulated once in LDS; each (Q, R) entry then sums the
 * packets' bands. */
__global__ void __launch_bounds__(256) k_rc_table(const JxsFrameDesc* __restrict__ frames, int nfr) {
    __shared__ uint16_t rec[JXS_NBL][JXS_HREC];
    __shared__ uint32_t t_ns[JXS_NBL][16], t_zr[JXS_NBL][16], t_d[JXS_NBL][16], t_g[JXS_NBL][16], t_s[JXS_NBL][16];
    __shared__ uint8_t bl_band[JXS_NBL];
    __shared__ uint32_t b_gain[JXS_NGB], b_prio[JXS_NGB], b_bl[JXS_NGB], b_nl[JXS_NGB], b_gw[JXS_NGB], b_sw[JXS_NGB];
    const int f = find_frame(frames, nfr, blockIdx.x);
    const JxsFrameDesc& F = frames[f];
    __shared__ __align__(16) JxsGeom sg;
    stage_geom(&sg, F.g);
    __syncthreads();
    const JxsGeom& g = sg;
    const uint32_t row = blockIdx.x - F.prec_base;
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    const uint32_t nbl = g.n_bl;
    for (uint32_t gi = threadIdx.x; gi < g.nb_all; gi += blockDim.x) {
        uint32_t nl = 0;
        if (g.gb[gi] != 0xFF) {
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            nl = bd.lines[type];
            b_gain[gi] = bd.gain;
            b_prio[gi] = bd.prio;
            b_bl[gi] = bd.bl;
            b_gw[gi] = bd.gw;
            b_sw[gi] = bd.sw;
            for (uint32_t k = 0; k < bd.hl; k++)
                bl_band[bd.bl + k] = (uint8_t)gi;
        }
        b_nl[gi] = nl;
    }
    const uint32_t* src = (const uint32_t*)(F.hist + (size_t)row * nbl * JXS_HREC);
    {
        /* nbl * 16 <= 1024 words: four loads per thread, issued together */
        uint32_t v[4];
        const uint32_t words = nbl * JXS_HREC / 2;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t i = j * blockDim.x + threadIdx.x;
            v[j] = gld(src + min(i, words - 1)); /* clamped: every load in bounds */
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t i = j * blockDim.x + threadIdx.x;
            if (i < words)
                ((uint32_t*)&rec[0][0])[i] = v[j];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbl * 16; i += blockDim.x) {
        const uint32_t bl = i >> 4, t = i & 15;
        const LineBits b = line_bits(rec[bl], t, b_gw[bl_band[bl]]);
        t_ns[bl][t] = b.ns;
        t_zr[bl][t] = b.zeroed;
        t_d[bl][t] = b.data;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < g.nb_all * 16; i += blockDim.x) {
        const uint32_t gi = i >> 4, t = i & 15, nl = b_nl[gi];
        if (nl == 0)
            continue;
        const uint32_t bl = b_bl[gi], sw = b_sw[gi];
        uint32_t nosig = 0, sig = sw * nl;
        for (uint32_t k = 0; k < nl; k++) {
            nosig += t_ns[bl + k][t];
            sig += t_ns[bl + k][t] - t_zr[bl + k][t];
        }
        const bool use = g.sig && nosig > sig;
        for (uint32_t k = 0; k < nl; k++) {
            t_g[bl + k][t] = use ? t_ns[bl + k][t] - t_zr[bl + k][t] : t_ns[bl + k][t];
            t_s[bl + k][t] = use ? sw : 0u;
        }
    }
    __syncthreads();
    uint32_t* tab = F.table + (size_t)row * JXS_NQ * JXS_NR;
    const uint32_t nq = g.maxq + 1, nr = g.maxr + 1, npk = g.npk;
    for (uint32_t idx = threadIdx.x; idx < nq * nr; idx += blockDim.x) {
        const uint32_t q = idx / nr, r = idx % nr;
        uint32_t total = 0;
        for (uint32_t p = 0; p < npk; p++) {
            const uint32_t l = g.pk_line[p];
            uint32_t d = 0, gb = 0, sb = 0;
            for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
                if (l >= b_nl[gi])
                    continue;
                const uint32_t gt = truncation(b_gain[gi], b_prio[gi], q, r), bl = b_bl[gi] + l;
                d += t_d[bl][gt];
                gb += t_g[bl][gt];
                sb += t_s[bl][gt];
            }
            const uint32_t db = (d + 7) >> 3, gbb = (gb + 7) >> 3, sbb = (sb + 7) >> 3;
            const uint32_t raw = g.raw_bytes[type][p];
            total += db + (sbb + gbb > raw ? raw : sbb + gbb);
        }
        tab[q * JXS_NR + r] = total;
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

#define SEL_LDS 8192 /* u32 of staged rate tables per slice block */

/* One block per slice: the slice's rate tables are staged in LDS (as many
 * precincts at a time as fit), then thread 0 walks the precincts in order:
 * budgets with the +10 % first precinct and the padding carry
 * (PackStageProcess.c:600-684), the binary searches of Q then R
 * (RateControl.c:889-1083) and the precinct offsets. */
#define SEL_THREADS 256
__global__ void __launch_bounds__(SEL_THREADS) k_select(const JxsFrameDesc* __restrict__ frames, int nfr,
                                                        uint32_t total_slices) {
    __shared__ uint32_t tabs[SEL_LDS];
    const uint32_t gs = blockIdx.x;
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
    __shared__ __align__(16) JxsGeom sg;
    stage_geom(&sg, F.g);
    __syncthreads();
    const JxsGeom& g = sg;
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
    const uint32_t per = (g.maxq + 1) * JXS_NR; /* staged words per precinct */
    const uint32_t chunk = SEL_LDS / per;       /* >= 8 */
    uint32_t carry = 0, off = F.slice_off[s] + 6;
    bool failed = false;
    for (uint32_t i0 = 0; i0 < nprec; i0 += chunk) {
        const uint32_t n = min(chunk, nprec - i0);
        const uint32_t words = n * per;
        const uint32_t* src = F.table + (size_t)(first_row + i0) * JXS_NQ * JXS_NR;
        __syncthreads();
        /* consecutive precincts' tables are JXS_NQ * JXS_NR words apart */
        for (uint32_t b0 = 0; b0 < words; b0 += blockDim.x * 8) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t idx = b0 + j * blockDim.x + threadIdx.x;
                const uint32_t ic = min(idx, words - 1); /* clamped: every load in bounds */
                v[j] = gld(src + (size_t)(ic / per) * JXS_NQ * JXS_NR + ic % per);
            }
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t idx = b0 + j * blockDim.x + threadIdx.x;
                if (idx < words)
                    tabs[idx] = v[j];
            }
        }
        __syncthreads();
        for (uint32_t i = i0; i < i0 + n && threadIdx.x == 0; i++) {
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
            const uint32_t* tab = tabs + (i - i0) * per;
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
}

/* ------------------------------------------------------------------ */
/* packet plan                                                         */
/* ------------------------------------------------------------------ */
/* Inclusive sum over the 64 lanes of a wave with DPP moves: shifts by 1, 2,
 * 4, 8 inside each row of 16 lanes, then the row totals broadcast from lane
 * 15 into rows 1 and 3 and from lane 31 into rows 2 and 3.  Every lane of
 * the wave must be active. */
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t /*lane*/) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(1), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(2), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(4), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(8), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_BCAST15, 0xA, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_BCAST31, 0xC, 0xF, false);
    return (uint32_t)x;
}

__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, WAVE - 1);
}

/* One single-wave block per precinct, after rate control: truncation and
 * coding method of every band for the chosen (Q, R), the sub-packet sizes,
 * RAW choice and codestream offset of every packet (PackPrecinct.c:235-572,
 * RateControl.c:457-781), and the bytes around the packets: slice header
 * (PackHeaders.c:142-156), precinct header (PackHeaders.c:158-196),
 * padding, end-of-codestream marker and, for the first precinct, the frame
 * header.  A precinct whose packets would not add up to the size the rate
 * control chose is flagged and nothing of it is written. */
__global__ void __launch_bounds__(64) k_plan(const JxsFrameDesc* __restrict__ frames, int nfr) {
    __shared__ uint32_t bd_d[JXS_NBL], bd_g[JXS_NBL], bd_s[JXS_NBL];
    __shared__ uint8_t b_ms[JXS_NGB], b_gt[JXS_NGB];
    __shared__ uint32_t k_rel[JXS_NPK], k_sb[JXS_NPK], k_gb[JXS_NPK], k_db[JXS_NPK], k_fl[JXS_NPK];
    __shared__ uint32_t ub_s[JXS_NU], ub_g[JXS_NU], ub_d[JXS_NU];
    const int f = find_frame(frames, nfr, blockIdx.x);
    const JxsFrameDesc& F = frames[f];
    __shared__ __align__(16) JxsGeom sg;
    stage_geom(&sg, F.g);
    __syncthreads();
    const JxsGeom& g = sg;
    const uint32_t row = blockIdx.x - F.prec_base;
    const uint32_t type = row + 1 >= g.rows ? 1 : 0;
    const uint32_t lane = threadIdx.x;
    const JxsPrec P = F.prec[row];
    JxsPlan& PL = F.plan[row];
    if (P.err) {
        /* nothing of a failed precinct is packed */
        if (lane < g.npk)
            PL.pk[lane].flags = 0;
        for (uint32_t u = lane; u < g.n_units; u += WAVE)
            {
                JxsUnit& U = F.units[(size_t)row * g.n_units + u];
                U.len[0] = U.len[1] = U.len[2] = 0;
                U.hlen = 0;
            }
        return;
    }
    const uint16_t* recs = F.hist + (size_t)row * g.n_bl * JXS_HREC;
    if (lane < g.nb_all) {
        const uint32_t gi = lane;
        uint32_t ms = 0, gt = 0;
        if (g.gb[gi] != 0xFF) {
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            gt = truncation(bd.gain, bd.prio, P.q, P.r);
            const uint32_t nl = bd.lines[type]; /* 0, 1 or 2 */
            LineBits l0 = {0, 0, 0}, l1 = {0, 0, 0};
            if (nl > 0)
                l0 = line_bits(recs + (size_t)bd.bl * JXS_HREC, gt, bd.gw);
            if (nl > 1)
                l1 = line_bits(recs + (size_t)(bd.bl + 1) * JXS_HREC, gt, bd.gw);
            const uint32_t nosig = l0.ns + l1.ns;
            const uint32_t sig = bd.sw * nl + (l0.ns - l0.zeroed) + (l1.ns - l1.zeroed);
            ms = (g.sig && nl > 0 && nosig > sig) ? 1u : 0u;
            if (nl > 0) {
                bd_d[bd.bl] = l0.data;
                bd_g[bd.bl] = ms ? l0.ns - l0.zeroed : l0.ns;
                bd_s[bd.bl] = ms ? bd.sw : 0u;
            }
            if (nl > 1) {
                bd_d[bd.bl + 1] = l1.data;
                bd_g[bd.bl + 1] = ms ? l1.ns - l1.zeroed : l1.ns;
                bd_s[bd.bl + 1] = ms ? bd.sw : 0u;
            }
        }
        b_ms[gi] = (uint8_t)ms;
        b_gt[gi] = (uint8_t)gt;
        PL.band[gi] = (uint8_t)(gt | (ms << 4));
    }
    __syncthreads();
    uint32_t size = 0, sb = 0, gbb = 0, db = 0, flags = 0;
    if (lane < g.npk) {
        const uint32_t p = lane, l = g.pk_line[p];
        uint32_t d = 0, gb = 0, s = 0, any = 0;
        for (uint32_t gi = g.pk_bs[p]; gi < g.pk_be[p]; gi++) {
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            if (l >= bd.lines[type])
                continue;
            any = 1;
            d += bd_d[bd.bl + l];
            gb += bd_g[bd.bl + l];
            s += bd_s[bd.bl + l];
        }
        db = (d + 7) >> 3;
        gbb = (gb + 7) >> 3;
        sb = (s + 7) >> 3;
        const uint32_t raw = g.raw_bytes[type][p];
        if (sb + gbb > raw) {
            flags = PKT_RAW;
            sb = 0;
            gbb = raw;
        }
        if (any) {
            flags |= PKT_ANY;
            size = g.pkt_hdr_bytes + sb + gbb + db;
        }
    }
    const uint32_t incl = wave_incl_scan(size, lane);
    const uint32_t sum = wave_last(incl);
    const uint32_t eoc = row + 1 == g.rows ? 2u : 0u;
    const bool valid = g.prec_hdr_bytes + sum + P.pad == P.total && P.offset >= 6 &&
        (uint64_t)P.offset + P.total + eoc <= g.lcod;
    if (!valid) {
        if (lane < g.npk)
            PL.pk[lane].flags = 0;
        for (uint32_t u = lane; u < g.n_units; u += WAVE)
            {
                JxsUnit& U = F.units[(size_t)row * g.n_units + u];
                U.len[0] = U.len[1] = U.len[2] = 0;
                U.hlen = 0;
            }
        if (lane == 0)
            atomicOr(F.status, 2u);
        return;
    }
    const uint32_t rel = g.prec_hdr_bytes + incl - size; /* packet offset from the precinct's first byte */
    if (lane < g.npk) {
        JxsPkt K;
        K.off = P.offset + rel;
        K.s = sb;
        K.g = gbb;
        K.d = db;
        K.flags = flags;
        PL.pk[lane] = K;
    }
    /* packing units, one lane each: the bits of the band line's three
     * segments (significance, bit-plane counts, data), then their places
     * after the earlier band lines of the same packet */
    if (lane < g.npk) {
        k_rel[lane] = rel;
        k_sb[lane] = sb;
        k_gb[lane] = gbb;
        k_db[lane] = db;
        k_fl[lane] = flags;
    }
    __syncthreads();
    for (uint32_t u = lane; u < g.n_units; u += WAVE) {
        const uint32_t p = g.u_p[u], gi = g.u_gi[u], l = g.pk_line[p];
        const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
        const bool raw = (k_fl[p] & PKT_RAW) != 0;
        uint32_t s = 0, gb = 0, d = 0;
        if (l < bd.lines[type]) {
            const uint32_t bl = bd.bl + l;
            s = raw ? 0u : bd_s[bl];
            gb = raw ? 4 * bd.gw : bd_g[bl];
            d = bd_d[bl];
        }
        ub_s[u] = s;
        ub_g[u] = gb;
        ub_d[u] = d;
    }
    __syncthreads();
    {
        const uint64_t pbit = 8 * (uint64_t)P.offset; /* precinct's first bit in the codestream */
        JxsUnit* UR = F.units + (size_t)row * g.n_units;
        for (uint32_t u = lane; u < g.n_units; u += WAVE) {
            const uint32_t p = g.u_p[u], gi = g.u_gi[u], l = g.pk_line[p], j = gi - g.pk_bs[p];
            const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
            const bool raw = (k_fl[p] & PKT_RAW) != 0;
            uint32_t a[3];
            a[0] = 8 * (k_rel[p] + g.pkt_hdr_bytes);
            a[1] = a[0] + 8 * k_sb[p];
            a[2] = a[1] + 8 * k_gb[p];
            for (uint32_t k = 0; k < j; k++) {
                a[0] += ub_s[u - j + k];
                a[1] += ub_g[u - j + k];
                a[2] += ub_d[u - j + k];
            }
            const uint32_t len[3] = {ub_s[u], ub_g[u], ub_d[u]};
            JxsUnit U = {};
            if (l < bd.lines[type]) {
                const uint32_t y = row * bd.hl + l;
                U.gl = F.gcli + bd.gcli_off + (size_t)y * bd.gw;
                U.sl = F.smax + bd.smax_off + (size_t)y * bd.sw;
                U.cl = F.coeff + bd.coeff_off + (size_t)y * bd.cs;
                U.gt = b_gt[gi];
                U.ms = b_ms[gi];
                U.raw = raw ? 1u : 0u;
                U.qpih = g.qpih;
                U.w = bd.w;
                U.gw = bd.gw;
                U.sw = bd.sw;
            }
            for (uint32_t k = 0; k < 3; k++) {
                const uint64_t first = pbit + a[k];
                U.ob[k] = (uint32_t*)F.out + (first >> 5);
                U.lo[k] = (uint32_t)(first & 31);
                U.len[k] = len[k];
            }
            if (j == 0 && (k_fl[p] & PKT_ANY)) {
                /* packet header (PackHeaders.c:98-127): RAW flag, data bytes,
                 * bit-plane-count bytes, sign bytes (0) */
                const uint64_t first = pbit + 8 * k_rel[p];
                U.hob = (uint32_t*)F.out + (first >> 5);
                U.hlo = (uint32_t)(first & 31);
                U.hlen = 8 * g.pkt_hdr_bytes;
                U.hdr = g.short_hdr ? ((uint64_t)raw << 39) | ((uint64_t)k_db[p] << 24) | ((uint64_t)k_gb[p] << 11)
                                    : ((uint64_t)raw << 55) | ((uint64_t)k_db[p] << 35) | ((uint64_t)k_gb[p] << 15);
            }
            U.status = F.status;
            UR[u] = U;
        }
    }
    uint8_t* o = F.out;
    /* packet bodies and padding start as zeros: the packing units OR their
     * first and last words in */
    {
        const uint32_t z0 = P.offset + g.prec_hdr_bytes, z1 = P.offset + P.total;
        const uint32_t a0 = (z0 + 3) & ~3u, a1 = z1 & ~3u;
        if (a0 < a1) {
            for (uint32_t wd = a0 / 4 + lane; wd < a1 / 4; wd += WAVE)
                gst((uint32_t*)o + wd, 0u);
            if (z0 + lane < a0)
                gst(o + z0 + lane, (uint8_t)0);
            if (a1 + lane < z1)
                gst(o + a1 + lane, (uint8_t)0);
        } else {
            for (uint32_t k = z0 + lane; k < z1; k += WAVE)
                gst(o + k, (uint8_t)0);
        }
    }
    /* slice header */
    if (row % g.pps == 0 && lane < 6) {
        const uint32_t s = row / g.pps;
        /* FF20 (SLH marker), Lslh = 4, slice index */
        const uint32_t v = lane == 0 ? 0xFFu : lane == 1 ? 0x20u : lane == 3 ? 0x04u : lane == 4 ? (s >> 8) : lane == 5 ? s : 0u;
        o[P.offset - 6 + lane] = (uint8_t)v;
    }
    /* precinct header: Lprc, Q, R, then 2 bits per existing band */
    const uint32_t lprc = P.total - g.prec_hdr_bytes;
    for (uint32_t k = lane; k < g.prec_hdr_bytes; k += WAVE) {
        uint32_t v;
        if (k < 3)
            v = lprc >> (16 - 8 * k);
        else if (k == 3)
            v = P.q;
        else if (k == 4)
            v = P.r;
        else {
            v = 0;
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t e = 4 * (k - 5) + j;
                if (e < g.nb_exist && b_ms[g.ex_gid[e]])
                    v |= 2u << (6 - 2 * j);
            }
        }
        o[P.offset + k] = (uint8_t)v;
    }
    if (eoc && lane < 2)
        o[P.offset + P.total + lane] = lane ? 0x11 : 0xFF;
    if (row == 0)
        for (uint32_t k = lane; k < g.hdr_len; k += WAVE)
            o[k] = g.hdr[k];
}

/* ------------------------------------------------------------------ */
/* packing                                                             */
/* ------------------------------------------------------------------ */
#define PK_WAVES 4
#define RINGB 512      /* words of a wave's significance / bit-plane-count ring (2 KiB): a step adds <= 128 */
#define RINGC 1024     /* words of a wave's data ring (4 KiB): a step adds <= 512 */
#define RING_FLUSH 256 /* flush once this many complete words are pending */

/* Orders this wave's LDS accesses around cross-lane reads of a ring. */
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* The n (1..64) low bits of v, MSB first, placed at bit offset sh (0..31)
 * of three consecutive 32-bit words (MSB-first bit order). */
__device__ __forceinline__ void split_bits(uint64_t v, uint32_t n, uint32_t sh, uint32_t& w0, uint32_t& w1,
                                           uint32_t& w2) {
    const uint64_t a = v << (64 - n);
    w0 = (uint32_t)(a >> (32 + sh));
    const uint64_t rest = a << (32 - sh);
    w1 = (uint32_t)(rest >> 32);
    w2 = (uint32_t)rest;
}

/* The output of one segment: bits [lo, hi) counted from the dword at ob
 * (the dword holding the segment's first bit), assembled in an LDS ring of
 * mask + 1 words; ring word w is output dword ob[w]. */
struct WaveOut {
    uint32_t* ring;
    uint32_t mask;
    uint32_t* ob;
    uint32_t lo, hi, pos, flushed, lane;
    __device__ __forceinline__ void start(uint32_t* r, uint32_t m, uint32_t* o, uint32_t first, uint32_t len,
                                          uint32_t ln) {
        ring = r;
        mask = m;
        ob = o;
        lo = first;
        hi = first + len;
        pos = first;
        flushed = 0;
        lane = ln;
        const uint32_t words = min((hi + 31) / 32, mask + 1);
        for (uint32_t k = lane; k < words; k += WAVE)
            ring[k] = 0;
    }
    /* ORs the n (1..64) low bits of v into the ring at bit p (words that
     * receive only zeros are skipped) */
    __device__ __forceinline__ void put(uint32_t p, uint64_t v, uint32_t n) {
        uint32_t w0, w1, w2;
        split_bits(v, n, p & 31, w0, w1, w2);
        const uint32_t wd = p >> 5;
        if (w0)
            atomicOr(&ring[wd & mask], w0);
        if (w1)
            atomicOr(&ring[(wd + 1) & mask], w1);
        if (w2)
            atomicOr(&ring[(wd + 2) & mask], w2);
    }
    /* Writes ring words [from, to) out and clears them.  A word wholly inside
     * the segment is stored; the segment's first and last words, which it
     * may share with its neighbours, are ORed into the zeroed output
     * (device-scope atomics: the neighbours may run on other XCDs). */
    __device__ __forceinline__ void flush(uint32_t from, uint32_t to) {
        for (uint32_t w = from + lane; w < to; w += WAVE) {
            uint32_t* slot = &ring[w & mask];
            const uint32_t v = *slot;
            *slot = 0;
            const uint32_t b0 = 32 * w;
            if (b0 >= lo && b0 + 32 <= hi)
                gst(ob + w, __builtin_bswap32(v));
            else if (v)
                atomicOr(ob + w, __builtin_bswap32(v));
        }
    }
    __device__ __forceinline__ void maybe_flush() {
        if ((pos >> 5) - flushed >= RING_FLUSH) {
            wave_sync();
            flush(flushed, pos >> 5);
            flushed = pos >> 5;
            wave_sync();
        }
    }
    __device__ __forceinline__ void finish() {
        wave_sync();
        flush(flushed, (hi + 31) / 32);
    }
};

/* Magnitude of one quantised coefficient shifted down by gt (Quant.c:9-45):
 * deadzone keeps the bit-planes above gt (m >> gt), uniform rounds to the
 * nearest of the 2^(gc-gt) levels (identity at gt = 0). */
template <bool UNIFORM>
__device__ __forceinline__ uint32_t quant_mag(uint32_t c, uint32_t gc, uint32_t gt) {
    const uint32_t m = c & 0x7FFFu;
    if (!UNIFORM)
        return m >> gt;
    const uint32_t sc = gc - gt + 1;
    const uint32_t u = (((m << sc) - m + (1u << gc)) >> (gc + 1)) & 0xFFFFu;
    return gt == 0 ? m : u;
}

/* bit i of x (x < 2^8) -> bit 4i */
__device__ __forceinline__ uint32_t spread8(uint32_t x) {
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x;
}

/* Bit-plane nibbles of four magnitudes (< 2^15): nibble j holds bit j of
 * q0, q1, q2, q3 from its MSB down.  The high word (planes 8..14) is built
 * only when some lane of the wave needs it. */
__device__ __forceinline__ uint64_t interleave4(uint32_t q0, uint32_t q1, uint32_t q2, uint32_t q3, bool any_high) {
    const uint32_t lo = (spread8(q0 & 0xFFu) << 3) | (spread8(q1 & 0xFFu) << 2) | (spread8(q2 & 0xFFu) << 1) |
        spread8(q3 & 0xFFu);
    uint32_t hi = 0;
    if (any_high)
        hi = (spread8(q0 >> 8) << 3) | (spread8(q1 >> 8) << 2) | (spread8(q2 >> 8) << 1) | spread8(q3 >> 8);
    return ((uint64_t)hi << 32) | lo;
}

#define IPL 4 /* consecutive items coded by one lane per step */

/* Significance sub-packet segment of one band line: one bit per group of
 * 8 GCLIs, 1 when its maximum is <= gt (PackPrecinct.c, significance
 * flags).  Lane L codes flags 4L .. 4L+3 of a step of 256. */
__device__ __forceinline__ void emit_sig(WaveOut& W, const uint8_t* sl, uint32_t sw, uint32_t gt) {
    for (uint32_t base = 0; base < sw; base += IPL * WAVE) {
        const uint32_t i0 = base + IPL * W.lane;
        uint32_t a[IPL];
#pragma unroll
        for (uint32_t k = 0; k < IPL; k++)
            a[k] = gld(sl + min(i0 + k, sw - 1));
        uint32_t nb = 0;
        uint64_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < IPL; k++)
            if (i0 + k < sw) {
                v = (v << 1) | (a[k] <= gt ? 1u : 0u);
                nb++;
            }
        const uint32_t incl = wave_incl_scan(nb, W.lane);
        if (nb && v)
            W.put(W.pos + incl - nb, v, nb);
        W.pos += wave_last(incl);
        W.maybe_flush();
    }
}

/* Bit-plane-count and data segments of one band line together, 64 * N
 * groups per step: lane L codes groups N*L .. N*L+N-1 of the step (N = 4
 * for wide lines, 1 for lines of at most 64 groups).  Bit-plane counts are
 * unary (gcli - gt ones and a zero, none for a group inside an
 * insignificant significance group) or raw 4-bit GCLIs; a lane's codes are
 * concatenated (at most 64 bits).  Data: per coded group a sign nibble,
 * then bit-planes gcli-1 .. gt of the quantised magnitudes, 4 bits each;
 * groups no lane codes are skipped.  The inputs of a step are loaded before
 * any of them is coded. */
template <uint32_t N, bool UNIFORM>
__device__ __forceinline__ void emit_gd(WaveOut& B, WaveOut& C, const uint8_t* gl, const uint8_t* sl,
                                        const uint16_t* cl, uint32_t w, uint32_t gw, uint32_t gt, uint32_t ms,
                                        uint32_t raw) {
    const uint32_t lane = B.lane;
    for (uint32_t base = 0; base < gw; base += N * WAVE) {
        const uint32_t i0 = base + N * lane;
        uint32_t a[N], s8[N];
        uint2 cv[N];
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            const uint32_t idx = min(i0 + k, gw - 1); /* clamped: every load in bounds */
            a[k] = gld(gl + idx);
            s8[k] = ms ? gld(sl + idx / 8) : 0u;
            cv[k] = gld2(cl + 4 * idx);
        }
        /* bit-plane counts */
        uint32_t nb = 0;
        uint64_t v = 0;
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            if (i0 + k >= gw)
                break;
            uint32_t len, code;
            if (raw) {
                len = 4;
                code = a[k];
            } else if (ms && s8[k] <= gt) {
                len = 0;
                code = 0;
            } else {
                const uint32_t m = a[k] > gt ? a[k] - gt : 0u;
                len = m + 1;
                code = ((1u << m) - 1) << 1;
            }
            v = (v << len) | code;
            nb += len;
        }
        const uint32_t incl = wave_incl_scan(nb, lane);
        if (nb && v)
            B.put(B.pos + incl - nb, v, nb);
        B.pos += wave_last(incl);
        /* data */
        uint64_t dv[N];
        uint32_t dl[N], dn = 0;
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            const uint32_t i = i0 + k, gc = a[k];
            const bool coded = i < gw && gc > gt;
            dl[k] = 0;
            dv[k] = 0;
            if (__ballot(coded) == 0)
                continue;
            const uint32_t np = coded ? gc - gt : 0u;
            const uint32_t c[4] = {cv[k].x & 0xFFFFu, cv[k].x >> 16, cv[k].y & 0xFFFFu, cv[k].y >> 16};
            uint32_t q[4], signs = 0;
#pragma unroll
            for (uint32_t e = 0; e < 4; e++) {
                q[e] = (coded && 4 * i + e < w) ? quant_mag<UNIFORM>(c[e], gc, gt) : 0u;
                signs |= (q[e] && (c[e] & 0x8000u)) ? (8u >> e) : 0u;
            }
            const bool any_high = __ballot(np > 8) != 0;
            const uint64_t planes = interleave4(q[0], q[1], q[2], q[3], any_high);
            dl[k] = coded ? 4 * np + 4 : 0u;
            dv[k] = coded ? ((uint64_t)signs << (4 * np)) | planes : 0u;
            dn += dl[k];
        }
        const uint32_t dincl = wave_incl_scan(dn, lane);
        uint32_t at = C.pos + dincl - dn;
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            if (dl[k] && dv[k])
                C.put(at, dv[k], dl[k]);
            at += dl[k];
        }
        C.pos += wave_last(dincl);
        B.maybe_flush();
        C.maybe_flush();
    }
}

/* One wave per band line of a precinct (PackPrecinct.c:235-572): its
 * significance flags, bit-plane counts (unary or raw) and data bit-planes
 * with the quantisation (Quant.c:9-45) fused, each segment into its own
 * sub-packet; the first band line of a packet also writes the packet
 * header (PackHeaders.c:98-127, sign size 0).  Every unit is described by
 * one self-contained record of the packet plan, so a wave makes one
 * descriptor load; the GCLIs and coefficients of the line are read once.
 * Waves never wait on each other. */
__global__ void __launch_bounds__(64 * PK_WAVES) k_pack(const JxsUnit* __restrict__ units, uint32_t total_units) {
    __shared__ uint32_t ringb[PK_WAVES][RINGB];
    __shared__ uint32_t ringc[PK_WAVES][RINGC];
    const uint32_t wid = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    const uint32_t gu = blockIdx.x * PK_WAVES + wid;
    if (gu >= total_units)
        return;
    const CAS JxsUnit& U = ((const CAS JxsUnit*)units)[gu];
    bool bad = false;
    if (U.hlen) {
        uint32_t w0, w1, w2;
        split_bits(U.hdr, U.hlen, U.hlo, w0, w1, w2);
        const uint32_t piece = lane == 0 ? w0 : lane == 1 ? w1 : w2;
        if (lane < 3 && piece)
            atomicOr(U.hob + lane, __builtin_bswap32(piece));
    }
    WaveOut B, C;
    if (U.len[0]) {
        B.start(ringb[wid], RINGB - 1, U.ob[0], U.lo[0], U.len[0], lane);
        wave_sync();
        emit_sig(B, U.sl, U.sw, U.gt);
        B.finish();
        bad |= B.pos != B.hi;
    }
    if (U.len[1] | U.len[2]) {
        wave_sync();
        B.start(ringb[wid], RINGB - 1, U.ob[1], U.lo[1], U.len[1], lane);
        C.start(ringc[wid], RINGC - 1, U.ob[2], U.lo[2], U.len[2], lane);
        wave_sync();
        const bool narrow = U.gw <= WAVE;
        if (U.qpih) {
            if (narrow)
                emit_gd<1, true>(B, C, U.gl, U.sl, U.cl, U.w, U.gw, U.gt, U.ms, U.raw);
            else
                emit_gd<IPL, true>(B, C, U.gl, U.sl, U.cl, U.w, U.gw, U.gt, U.ms, U.raw);
        } else {
            if (narrow)
                emit_gd<1, false>(B, C, U.gl, U.sl, U.cl, U.w, U.gw, U.gt, U.ms, U.raw);
            else
                emit_gd<IPL, false>(B, C, U.gl, U.sl, U.cl, U.w, U.gw, U.gt, U.ms, U.raw);
        }
        B.finish();
        C.finish();
        bad |= B.pos != B.hi || C.pos != C.hi;
    }
    /* the plan's lengths are the segments' exact bit counts */
    if (bad && lane == 0)
        atomicOr(U.status, 2u);
}

/* ------------------------------------------------------------------ */
/* launchers (C linkage for the host code)                             */
/* ------------------------------------------------------------------ */
extern "C" {

hipError_t jxs_launch_dwt(const JxsDwtJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, uint32_t wmax,
                          hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    if (wmax == 0 || wmax > HMAX)
        return hipErrorInvalidValue;
    const size_t lds = 2 * (size_t)((wmax + 3) & ~3u) * sizeof(int32_t);
    hipLaunchKernelGGL(k_dwt, dim3(blocks), dim3(256), lds, st, jobs, bases, njobs, blocks, wmax);
    return hipGetLastError();
}

hipError_t jxs_launch_rc_table(const JxsFrameDesc* frames, int nfr, uint32_t precincts, hipStream_t st) {
    hipLaunchKernelGGL(k_rc_table, dim3(precincts), dim3(256), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_select(const JxsFrameDesc* frames, int nfr, uint32_t slices, hipStream_t st) {
    hipLaunchKernelGGL(k_select, dim3(slices), dim3(SEL_THREADS), 0, st, frames, nfr, slices);
    return hipGetLastError();
}

hipError_t jxs_launch_plan(const JxsFrameDesc* frames, int nfr, uint32_t precincts, hipStream_t st) {
    hipLaunchKernelGGL(k_plan, dim3(precincts), dim3(64), 0, st, frames, nfr);
    return hipGetLastError();
}

hipError_t jxs_launch_pack(const JxsUnit* units, uint32_t n, hipStream_t st) {
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3((n + PK_WAVES - 1) / PK_WAVES), dim3(64 * PK_WAVES), 0, st, units, n);
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
