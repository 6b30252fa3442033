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
 *   k_rate     one block per slice, one wave per precinct: the precinct's
 *              rate table (bits per band line and truncation) from the
 *              histogram records (RateControl.c:457-781), then, precinct
 *              after precinct in one wave, the budgets with the +10 % first
 *              precinct and the padding carry (PackStageProcess.c:600-684)
 *              and the reference's binary searches of Q and R
 *              (RateControl.c:889-1083, BinarySearch.c) on the table
 *              entries they visit, then per precinct the band truncations
 *              and coding methods, packet sizes and offsets, the packing
 *              units, slice / precinct headers, padding, end marker and
 *              frame header.
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
#include "jxs_wave.h"

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
/* A struct stored to global memory as 16-byte (or 8-byte) words: assigning
 * the struct through a global pointer would go through a generic-address
 * copy, i.e. flat stores. */
template <typename T>
__device__ __forceinline__ void gst_words(T* p, const T& v) {
    if constexpr (sizeof(T) % 16 == 0 && alignof(T) >= 8) {
        typedef uint32_t w4 __attribute__((ext_vector_type(4)));
        const w4* s = (const w4*)&v;
        GAS w4* d = (GAS w4*)p;
#pragma unroll
        for (uint32_t i = 0; i < sizeof(T) / 16; i++)
            d[i] = s[i];
    } else {
        static_assert(sizeof(T) % 8 == 0, "8-byte multiple");
        const uint64_t* s = (const uint64_t*)&v;
        GAS uint64_t* d = (GAS uint64_t*)p;
#pragma unroll
        for (uint32_t i = 0; i < sizeof(T) / 8; i++)
            d[i] = s[i];
    }
}
/* A read-only struct loaded through the constant address space (scalar
 * loads), word by word. */
template <typename T>
__device__ __forceinline__ T cld_words(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
    T v;
    const CAS uint32_t* s = (const CAS uint32_t*)p;
    uint32_t* d = (uint32_t*)&v;
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++)
        d[i] = s[i];
    return v;
}
/* A value the same in every lane of the wave, declared so: what depends
 * on it (addresses, loops, branches) then stays on the scalar unit. */
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

/* Block barrier for LDS traffic only: waits for this wave's LDS accesses,
 * not for its global loads and stores as __syncthreads() does (the blocks
 * that use it exchange nothing through global memory). */
__device__ __forceinline__ void lds_barrier() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

/* Device-scope atomic OR into global memory (a global, not a flat, atomic:
 * it does not hold up the wave's later LDS accesses). */
__device__ __forceinline__ void gor(uint32_t* p, uint32_t v) {
    (void)__hip_atomic_fetch_or((GAS uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
/* The shadow word of codestream dword p (a frame's shadow array lies at a
 * fixed byte distance from its codestream). */
__device__ __forceinline__ uint32_t* shadow_of(uint32_t* p, int64_t sdelta) {
    return (uint32_t*)((uint8_t*)p + sdelta);
}

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

/* Transform layout: every output row (pair) is a row slot of whole waves;
 * thread t of a slot owns the CH consecutive columns [t * CH, t * CH + CH)
 * of its row(s) in registers, through the vertical step, every horizontal
 * level and the band output.  Neighbouring threads exchange one sample per
 * lifting step through LDS. */
#define DW_MAXT 512              /* threads of the largest transform block */
#define DW_SLOTS (DW_MAXT / WAVE) /* row slots per block (at most one per wave) */
#define DW_NBAND 8  /* band rows of a slot: horizontal levels 0..4, low-pass, LH, HH */
#define DW_LF 5
#define DW_LH 6
#define DW_HH 7
#define DW_STAGED 3 /* band rows too short for whole groups per thread, staged in LDS */

/* Sample access for the transform.  Input samples are scaled as
 * NltEnc.c:33-71 does: x = (v & mask) << shift - offset; the int32 low-pass
 * planes of the second level are read as they are. */
struct Scale {
    uint32_t shift, mask;
    int32_t offset;
    __device__ __forceinline__ int32_t operator()(uint32_t v) const { return (int32_t)((v & mask) << shift) - offset; }
};

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }

template <int KIND>
__device__ __forceinline__ int32_t ld1(const void* row, uint32_t x, uint32_t step, const Scale& sc) {
    if (KIND == JXS_SRC_U16)
        return sc(gld((const uint16_t*)row + (size_t)x * step));
    if (KIND == JXS_SRC_U8)
        return sc(gld((const uint8_t*)row + (size_t)x * step));
    return gld((const int32_t*)row + (size_t)x * step);
}

/* Samples of a row for the transform, NC columns [x0, x0 + NC) per call.
 * With `vec` (contiguous rows, 16-byte aligned, width a multiple of 8 and at
 * least 8: checked by the host) the row is read 8 samples per load, every
 * load issued before any is used; a load past the row's end reads the row's
 * last 8 samples instead, whose values then stand in for columns that no
 * output uses.  Otherwise one load per sample, the column clamped into the
 * row. */
template <int KIND, int NC>
struct RowRaw {
    /* NC (8 or 16) samples: 8-sample loads (4-sample loads for int32 rows) */
    u32x4 q[KIND == JXS_SRC_I32 ? NC / 4 : NC / 8];
};
template <int KIND, int NC>
__device__ __forceinline__ RowRaw<KIND, NC> load_vec(const void* row, uint32_t x0, uint32_t w) {
    RowRaw<KIND, NC> r;
    if constexpr (KIND == JXS_SRC_I32) {
#pragma unroll
        for (int k = 0; k < NC / 4; k++)
            r.q[k] = gld((const u32x4*)((const int32_t*)row + min(x0 + 4 * k, w - 4)));
    } else if constexpr (KIND == JXS_SRC_U16) {
#pragma unroll
        for (int k = 0; k < NC / 8; k++)
            r.q[k] = gld((const u32x4*)((const uint16_t*)row + min(x0 + 8 * k, w - 8)));
    } else {
#pragma unroll
        for (int k = 0; k < NC / 8; k++) {
            const u32x2 v = gld((const u32x2*)((const uint8_t*)row + min(x0 + 8 * k, w - 8)));
            r.q[k] = {v.x, v.y, 0u, 0u};
        }
    }
    return r;
}
/* NC samples of component `comp` (0..2) of a packed 16-bit row (three
 * interleaved components; row and plane 16-byte aligned, width a multiple of
 * 8: checked by the host), as the RowRaw of a planar row: eight samples take
 * the 24 interleaved samples [3x, 3x + 24) in three 16-byte loads, and every
 * sample pair of the component is picked from two of their words.  `row`
 * points at the component's first sample (the row start + comp). */
template <int NC>
__device__ __forceinline__ RowRaw<JXS_SRC_U16, NC> load_vec3(const void* row, uint32_t x0, uint32_t w, uint32_t comp) {
    RowRaw<JXS_SRC_U16, NC> r;
    const uint16_t* base = (const uint16_t*)row - comp;
    /* pair p of eight samples = samples 2p, 2p + 1 = interleaved 6p + comp, 6p + 3 + comp:
     * low half from word 3p (comp 0, 1) or 3p + 1 (comp 2), high half from word 3p + 1 (comp 0) or 3p + 2 */
    const uint32_t sel = comp == 1 ? 0x05040302u : 0x07060100u;
#pragma unroll
    for (int k = 0; k < NC / 8; k++) {
        const uint16_t* p = base + 3 * (size_t)min(x0 + 8 * k, w - 8);
        const u32x4 a = gld((const u32x4*)p), b = gld((const u32x4*)(p + 8)), c = gld((const u32x4*)(p + 16));
        const uint32_t d[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
        uint32_t pr[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t lo = comp == 2 ? d[3 * q + 1] : d[3 * q];
            const uint32_t hi = comp == 0 ? d[3 * q + 1] : d[3 * q + 2];
            pr[q] = __builtin_amdgcn_perm(hi, lo, sel);
        }
        r.q[k] = {pr[0], pr[1], pr[2], pr[3]};
    }
    return r;
}

/* load_vec3 for the kernels of every input kind (the host sets packed vector
 * rows for 16-bit samples only). */
template <int KIND, int NC>
__device__ __forceinline__ RowRaw<KIND, NC> load_packed(const void* row, uint32_t x0, uint32_t w, uint32_t comp) {
    if constexpr (KIND == JXS_SRC_U16)
        return load_vec3<NC>(row, x0, w, comp);
    else
        return load_vec<KIND, NC>(row, x0, w);
}

template <int KIND, int NC>
__device__ __forceinline__ void unpack(const RowRaw<KIND, NC>& r, const Scale& sc, int32_t (&out)[NC]) {
    if constexpr (KIND == JXS_SRC_I32) {
#pragma unroll
        for (int k = 0; k < NC / 4; k++) {
            out[4 * k] = (int32_t)r.q[k].x;
            out[4 * k + 1] = (int32_t)r.q[k].y;
            out[4 * k + 2] = (int32_t)r.q[k].z;
            out[4 * k + 3] = (int32_t)r.q[k].w;
        }
    } else if constexpr (KIND == JXS_SRC_U16) {
#pragma unroll
        for (int k = 0; k < NC / 8; k++) {
            const uint32_t wd[4] = {r.q[k].x, r.q[k].y, r.q[k].z, r.q[k].w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                out[8 * k + 2 * j] = sc(wd[j] & 0xFFFFu);
                out[8 * k + 2 * j + 1] = sc(wd[j] >> 16);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < NC / 8; k++)
#pragma unroll
            for (int j = 0; j < 8; j++)
                out[8 * k + j] = sc(((j < 4 ? r.q[k].x : r.q[k].y) >> (8 * (j & 3))) & 0xFFu);
    }
}
template <int KIND, int NC>
__device__ __forceinline__ void load_scalar(const void* row, uint32_t x0, uint32_t w, uint32_t step, const Scale& sc,
                                            int32_t (&out)[NC]) {
#pragma unroll
    for (int k = 0; k < NC; k++)
        out[k] = ld1<KIND>(row, min(x0 + k, w - 1), step, sc);
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

/* The vertical 5/3 step on unscaled samples.  With x' = x * 2^s - offset
 * (s = 20 - depth >= 6) and D = 2 x(2i+1) - x(2i) - x(2i+2) (|D| < 2^15
 * for depths up to 14), the high sample is D * 2^(s-1) exactly, and the
 * low sample's rounded shifts drop no bits:
 *   lf = (8 x(2i) + P + C) * 2^(s-3) - offset,
 * with P = D(i-1), C = D(i) inside the plane; P = C = D(i) at the top row
 * pair; P = C = D(i-1) when there is no high row; x(2i+2) := x(2i) when it
 * lies past the end (the same values vert53 gives on scaled samples).  So
 * the differences are formed on 16-bit sample pairs, two columns per
 * instruction, and only the results are widened. */
__device__ __forceinline__ int32_t lo16(uint32_t w) { return (int32_t)(w << 16) >> 16; }
__device__ __forceinline__ int32_t hi16(uint32_t w) { return (int32_t)w >> 16; }
template <int KIND, int NC>
__device__ __forceinline__ uint32_t pair16(const RowRaw<KIND, NC>& r, int k) {
    /* samples 2k and 2k + 1 of the row as two 16-bit halves */
    if constexpr (KIND == JXS_SRC_U16) {
        const u32x4 q = r.q[k / 4];
        const int e = k % 4;
        return e == 0 ? q.x : e == 1 ? q.y : e == 2 ? q.z : q.w;
    } else {
        const u32x4 q = r.q[k / 4];
        const uint32_t v = (k % 4) < 2 ? q.x : q.y;
        const uint32_t b = 2 * (k % 2);
        return __builtin_amdgcn_perm(0u, v, 0x0C000C00u | ((b + 1) << 16) | b);
    }
}
template <int KIND, int NC>
__device__ __forceinline__ void vert_raw(const VertFlags& f, const RowRaw<KIND, NC> (&raw)[5], const Scale& sc,
                                         int32_t* lo, int32_t* hi) {
    const uint32_t m2 = sc.mask * 0x10001u, s = sc.shift;
#pragma unroll
    for (int k = 0; k < NC / 2; k++) {
        const u16x2 a = as_u16x2(pair16<KIND, NC>(raw[0], k) & m2), b = as_u16x2(pair16<KIND, NC>(raw[1], k) & m2);
        const u16x2 c = as_u16x2(pair16<KIND, NC>(raw[2], k) & m2), d = as_u16x2(pair16<KIND, NC>(raw[3], k) & m2);
        const u16x2 e = as_u16x2(pair16<KIND, NC>(raw[4], k) & m2);
        const u16x2 one = {1, 1};
        const uint32_t dc = as_u32((d << one) - c - e), dp = as_u32((b << one) - a - c);
        const uint32_t C = f.h ? dc : 0u;
        const uint32_t P = f.prev ? dp : C;
        const uint32_t C2 = f.h ? C : P;
        const uint32_t cw = as_u32(c);
        lo[2 * k] = (int32_t)((8 * (cw & 0xFFFFu) + lo16(P) + lo16(C2)) << (s - 3)) - sc.offset;
        lo[2 * k + 1] = (int32_t)((8 * (cw >> 16) + hi16(P) + hi16(C2)) << (s - 3)) - sc.offset;
        hi[2 * k] = lo16(C) * (1 << (s - 1));
        hi[2 * k + 1] = hi16(C) * (1 << (s - 1));
    }
}

/* Horizontal 5/3 lifting of one level on the CH / S samples v[k * S] a
 * thread holds (global sample index t * (CH / S) + k of a level of n
 * samples; n >= 2).  Predict, at odd positions:
 *   x(2i+1) -= (x(2i) + x(2i+2)) >> 1, x(2i+2) := x(2i) past the end;
 * update, at even positions:
 *   x(2i) += (x(2i-1) + x(2i+1) + 2) >> 2, a missing neighbour being
 * replaced by the other one.  The sample after the thread's last one (first
 * of the next thread) and the one before its first (the previous thread's
 * last, after the predict) come from the neighbours.
 *
 * The edge tests compare k with rem = n - t * M, the samples of the level
 * from the thread's first on (may be <= 0 or > M).  With 16 or more
 * columns per thread, rem is passed through an empty asm statement in each
 * step, so that the per-sample lane masks are recomputed where they are used
 * instead of being kept (and spilled out of the scalar registers) from one
 * step to the next. */
template <int CH>
__device__ __forceinline__ int32_t lift_rem(uint32_t n, uint32_t first) {
    int32_t rem = (int32_t)n - (int32_t)first;
    if constexpr (CH >= 16)
        asm volatile("" : "+v"(rem));
    return rem;
}
template <int CH, int S>
__device__ __forceinline__ void lift_predict(int32_t (&v)[CH], uint32_t t, uint32_t n, int32_t right_first) {
    constexpr int M = CH / S;
    const int32_t rem = lift_rem<CH>(n, t * M);
#pragma unroll
    for (int k = 1; k < M; k += 2) {
        const int32_t xl = v[(k - 1) * S];
        const int32_t xr = (k + 1 < M) ? v[(k + 1) * S] : right_first;
        const int32_t r = (k + 1 < rem) ? xr : xl;
        const int32_t nv = v[k * S] - ((xl + r) >> 1);
        v[k * S] = (k < rem) ? nv : v[k * S];
    }
}
template <int CH, int S>
__device__ __forceinline__ void lift_update(int32_t (&v)[CH], uint32_t t, uint32_t n, int32_t left_last) {
    constexpr int M = CH / S;
    const int32_t rem = lift_rem<CH>(n, t * M);
#pragma unroll
    for (int k = 0; k < M; k += 2) {
        const int32_t hr = v[(k + 1) * S];
        const int32_t hl0 = (k > 0) ? v[(k - 1) * S] : left_last;
        const int32_t hl = (k > 0 || t > 0) ? hl0 : hr;
        const int32_t hr2 = (k + 1 < rem) ? hr : hl;
        const int32_t nv = v[k * S] + ((hl + hr2 + 2) >> 2);
        v[k * S] = (k < rem) ? nv : v[k * S];
    }
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


/* Band row output of G groups of 4 coefficients per thread (groups g0 ..
 * g0 + G - 1 of band row y, count coefficients in the row): 16-bit
 * sign-magnitude coefficients (image_shift, NltEnc.c:12-31; zero past the
 * end of the row), the GCLI of every group and the maximum GCLI of every 8
 * groups (GcStageProcess.c:22-97, across 8 / G neighbouring lanes), and,
 * into hc, the counts of the band-line histogram record: GCLI values from
 * wave ballots (lanes 0..15 count one value each), maxima of full
 * significance groups below 15, and the trailing partial group's maximum
 * + 1 at HREC_TAIL.  Called by whole waves (act: this lane's row exists). */
template <int G>
__device__ __forceinline__ void emit_groups(const JxsBandOut& o, uint32_t y, const int32_t (&val)[4 * G], uint32_t g0,
                                            uint32_t count, uint32_t* hc, bool act, uint32_t lane) {
    uint16_t* crow = o.coeff + (size_t)y * o.cs;
    uint8_t* grow = o.gcli + (size_t)y * o.gw;
    uint8_t* srow = o.smax + (size_t)y * o.sw;
    uint32_t mx = 0, mycnt = 0;
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
        const uint32_t g = g0 + gi;
        const bool gv = act && g < o.gw;
        uint32_t v[4], orv = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int32_t x = val[4 * gi + k];
            const uint32_t a = (uint32_t)(x < 0 ? -x : x);
            const uint32_t mg = (a + 128) >> 8;
            v[k] = (4 * g + k < count && mg) ? (mg | (((uint32_t)x >> 31) << 15)) : 0u;
            orv |= v[k];
        }
        const uint32_t gc = gv ? bitlen(orv & 0x7FFFu) : 0u;
        if (gv) {
            gst2(crow + 4 * g, make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16)));
            gst(grow + g, (uint8_t)gc);
        }
        mx = max(mx, gc);
        if (hc) {
            const uint64_t vg = __ballot(gv);
            const uint64_t b0 = __ballot(gc & 1u), b1 = __ballot(gc & 2u), b2 = __ballot(gc & 4u), b3 = __ballot(gc & 8u);
            /* lanes whose GCLI equals this lane's index: bit k of the GCLI
             * matches bit k of the lane */
            const uint64_t e0 = (lane & 1u) ? b0 : ~b0, e1 = (lane & 2u) ? b1 : ~b1;
            const uint64_t e2 = (lane & 4u) ? b2 : ~b2, e3 = (lane & 8u) ? b3 : ~b3;
            mycnt += (uint32_t)__popcll(vg & e0 & e1 & e2 & e3);
        }
    }
    /* maximum over the 8 groups of a significance group: 8 / G lanes */
    if (G < 2)
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, 4, WAVE));
    if (G < 4)
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, 2, WAVE));
    if (G < 8)
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, 1, WAVE));
    const bool head = act && (g0 & 7) == 0 && g0 < o.gw;
    if (head)
        gst(srow + g0 / 8, (uint8_t)mx);
    if (hc) {
        if (head) {
            if (g0 / 8 >= o.gw / 8)
                hc[HREC_TAIL] = mx + 1; /* the partial significance group */
            else if (mx < 15)
                atomicAdd(&hc[16 + mx], 1u); /* maximum 15: implied by the group count */
        }
        if (lane < 16 && mycnt)
            atomicAdd(&hc[lane], mycnt);
    }
}

/* A band row staged in LDS (values st[0 .. count)), one group per thread of
 * the slot. */
__device__ __forceinline__ void emit_staged(const JxsBandOut& o, uint32_t y, const int32_t* st, uint32_t t,
                                            uint32_t count, uint32_t* hc, bool act, uint32_t lane) {
    int32_t val[4];
    const uint32_t last = count > 0 ? count - 1 : 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        val[k] = st[min(4 * t + k, last)];
    emit_groups<1>(o, y, val, t, count, hc, act, lane);
}

/* Per-block state of the transform (LDS). */
struct DwLds {
    int32_t exr[2][DW_MAXT]; /* first sample of every thread (low row, high row) */
    int32_t exl[2][DW_MAXT]; /* last sample of every thread after the predict */
    int32_t stage[DW_STAGED][2 * DW_MAXT]; /* short band rows: at most 2 samples per thread */
    uint32_t hc[DW_SLOTS][DW_NBAND][JXS_HREC];
};

/* What a slot's thread knows of its row. */
struct DwRow {
    uint32_t tid, lane, slot, t, tpr, ir;
    bool act, right, left;
    uint32_t st_count[DW_STAGED]; /* staged band rows: coefficient count (0: none) */
    uint32_t st_band[DW_STAGED];
};

/* The low-pass samples lo[k * S] (k < CH / S) after the last horizontal level
 * of the low row: the LL band, or the int32 plane of the next vertical
 * level. */
template <int CH, int S>
__device__ __forceinline__ void dw_lowpass(const CAS JxsDwtJob& J, DwLds& L, DwRow& R, const int32_t (&lo)[CH],
                                           uint32_t nlev) {
    constexpr int M = CH / S;
    if (J.lf_is_band) {
        const JxsBandOut ob = band_out(J.lf);
        if constexpr (M >= 4) {
            int32_t v[M];
#pragma unroll
            for (int k = 0; k < M; k++)
                v[k] = lo[k * S];
            emit_groups<M / 4>(ob, R.ir, v, R.t * (M / 4), nlev, ob.hist ? L.hc[R.slot][DW_LF] : nullptr, R.act,
                               R.lane);
        } else {
#pragma unroll
            for (int k = 0; k < M; k++)
                L.stage[DW_STAGED - 1][R.slot * R.tpr * 2 + R.t * M + k] = lo[k * S];
            R.st_count[DW_STAGED - 1] = nlev;
            R.st_band[DW_STAGED - 1] = DW_LF;
        }
    } else if (R.act) {
        int32_t* dst = J.lf_i32 + (size_t)R.ir * J.lf_stride;
        const uint32_t x0 = R.t * M;
        if constexpr (M % 4 == 0) {
            if ((J.lf_stride & 3) == 0 && x0 + M <= nlev) {
                /* 16-byte stores: plane 16-byte aligned, rows a multiple of 4 samples */
#pragma unroll
                for (int k = 0; k < M; k += 4) {
                    const u32x4 v = {(uint32_t)lo[k * S], (uint32_t)lo[(k + 1) * S], (uint32_t)lo[(k + 2) * S],
                                     (uint32_t)lo[(k + 3) * S]};
                    gst((u32x4*)(dst + x0 + k), v);
                }
                return;
            }
        }
#pragma unroll
        for (int k = 0; k < M; k++) {
            const uint32_t x = x0 + k;
            if (x < nlev)
                gst(dst + x, lo[k * S]);
        }
    }
}

/* Horizontal level LVL of the low row (stride S = 2^LVL between the samples
 * of this level in lo[]), the high row's single level with level 0, the
 * output of the level's high band, then the next level or the low-pass. */
template <int CH, int LVL>
__device__ __forceinline__ void dw_level(const CAS JxsDwtJob& J, DwLds& L, DwRow& R, int32_t (&lo)[CH],
                                         int32_t (&hi)[CH], bool has_hi, uint32_t w, uint32_t nlev) {
    constexpr int S = 1 << LVL, M = CH / S;
    const uint32_t tid = R.tid;
    L.exr[0][tid] = lo[0];
    if constexpr (LVL == 0)
        L.exr[1][tid] = hi[0];
    lds_barrier();
    lift_predict<CH, S>(lo, R.t, nlev, R.right ? L.exr[0][tid + 1] : 0);
    if constexpr (LVL == 0)
        if (has_hi)
            lift_predict<CH, 1>(hi, R.t, w, R.right ? L.exr[1][tid + 1] : 0);
    L.exl[0][tid] = lo[(M - 1) * S];
    if constexpr (LVL == 0)
        L.exl[1][tid] = hi[CH - 1];
    lds_barrier();
    lift_update<CH, S>(lo, R.t, nlev, R.left ? L.exl[0][tid - 1] : 0);
    if constexpr (LVL == 0)
        if (has_hi)
            lift_update<CH, 1>(hi, R.t, w, R.left ? L.exl[1][tid - 1] : 0);
    const uint32_t nh = nlev / 2;
    const JxsBandOut ob = band_out(J.hf[LVL]);
    if constexpr (M >= 8) {
        int32_t v[M / 2];
#pragma unroll
        for (int k = 0; k < M / 2; k++)
            v[k] = lo[(2 * k + 1) * S];
        emit_groups<M / 8>(ob, R.ir, v, R.t * (M / 8), nh, ob.hist ? L.hc[R.slot][LVL] : nullptr, R.act, R.lane);
    } else {
        /* M = 4 or 2: the first or second short level */
        constexpr int SI = M == 4 ? 0 : 1;
#pragma unroll
        for (int k = 0; k < M / 2; k++)
            L.stage[SI][R.slot * R.tpr * 2 + R.t * (M / 2) + k] = lo[(2 * k + 1) * S];
        R.st_count[SI] = nh;
        R.st_band[SI] = LVL;
    }
    if constexpr (LVL == 0) {
        if (has_hi) {
            /* the high row: LH from its low samples, HH from its high ones */
            int32_t vl[CH / 2], vh[CH / 2];
#pragma unroll
            for (int k = 0; k < CH / 2; k++) {
                vl[k] = hi[2 * k];
                vh[k] = hi[2 * k + 1];
            }
            const JxsBandOut olh = band_out(J.lh), ohh = band_out(J.hh);
            emit_groups<CH / 8>(olh, R.ir, vl, R.t * (CH / 8), w - w / 2, olh.hist ? L.hc[R.slot][DW_LH] : nullptr,
                                R.act, R.lane);
            emit_groups<CH / 8>(ohh, R.ir, vh, R.t * (CH / 8), w / 2, ohh.hist ? L.hc[R.slot][DW_HH] : nullptr,
                                R.act, R.lane);
        }
    }
    const uint32_t nl = nlev - nh;
    if constexpr (M > 2) {
        if ((uint32_t)LVL + 1 < J.lev_l) {
            dw_level<CH, LVL + 1>(J, L, R, lo, hi, false, w, nl);
            return;
        }
    }
    dw_lowpass<CH, 2 * S>(J, L, R, lo, nl);
}

/* One block of the transform for one input kind and chunk width CH: the
 * rows of J.rpb slots, each of J.tpr threads.  The number of barriers is the
 * same for every thread of the block (the job fixes it). */
template <int KIND, int CH>
__device__ __forceinline__ void dwt_slots(const CAS JxsDwtJob& J, uint32_t bi, DwLds& L) {
    DwRow R;
    R.tid = threadIdx.x;
    R.lane = R.tid & (WAVE - 1);
    R.tpr = J.tpr;
    R.slot = wave_uniform(R.tid / R.tpr); /* a slot is whole waves */
    R.t = R.tid - R.slot * R.tpr;
    const uint32_t i = bi * J.rpb + R.slot; /* output row (pair) */
    R.act = R.slot < J.rpb && i < J.rows;
    R.ir = wave_uniform(R.act ? i : 0); /* inactive slots work on row 0 and write nothing */
    R.right = R.t + 1 < R.tpr;
    R.left = R.t > 0;
#pragma unroll
    for (int k = 0; k < DW_STAGED; k++)
        R.st_count[k] = R.st_band[k] = 0;
    const uint32_t w = J.w, n = J.h, step = J.step, x0 = R.t * CH;
    const bool vert = J.vert != 0, vec = J.vec != 0;
    const Scale sc = {J.shift, J.mask, J.offset};
    const uint8_t* base = (const uint8_t*)J.src;
    const size_t pitch = (size_t)J.stride * (KIND == JXS_SRC_U8 ? 1u : KIND == JXS_SRC_U16 ? 2u : 4u);
    const uint32_t ir = R.ir;
    const VertFlags f = {ir > 0, ir < n / 2, 2 * ir + 2 < n};
    const bool has_hi = vert && f.h;
    int32_t lo[CH], hi[CH];
    if (KIND == JXS_SRC_U16 && vert && J.vec >= 2) {
        /* packed 16-bit rows: the same unscaled vertical step on component
         * J.vec - 2 of input rows 2i-2 .. 2i+2, eight columns at a time */
        const int i2 = (int)(2 * ir), last = (int)n - 1;
        const uint8_t* r[5];
        r[0] = base + (size_t)max(i2 - 2, 0) * pitch;
        r[1] = base + (size_t)max(i2 - 1, 0) * pitch;
        r[2] = base + (size_t)i2 * pitch;
        r[3] = base + (size_t)min(i2 + 1, last) * pitch;
        r[4] = base + (size_t)(f.next ? i2 + 2 : i2) * pitch;
        const uint32_t comp = J.vec - 2;
#pragma unroll
        for (int h0 = 0; h0 < CH; h0 += 8) {
            RowRaw<KIND, 8> raw[5];
#pragma unroll
            for (int j = 0; j < 5; j++)
                raw[j] = load_packed<KIND, 8>(r[j], x0 + h0, w, comp);
            vert_raw<KIND, 8>(f, raw, sc, lo + h0, hi + h0);
        }
    } else if (KIND != JXS_SRC_I32 && vert && vec) {
        /* input rows 2i-2 .. 2i+2 (clamped into the plane, the flags drop
         * what lies outside; row 2i+2 := row 2i past the end), unscaled */
        const int i2 = (int)(2 * ir), last = (int)n - 1;
        const uint8_t* r[5];
        r[0] = base + (size_t)max(i2 - 2, 0) * pitch;
        r[1] = base + (size_t)max(i2 - 1, 0) * pitch;
        r[2] = base + (size_t)i2 * pitch;
        r[3] = base + (size_t)min(i2 + 1, last) * pitch;
        r[4] = base + (size_t)(f.next ? i2 + 2 : i2) * pitch;
        constexpr int NC = CH < 16 ? CH : 16;
#pragma unroll
        for (int h0 = 0; h0 < CH; h0 += NC) {
            RowRaw<KIND, NC> raw[5];
#pragma unroll
            for (int j = 0; j < 5; j++)
                raw[j] = load_vec<KIND, NC>(r[j], x0 + h0, w);
            vert_raw<KIND, NC>(f, raw, sc, lo + h0, hi + h0);
        }
    } else if (vert) {
        /* input rows 2i-2 .. 2i+2, clamped into the plane (the flags drop
         * what lies outside), 16 columns at a time */
        const uint8_t* r[5];
#pragma unroll
        for (int j = 0; j < 5; j++)
            r[j] = base + (size_t)min(max((int)(2 * ir) + j - 2, 0), (int)n - 1) * pitch;
        constexpr int NC = CH < 16 ? CH : 16;
#pragma unroll
        for (int h0 = 0; h0 < CH; h0 += NC) {
            int32_t x[5][NC];
            if (vec) {
                RowRaw<KIND, NC> raw[5];
#pragma unroll
                for (int j = 0; j < 5; j++)
                    raw[j] = load_vec<KIND, NC>(r[j], x0 + h0, w);
#pragma unroll
                for (int j = 0; j < 5; j++)
                    unpack<KIND, NC>(raw[j], sc, x[j]);
            } else {
#pragma unroll
                for (int j = 0; j < 5; j++)
                    load_scalar<KIND, NC>(r[j], x0 + h0, w, step, sc, x[j]);
            }
#pragma unroll
            for (int k = 0; k < NC; k++)
                vert53(f, x[0][k], x[1][k], x[2][k], x[3][k], x[4][k], lo[h0 + k], hi[h0 + k]);
        }
    } else {
        constexpr int NC = CH < 16 ? CH : 16;
#pragma unroll
        for (int h0 = 0; h0 < CH; h0 += NC) {
            int32_t x[NC];
            if (KIND == JXS_SRC_U16 && J.vec >= 2)
                unpack<KIND, NC>(load_packed<KIND, NC>(base + (size_t)ir * pitch, x0 + h0, w, J.vec - 2), sc, x);
            else if (vec)
                unpack<KIND, NC>(load_vec<KIND, NC>(base + (size_t)ir * pitch, x0 + h0, w), sc, x);
            else
                load_scalar<KIND, NC>(base + (size_t)ir * pitch, x0 + h0, w, step, sc, x);
#pragma unroll
            for (int k = 0; k < NC; k++) {
                lo[h0 + k] = x[k];
                hi[h0 + k] = 0;
            }
        }
    }
    dw_level<CH, 0>(J, L, R, lo, hi, has_hi, w, w);
    /* short band rows, one group per thread, from LDS */
    lds_barrier();
#pragma unroll
    for (int k = 0; k < DW_STAGED; k++) {
        if (R.st_count[k] == 0)
            continue;
        const uint32_t b = R.st_band[k];
        const JxsBandOut ob = band_out(b == DW_LF ? J.lf : J.hf[b]);
        emit_staged(ob, ir, &L.stage[k][R.slot * R.tpr * 2], R.t, R.st_count[k], ob.hist ? L.hc[R.slot][b] : nullptr,
                    R.act && R.t < ob.gw, R.lane);
    }
    /* histogram records of the band rows written, by the first wave of
     * every slot, band after band (uniform descriptors) */
    lds_barrier();
    if (R.act && R.t < WAVE) {
        for (uint32_t b = 0; b < DW_NBAND; b++) {
            bool used;
            const CAS JxsBandOut* od;
            if (b < DW_LF) {
                used = b < J.lev_l;
                od = &J.hf[b];
            } else if (b == DW_LF) {
                used = J.lf_is_band != 0;
                od = &J.lf;
            } else {
                used = has_hi;
                od = b == DW_LH ? &J.lh : &J.hh;
            }
            uint16_t* hist = od->hist;
            if (!used || !hist)
                continue;
            if (R.lane < JXS_HREC) {
                const uint32_t k = R.lane, c = L.hc[R.slot][b][k];
                /* a band has 1 or 2 lines per precinct (checked by the host) */
                const uint32_t hs = od->hl >> 1;
                uint16_t* rec = hist + ((size_t)(ir >> hs) * od->n_bl + od->bl + (ir & hs)) * JXS_HREC;
                gst(rec + k, (uint16_t)(k == HREC_TAIL ? (c ? c - 1 : 0xFFFFu) : c));
            }
        }
    }
}

/* One block per J.rpb output rows (row pairs with a vertical level) of one
 * job: with a vertical level, row pair i = low row i and high row i from
 * input rows 2i-2 .. 2i+2; the low row goes through the job's horizontal
 * levels (bands of every level, then the LL band or the int32 low-pass plane
 * of the next vertical level), the high row through one level into the LH
 * and HH bands.  Without a vertical level, input row i through the
 * horizontal levels.  The first launch of a plan also clears the frames'
 * error flags (status, nstatus words).  Register budgets, for occupancy:
 * the 32-column variant is held to 168 VGPRs (3 waves per SIMD instead of
 * 2, no spill: its 8K packed-RGB launch ran 6 % faster), the 16-column one
 * to 128 (4 waves instead of 3, no spill: 4K second level 10.8 -> 9.3 us per
 * batched frame), the 8-column one to 64 (8 waves, 4 blocks per CU instead
 * of 3, at the cost of a few spills: 4K first level 29.9 -> 28.1 us per
 * plain frame, 23.7 -> 21.5 batched). */
template <int CH, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(CH == 32 ? 3 : CH == 16 ? 4 : 8))) k_dwt(const JxsDwtJob* __restrict__ jobs,
                                                    const uint32_t* __restrict__ bases, int njobs, uint32_t nblocks,
                                                    uint32_t* status, uint32_t nstatus) {
    __shared__ DwLds L;
    const uint32_t blk = xcd_remap(blockIdx.x, nblocks);
    const CAS JxsDwtJob* jl = (const CAS JxsDwtJob*)jobs;
    const CAS JxsDwtJob* jp = jl + find_job_u32(bases, njobs, blk);
    uint32_t bi = blk - jp->base;
    if (jp->il > 1) {
        const uint32_t il = jp->il, first = jp->il_first, rel = blk - jl[first].base;
        jp = jl + first + rel % il;
        bi = rel / il;
    }
    const CAS JxsDwtJob& J = *jp;
    if (status && blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < nstatus; k += NT)
            gst(status + k, 0u);
    for (uint32_t e = threadIdx.x; e < DW_SLOTS * DW_NBAND * JXS_HREC; e += NT)
        (&L.hc[0][0][0])[e] = 0;
    lds_barrier();
    const uint32_t kind = J.kind;
    if (kind == JXS_SRC_U16)
        dwt_slots<JXS_SRC_U16, CH>(J, bi, L);
    else if (kind == JXS_SRC_U8)
        dwt_slots<JXS_SRC_U8, CH>(J, bi, L);
    else
        dwt_slots<JXS_SRC_I32, CH>(J, bi, L);
}

/* Copies a frame's geometry into LDS (all NTH threads of the block take
 * part; the caller synchronises).  The rate-control kernel indexes it with
 * data-dependent band and packet numbers. */
#define GEOM_THREADS 256 /* block size of the kernels that stage the geometry */
template <uint32_t NTH = GEOM_THREADS>
__device__ __forceinline__ void stage_geom(JxsGeom* dst, const JxsGeom* src) {
    const u32x4* s = (const u32x4*)src; /* read through global, not flat, loads */
    u32x4* d = (u32x4*)dst;
    constexpr uint32_t n = sizeof(JxsGeom) / 16;
    constexpr uint32_t per = (n + NTH - 1) / NTH; /* copies per thread, loaded together */
    const uint32_t t = threadIdx.x;
    /* indexes clamped into the struct: every load is in bounds, only the
     * stores are predicated */
    u32x4 v[per];
#pragma unroll
    for (uint32_t j = 0; j < per; j++)
        v[j] = gld(s + min(t + j * NTH, n - 1));
#pragma unroll
    for (uint32_t j = 0; j < per; j++)
        if (t + j * NTH < n)
            d[t + j * NTH] = v[j];
}

/* ------------------------------------------------------------------ */
/* rate control                                                        */
/* ------------------------------------------------------------------ */
/* The reference's binary search (BinarySearch.c:9-78) over [0, end], replayed
 * on the outcome of every probe given at once as bit masks: probing v
 * answers "too small" when bit v of m1 is set, "out of range" when bit v of
 * m3 is set, "too big" otherwise.  `below` keeps the largest "too small"
 * probe as the answer, otherwise the smallest "too big" one; the first probe
 * is at step0 (or end / 2).  Uniform values only: the loop runs on the scalar
 * unit. */
__device__ __forceinline__ bool bsearch_masks(uint32_t m1, uint32_t m3, int32_t end, bool below, int32_t step0,
                                              uint32_t& out) {
    /* The common case in closed form: "out of range" only at the top of
     * [0, end] (from position a on) and "too small" exactly at the first k
     * positions below a.  The search then ends at the first "too big"
     * position (k, if k < a) or at the last "too small" one (k - 1, if
     * k > 0), whatever the first probe; the replay below agrees with this
     * for every end <= 31, a and k (checked exhaustively). */
    {
        const uint32_t full = end >= 31 ? ~0u : ((1u << (end + 1)) - 1);
        const uint32_t m3f = m3 & full;
        const uint32_t a = m3f ? (uint32_t)__builtin_ctz(m3f) : (uint32_t)end + 1;
        const uint32_t valid = a >= 32 ? ~0u : ((1u << a) - 1);
        const uint32_t sm = m1 & valid;
        if ((m3f | valid) == full && (sm & (sm + 1)) == 0) {
            const uint32_t k = (uint32_t)__builtin_popcount(sm);
            if (below) {
                out = k - 1;
                return k > 0;
            }
            out = k < a ? k : ~0u;
            return k < a;
        }
    }
    int32_t beg = 0, best = -1;
    int32_t step = step0 ? step0 : end / 2;
    if (step > end - beg)
        step = (end - beg + 1) / 2;
    int32_t last = step + beg; /* beg <= last <= end <= 31 while the search runs */
    for (;;) {
        const bool oor = (m3 >> last) & 1u;
        const bool small = !oor && ((m1 >> last) & 1u);
        const bool big = !oor && !small;
        if ((small && below && (best == -1 || best < last)) || (big && !below && (best == -1 || best > last)))
            best = last;
        if (small ? end < last + 1 : beg > last - 1)
            break;
        if (small)
            beg = last + 1;
        else
            end = last - 1;
        if (step > end - beg)
            step = (end - beg + 1) / 2;
        last = step + beg;
    }
    out = (uint32_t)best;
    return best != -1;
}

/* Inclusive sum over the 64 lanes of a wave (every lane active). */
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t /*lane*/) { return wave_scan_dpp(v); }

__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, WAVE - 1);
}

/* Orders this wave's LDS accesses around reads of what other lanes of the
 * same wave wrote. */
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_of(uint32_t v, uint32_t k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k);
}
/* A value the same in every lane, declared so: the compiler then keeps what
 * depends on it, loops and branches included, on the scalar unit. */
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

/* ------------------------------------------------------------------ */
/* rate control and packet plan                                        */
/* ------------------------------------------------------------------ */
#define RT_STRIDE JXS_RT_STRIDE
#define RT_ZERO 16   /* the zero entry, read for band lines absent from the precinct */
#define RT_NW 16     /* waves of a k_rate block */
#define RT_NP 8      /* precinct rate tables a k_rate block holds at once */

/* Rate table of one precinct, built by one wave from the transform's
 * band-line histogram records: for every band line and truncation t, the
 * data bits (bit-planes above t and the sign bits), the bit-plane-count bits
 * and the significance bits of the band's coding method at t.  These are the
 * quantities RateControl.c:18-154 counts coefficient group by group and
 * RateControl.c:457-520 compares to choose the method. */
struct RateTab {
    uint32_t d[JXS_NBL * RT_STRIDE];
    uint32_t g[JXS_NBL * RT_STRIDE];
    uint32_t s[JXS_NBL * RT_STRIDE];
};

/* The histogram record of band line `line` of a precinct (64 bytes). */
struct Rec {
    u32x4 a, b, c, e;
};
__device__ __forceinline__ Rec load_rec(const uint16_t* recs, uint32_t line) {
    const u32x4* src = (const u32x4*)(recs + (size_t)line * JXS_HREC);
    Rec r;
    r.a = gld(src);
    r.b = gld(src + 1);
    r.c = gld(src + 2);
    r.e = gld(src + 3);
    return r;
}

/* Rate-table entries t in [T0, T0 + NT) of band line `lane` (one lane per
 * band line; every lane of the wave takes part, lanes past the last line
 * store nothing).  With h(v) the groups whose GCLI is v, at truncation t the
 * unary bit-plane counts take sum_{v<=t} h(v) + sum_{v>t} (v - t) h(v) bits,
 * the data sum_{v>t} 4 (v - t + 1) h(v) bits (v - t magnitude bit-planes and
 * a sign per coefficient), and significance coding removes the groups of the
 * significance groups of 8 whose maximum GCLI is at most t (and of the
 * trailing partial group) at the cost of one flag per significance group;
 * all are running sums over t, started from their sums below T0.  The
 * coding method compares the two counts over the band's lines in this
 * precinct (the other line of a two-line band is the neighbouring lane). */
template <uint32_t T0, uint32_t NT>
__device__ __forceinline__ void rate_lines_part(RateTab& T, const Rec& rc, uint32_t n_bl, uint32_t li, uint32_t type,
                                                uint32_t sig_on, uint32_t lane) {
    const uint32_t w[16] = {rc.a.x, rc.a.y, rc.a.z, rc.a.w, rc.b.x, rc.b.y, rc.b.z, rc.b.w,
                            rc.c.x, rc.c.y, rc.c.z, rc.c.w, rc.e.x, rc.e.y, rc.e.z, rc.e.w};
    uint32_t tot = 0, vtot = 0;
#pragma unroll
    for (uint32_t v = 0; v < 16; v++) {
        const uint32_t h = (w[v >> 1] >> (16 * (v & 1))) & 0xFFFFu;
        tot += h;
        vtot += v * h;
    }
    const uint32_t gw = li & 0xFFFu, full = gw >> 3, part = gw & 7, sw = (li >> LI_SW) & 0xFFu;
    const uint32_t nl = (li >> (LI_NL + 2 * type)) & 3u;
    const bool two = nl == 2;
    const int partner = (int)(two ? (((li >> LI_SECOND) & 1u) ? lane - 1 : lane + 1) : lane);
    const uint32_t tail = w[15] >> 16; /* 0xFFFF: no partial group */
    uint32_t cnt = 0, vcnt = 0, sig = 0;
#pragma unroll
    for (uint32_t v = 0; v < T0; v++) { /* T0 <= 12: every v < 15 */
        const uint32_t h = (w[v >> 1] >> (16 * (v & 1))) & 0xFFFFu;
        cnt += h;
        vcnt += v * h;
        sig += (w[8 + (v >> 1)] >> (16 * (v & 1))) & 0xFFFFu;
    }
    uint32_t data[NT], ns[NT], zr[NT];
#pragma unroll
    for (uint32_t j = 0; j < NT; j++) {
        const uint32_t t = T0 + j;
        const uint32_t h = (w[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
        cnt += h;
        vcnt += t * h;
        const uint32_t above = tot - cnt, vabove = vtot - vcnt;
        data[j] = vabove + above - t * above; /* sum_{v>t} (v - t + 1) h(v) */
        if (t < 15)
            sig += (w[8 + (t >> 1)] >> (16 * (t & 1))) & 0xFFFFu;
        else
            sig = full;
        ns[j] = data[j] + cnt;
        zr[j] = 8 * sig + (tail <= t ? part : 0u);
    }
    /* the other line's counts, all exchanged before any is used */
    uint32_t pn[NT], pz[NT];
#pragma unroll
    for (uint32_t j = 0; j < NT; j++) {
        pn[j] = (uint32_t)__shfl((int)ns[j], partner);
        pz[j] = (uint32_t)__shfl((int)zr[j], partner);
    }
    if (lane < n_bl) {
        const uint32_t base = lane * RT_STRIDE + T0;
#pragma unroll
        for (uint32_t j = 0; j < NT; j++) {
            const uint32_t nosig = ns[j] + (two ? pn[j] : 0u);
            const uint32_t withsig = sw * nl + (ns[j] - zr[j]) + (two ? pn[j] - pz[j] : 0u);
            const bool use = sig_on && nosig > withsig;
            T.d[base + j] = 4 * data[j];
            T.g[base + j] = use ? ns[j] - zr[j] : ns[j];
            T.s[base + j] = use ? sw : 0u;
        }
        if (T0 == 0) {
            T.d[lane * RT_STRIDE + RT_ZERO] = 0;
            T.g[lane * RT_STRIDE + RT_ZERO] = 0;
            T.s[lane * RT_STRIDE + RT_ZERO] = 0;
        }
    }
}

/* The share `sub` (of 2^lg, a wave each) of a precinct's rate table: one,
 * two or all four quarters of the truncations, a quarter at a time (which
 * keeps few values live). */
__device__ __forceinline__ void rate_lines(RateTab& T, const Rec& rc, uint32_t n_bl, uint32_t li, uint32_t type,
                                           uint32_t sig_on, uint32_t lane, uint32_t lg, uint32_t sub) {
    const uint32_t nq = 4u >> lg;
    for (uint32_t k = 0; k < nq; k++) {
        const uint32_t qi = sub * nq + k;
        if (qi == 0)
            rate_lines_part<0, 4>(T, rc, n_bl, li, type, sig_on, lane);
        else if (qi == 1)
            rate_lines_part<4, 4>(T, rc, n_bl, li, type, sig_on, lane);
        else if (qi == 2)
            rate_lines_part<8, 4>(T, rc, n_bl, li, type, sig_on, lane);
        else
            rate_lines_part<12, 4>(T, rc, n_bl, li, type, sig_on, lane);
    }
}

/* Bytes of the packets of walk entries [lo, hi) of a precinct at (q, r)
 * (precinct_get_budget_bytes, RateControl.c:700-781; the entries are whole
 * packets): per packet the data bytes plus the smaller of the significance
 * + bit-plane-count bytes and the packet's RAW size.  Each lane answers its
 * own (q, r).  The band lines are walked in packet order from registers:
 * lane k of rq holds the k-th one's walk word, of rw its packet's RAW bytes.
 * The walk takes 8 lines per step, their 24 table reads issued together;
 * entries at or past hi are read (in bounds) but not counted. */
__device__ __forceinline__ uint32_t rate_query_seg(uint32_t rq, uint32_t rw, uint32_t lo, uint32_t hi,
                                                   const RateTab& T, uint32_t q, uint32_t r) {
    uint32_t total = 0, d = 0, gb = 0, s = 0;
    for (uint32_t k0 = lo; k0 < hi; k0 += 8) {
        uint32_t vd[8], vg[8], vs[8];
#pragma unroll
        for (uint32_t kk = 0; kk < 8; kk++) {
            const uint32_t x = lane_of(rq, min(k0 + kk, (uint32_t)(WAVE - 1)));
            const int32_t t = (int32_t)q - (int32_t)((x >> RQ_GAIN) & 0xFFu) - (((x >> RQ_PRIO) & 0xFFu) < r ? 1 : 0);
            const uint32_t o = (x & 0x7FFu) + (uint32_t)min(max(t, 0), 15);
            vd[kk] = T.d[o];
            vg[kk] = T.g[o];
            vs[kk] = T.s[o];
        }
#pragma unroll
        for (uint32_t kk = 0; kk < 8; kk++) {
            if (k0 + kk < hi) {
                const uint32_t x = lane_of(rq, k0 + kk);
                d += vd[kk];
                gb += vg[kk];
                s += vs[kk];
                if ((x >> RQ_END) & 1u) {
                    total += ((d + 7) >> 3) + min(((s + 7) >> 3) + ((gb + 7) >> 3), lane_of(rw, k0 + kk));
                    d = gb = s = 0;
                }
            }
        }
    }
    return total;
}

/* The whole walk of n entries. */
__device__ __forceinline__ uint32_t rate_query(uint32_t rq, uint32_t rw, uint32_t n, const RateTab& T, uint32_t q,
                                               uint32_t r) {
    return rate_query_seg(rq, rw, 0, n, T, q, r);
}

/* The binary search of Q (RateControl.c:889-1083): lane q of cq holds the
 * precinct's bytes at (q, 0), bit q of eq is set when (q, 0) truncates
 * every band away; too small a Q is one whose bytes exceed avail. */
__device__ __forceinline__ bool search_q(uint32_t maxq, uint32_t eq, uint32_t cq, uint32_t avail, uint32_t& q) {
    const uint32_t m1 = (uint32_t)__ballot(cq > avail); /* lanes 0..31 */
    return bsearch_masks(m1, eq, (int32_t)maxq, false, 6 > maxq ? 0 : 6, q);
}

/* The binary search of R at Q = q: lane base + r of cr holds the bytes at
 * (q, r), bit r of er is set when (q, r) truncates every band away; an R
 * whose bytes fit avail is kept as the answer. */
__device__ __forceinline__ bool search_r(uint32_t maxr, uint32_t er, uint32_t cr, uint32_t base, uint32_t avail,
                                         uint32_t& r) {
    const uint32_t m1 = (uint32_t)(__ballot(cr <= avail) >> base);
    return bsearch_masks(m1, er, (int32_t)maxr, true, 6 >= maxr ? 0 : 6, r);
}

/* A packing unit that writes nothing. */
__device__ __forceinline__ void unit_empty(JxsUnit* u) {
    GAS JxsUnit* g = (GAS JxsUnit*)u;
    g->len[0] = g->len[1] = g->len[2] = 0;
    g->hlen = 0;
}

/* Sum of v over lanes [first, last] of the wave, given incl = wave_incl_scan(v). */
__device__ __forceinline__ uint32_t seg_sum(uint32_t incl, uint32_t v, uint32_t first, uint32_t last) {
    return (uint32_t)__shfl((int)incl, (int)last) - (uint32_t)__shfl((int)(incl - v), (int)first);
}
/* Sum of v over lanes [first, lane) of the wave. */
__device__ __forceinline__ uint32_t seg_before(uint32_t incl, uint32_t v, uint32_t first) {
    return (incl - v) - (uint32_t)__shfl((int)(incl - v), (int)first);
}

/* Packet plan of one precinct (one wave, one lane per packing unit), once
 * its (Q, R) is chosen and each unit's significance, bit-plane-count and
 * data bits at its band's truncation are known (vs, vg, vd; zeros for a
 * line absent from the precinct): truncation and coding method of every band, the
 * sub-packet sizes, RAW choice and codestream offset of every packet
 * (PackPrecinct.c:235-572, RateControl.c:457-781), the packing units, and the
 * bytes around the packets: slice header (PackHeaders.c:142-156), precinct
 * header (PackHeaders.c:158-196), padding, end-of-codestream marker and, for
 * the first precinct, the frame header.  Sums over a packet's band lines and
 * over the packets come from wave prefix sums.  A precinct whose packets
 * would not add up to the size the rate control chose is flagged and nothing
 * of it is written. */
__device__ __forceinline__ void plan_precinct(const JxsFrameDesc& F, const JxsGeom& g, uint32_t row, uint32_t type,
                                              const JxsPrec& P, uint32_t vs, uint32_t vg, uint32_t vd, uint32_t* ms_band,
                                              uint32_t ui, uint32_t lane) {
    const uint32_t nu = g.n_units;
    const bool unit = lane < nu;
    /* this lane's unit record: unit-major order, widest lines first */
    JxsUnit* UR = F.units + (size_t)g.u_rank[lane] * g.rows + row;
    if (P.err) {
        /* nothing of a failed precinct is packed */
        if (unit)
            unit_empty(UR);
        return;
    }
    const uint32_t p = unit ? g.u_p[lane] : 0u, gi = unit ? g.u_gi[lane] : 0u, l = g.pk_line[p];
    const uint32_t j = unit ? gi - g.pk_bs[p] : 0u; /* band line's place in its packet */
    const uint32_t first = lane - j, last = unit ? first + (g.pk_be[p] - g.pk_bs[p]) - 1 : lane;
    const bool pres = unit && ((ui >> (UI_NORMAL + type)) & 1u);
    const uint32_t gt = truncation((ui >> 8) & 0xFFu, (ui >> 16) & 0xFFu, P.q, P.r);
    const uint32_t ms = vs != 0 ? 1u : 0u;
    if (unit && l == 0)
        ms_band[gi] = ms;
    /* packet sizes: every lane of a packet computes them */
    const uint32_t ps_ = seg_sum(wave_incl_scan(vs, lane), vs, first, last);
    const uint32_t pg = seg_sum(wave_incl_scan(vg, lane), vg, first, last);
    const uint32_t pd = seg_sum(wave_incl_scan(vd, lane), vd, first, last);
    const uint64_t pm = __ballot(pres);
    const uint32_t n_in = last - first + 1;
    const bool any = unit && ((pm >> first) & (n_in >= 64 ? ~0ull : ((1ull << n_in) - 1))) != 0;
    uint32_t sb = (ps_ + 7) >> 3, gbb = (pg + 7) >> 3;
    const uint32_t db = (pd + 7) >> 3, raw_p = g.raw_bytes[type][p];
    const bool raw = sb + gbb > raw_p;
    if (raw) {
        sb = 0;
        gbb = raw_p;
    }
    const uint32_t size = any ? g.pkt_hdr_bytes + sb + gbb + db : 0u;
    const uint32_t psz = (unit && j == 0) ? size : 0u;
    const uint32_t pincl = wave_incl_scan(psz, lane);
    const uint32_t sum = wave_last(pincl);
    const uint32_t eoc = row + 1 == g.rows ? 2u : 0u;
    const bool valid = g.prec_hdr_bytes + sum + P.pad == P.total && P.offset >= 6 &&
        (uint64_t)P.offset + P.total + eoc <= g.lcod;
    if (!valid) {
        if (unit)
            unit_empty(UR);
        if (lane == 0)
            gor(F.status, 2u);
        return;
    }
    /* the packet's offset from the precinct's first byte */
    const uint32_t rel = g.prec_hdr_bytes + (uint32_t)__shfl((int)(pincl - psz), (int)first);
    /* the band line's three segments (RAW packets: no significance, 4 bits
     * per group for the bit-plane counts) and their places after the earlier
     * band lines of the packet */
    const JxsBand& bd = g.c[g.gc[gi]].b[g.gb[gi]];
    const uint32_t us = raw ? 0u : vs, ug = raw ? (pres ? 4 * bd.gw : 0u) : vg, ud = vd;
    const uint32_t bs = seg_before(wave_incl_scan(us, lane), us, first);
    const uint32_t bg = seg_before(wave_incl_scan(ug, lane), ug, first);
    const uint32_t bdd = seg_before(wave_incl_scan(ud, lane), ud, first);
    wave_sync(); /* ms_band complete */
    if (unit) {
        const uint64_t pbit = 8 * (uint64_t)P.offset; /* precinct's first bit in the codestream */
        const uint32_t a0 = 8 * (rel + g.pkt_hdr_bytes) + bs;
        const uint32_t a1 = 8 * (rel + g.pkt_hdr_bytes + sb) + bg;
        const uint32_t a2 = 8 * (rel + g.pkt_hdr_bytes + sb + gbb) + bdd;
        JxsUnit U = {};
        if (pres) {
            const uint32_t y = row * bd.hl + l;
            U.gl = F.gcli + bd.gcli_off + (size_t)y * bd.gw;
            U.sl = F.smax + bd.smax_off + (size_t)y * bd.sw;
            U.cl = F.coeff + bd.coeff_off + (size_t)y * bd.cs;
            U.gt = gt;
            U.ms = ms;
            U.raw = raw ? 1u : 0u;
            U.qpih = g.qpih;
            U.w = bd.w;
            U.gw = bd.gw;
            U.sw = bd.sw;
        }
        const uint64_t f0 = pbit + a0, f1 = pbit + a1, f2 = pbit + a2;
        U.ob[0] = (uint32_t*)F.out + (f0 >> 5);
        U.ob[1] = (uint32_t*)F.out + (f1 >> 5);
        U.ob[2] = (uint32_t*)F.out + (f2 >> 5);
        U.lo[0] = (uint32_t)(f0 & 31);
        U.lo[1] = (uint32_t)(f1 & 31);
        U.lo[2] = (uint32_t)(f2 & 31);
        U.len[0] = us;
        U.len[1] = ug;
        U.len[2] = ud;
        if (j == 0 && any) {
            /* packet header (PackHeaders.c:98-127): RAW flag, data bytes,
             * bit-plane-count bytes, sign bytes (0) */
            const uint64_t hb = pbit + 8 * rel;
            U.hob = (uint32_t*)F.out + (hb >> 5);
            U.hlo = (uint32_t)(hb & 31);
            U.hlen = 8 * g.pkt_hdr_bytes;
            U.hdr = g.short_hdr ? ((uint64_t)raw << 39) | ((uint64_t)db << 24) | ((uint64_t)gbb << 11)
                                : ((uint64_t)raw << 55) | ((uint64_t)db << 35) | ((uint64_t)gbb << 15);
        }
        U.status = F.status;
        U.sdelta = (int64_t)((uint8_t*)F.shadow - F.out);
        gst_words(UR, U); /* units are 144-byte records of a 256-byte aligned array */
        /* the shadow words the packer ORs into: first and last dword of each
         * segment, the packet header's dwords */
#pragma unroll
        for (uint32_t k = 0; k < 3; k++)
            if (U.len[k]) {
                gst(shadow_of(U.ob[k], U.sdelta), 0u);
                gst(shadow_of(U.ob[k] + ((U.lo[k] + U.len[k] - 1) >> 5), U.sdelta), 0u);
            }
        if (U.hlen)
            for (uint32_t k = 0; k <= ((U.hlo + U.hlen - 1) >> 5); k++)
                gst(shadow_of(U.hob + k, U.sdelta), 0u);
    }
    uint8_t* o8 = F.out;
    /* packet bodies and padding start as zeros: the packing units OR their
     * first and last words in */
    {
        /* 16-byte stores between the 16-byte aligned addresses inside
         * [z0, z1), single bytes before and after */
        uint8_t* z0 = o8 + P.offset + g.prec_hdr_bytes;
        uint8_t* z1 = o8 + P.offset + P.total;
        uint8_t* a0 = (uint8_t*)(((uintptr_t)z0 + 15) & ~(uintptr_t)15);
        uint8_t* a1 = (uint8_t*)((uintptr_t)z1 & ~(uintptr_t)15);
        if (a0 < a1) {
            const u32x4 zero = {0u, 0u, 0u, 0u};
            const uint32_t nq = (uint32_t)((a1 - a0) / 16);
            for (uint32_t q4 = lane; q4 < nq; q4 += WAVE)
                gst((u32x4*)a0 + q4, zero);
            if (z0 + lane < a0)
                gst(z0 + lane, (uint8_t)0);
            if (a1 + lane < z1)
                gst(a1 + lane, (uint8_t)0);
        } else {
            for (uint8_t* b = z0 + lane; b < z1; b += WAVE)
                gst(b, (uint8_t)0);
        }
    }
    /* slice header */
    if (row % g.pps == 0 && lane < 6) {
        const uint32_t s = row / g.pps;
        /* FF20 (SLH marker), Lslh = 4, slice index */
        const uint32_t v = lane == 0 ? 0xFFu : lane == 1 ? 0x20u : lane == 3 ? 0x04u : lane == 4 ? (s >> 8) : lane == 5 ? s : 0u;
        gst(o8 + P.offset - 6 + lane, (uint8_t)v);
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
            for (uint32_t e4 = 0; e4 < 4; e4++) {
                const uint32_t e = 4 * (k - 5) + e4;
                if (e < g.nb_exist && ms_band[g.ex_gid[e]])
                    v |= 2u << (6 - 2 * e4);
            }
        }
        gst(o8 + P.offset + k, (uint8_t)v);
    }
    if (eoc && lane < 2)
        gst(o8 + P.offset + P.total + lane, (uint8_t)(lane ? 0x11 : 0xFF));
    if (row == 0)
        for (uint32_t k = lane; k < g.hdr_len; k += WAVE)
            gst(o8 + k, g.hdr[k]);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += (uint32_t)__shfl_xor((int)v, o, WAVE);
    return v;
}
/* v, as a value the compiler cannot see through (nothing computed from it
 * is hoisted out of a loop). */
__device__ __forceinline__ uint32_t opaque_v(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, o, WAVE));
    return v;
}

/* Precinct budgets of a slice under rate control modes 2 and 3
 * (PackStageProcess.c:340-481), by wave 0 with one lane per precinct, from
 * ds_i = headers + data bytes at the slice's (Q, R), stored in the precincts'
 * `total` words: the leftover of the slice's data budget goes to the last
 * precinct; mode 3 raises every precinct below a quarter of the largest ds
 * to that quarter, taking the bytes from the leftover, else one byte at a
 * time from the first precincts on (the closed form: D / n each, one more
 * for the first D mod n); what is left then goes 10 % of the first
 * precinct's bytes to the first precinct and the rest evenly, one byte more
 * for the first ones.  The budgets replace the `total` words. */
__device__ void slice_budgets(const JxsFrameDesc& F, uint32_t first_row, uint32_t nprec, uint32_t mode, uint32_t left,
                              uint32_t dmax, uint32_t lane) {
    const uint32_t mn = mode == 3 ? dmax / 4 : 0u;
    uint32_t get = 0;
    for (uint32_t i = lane; i < nprec; i += WAVE) {
        const uint32_t ds = gld(&F.prec[first_row + i].total);
        get += ds < mn ? mn - ds : 0u;
    }
    get = uni(wave_sum(get));
    uint32_t pad1 = left, dec = 0;
    if (get) {
        if (left >= get) {
            pad1 = left - get;
        } else {
            pad1 = 0;
            dec = get - left;
        }
    }
    const uint32_t dq = dec / nprec, dr = dec - dq * nprec;
    /* the first precinct after the mode-3 step, for its 10 % */
    const uint32_t ds0 = uni(gld(&F.prec[first_row].total));
    const uint32_t b0 = ds0 + (ds0 < mn ? mn - ds0 : 0u) - (dq + (0 < dr ? 1u : 0u)) + (nprec == 1 ? pad1 : 0u);
    uint32_t af = 0, lb = 0;
    if (pad1 > 0) {
        af = min(b0 * 10 / 100, pad1);
        lb = pad1 - af;
    }
    const uint32_t lq = lb / nprec, lr = lb - lq * nprec;
    for (uint32_t i = lane; i < nprec; i += WAVE) {
        const uint32_t ds = gld(&F.prec[first_row + i].total);
        const uint32_t b = ds + (ds < mn ? mn - ds : 0u) - (dq + (i < dr ? 1u : 0u)) + (i == 0 ? af : 0u) + lq +
            (i < lr ? 1u : 0u);
        gst(&F.prec[first_row + i].total, b);
    }
}

/* Rate control modes 2 and 3 (one (Q, R) per slice sizing the precinct
 * budgets): one block per slice (RT_NW waves), the rate control and packet
 * plan of its precincts, 2^lg waves per precinct (4 while a slice has at
 * most 4 precincts, 2 up to 8, else 1) and up to RT_NP precincts at a time:
 *   1. the waves of a precinct build its rate table from the histogram
 *      records, each a share of the truncations;
 *   2. the slice's Q, then R, by the per-precinct binary searches on the
 *      slice's summed bytes, each wave walking a run of whole packets of its
 *      precinct, then every precinct's budget (PackStageProcess.c:340-481);
 *   3. per precinct, the bytes at (q, 0) for every q and the (q, r) rows at
 *      the Q its budget gets without a padding carry and at that Q - 1 (a
 *      carry only ever adds bytes, rarely more than one step of Q's worth);
 *   4. the first wave of each precinct runs the reference's binary searches
 *      (RateControl.c:889-1083, BinarySearch.c) at that budget and notes the
 *      smallest column or row entry above it: a search's outcome depends on
 *      the budget only through which entries exceed it, so any budget up to
 *      that entry gets the same (Q, R);
 *   5. wave 0 walks the precincts in order with the padding carry
 *      (PackStageProcess.c:600-684), taking the noted (Q, R) while the
 *      carried budget stays below the noted entry, else searching again on
 *      the column and rows (a row computed now if the carry moved Q
 *      further);
 *   6. the first wave of each precinct plans it. */
__global__ void __launch_bounds__(WAVE * RT_NW) k_rate(const JxsFrameDesc* __restrict__ frames, int nfr,
                                                      uint32_t total_slices) {
    __shared__ __align__(16) JxsGeom sg;
    __shared__ RateTab tabs[RT_NP];
    __shared__ uint32_t part[RT_NW][WAVE];
    __shared__ uint32_t col[RT_NP][WAVE], rowq[RT_NP][WAVE];
    __shared__ uint32_t spec[RT_NP], sp_r[RT_NP], sp_data[RT_NP], sp_thr[RT_NP], slice_qr;
    __shared__ JxsPrec ps[RT_NP];
    __shared__ uint32_t ms_band[RT_NP][JXS_NGB];
    __shared__ u32x4 recs[RT_NP][4][WAVE]; /* the first precincts' histogram records, quarter k of line l at [k][l] */
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
    /* the frame descriptor in scalar registers */
    const JxsFrameDesc F = cld_words(frames + fl);
    if (F.ordered || !F.slice_rc)
        return; /* coded by k_enc_slice, or by k_rate_tab / k_rate_plan */
    /* the wave index is made visibly uniform: everything a wave derives from
     * it (its precinct, the searches, the plan) then stays in scalar
     * registers and scalar branches */
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1);
    const uint32_t wid = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid / WAVE));
    const uint32_t pps = F.pps, rows = F.rows, n_bl = F.n_bl;
    const uint32_t s = gs - F.slice_base;
    const uint32_t first_row = s * pps;
    const uint32_t nprec = (s == F.slices - 1) ? rows - first_row : pps;
    const uint32_t lg = pps <= 4 ? 2u : pps <= 8 ? 1u : 0u, wpp = 1u << lg;
    const uint32_t np = min((uint32_t)RT_NW >> lg, (uint32_t)RT_NP); /* precincts at a time */
    const uint32_t slot = wid >> lg, sub = wid & (wpp - 1);
    const bool active = slot < np;
    /* the first precincts' records are staged in LDS with the geometry:
     * the waves of a precinct load a share of each line's record */
    const uint32_t rec_line = min(lane, n_bl - 1);
    {
        const u32x4* src = (const u32x4*)(F.hist + ((size_t)min(first_row + slot, rows - 1) * n_bl + rec_line) * JXS_HREC);
        const uint32_t nk = 4u >> lg;
        u32x4 v[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (active && k < nk)
                v[k] = gld(src + sub * nk + k);
        stage_geom<WAVE * RT_NW>(&sg, F.g);
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (active && k < nk)
                recs[slot][sub * nk + k][lane] = v[k];
    }
    __syncthreads();
    const JxsGeom& g = sg;
    /* per-lane geometry words are read from LDS where they are used, which
     * keeps them out of the registers live across the kernel */
    const uint32_t maxq = uni(g.maxq), maxr = uni(g.maxr);
    /* bit q: (q, 0) truncates every band away */
    const uint32_t eq = (uint32_t)__ballot(g.empty_mask[lane & (JXS_NQ - 1)] & 1u);
    const uint32_t sbud = gld(F.slice_bud + s); /* global, not flat: a uniform value */
    const uint32_t rc_mode = uni(g.rc_mode);
    uint32_t carry = 0, off = gld(F.slice_off + s) + 6;
    bool failed = false;
    /* all of the slice's tables fit at once: built a single time */
    const bool single = nprec <= np;
    bool built = false;
    /* the tables of precincts i0 .. i0 + np - 1 (the caller synchronises) */
    auto build = [&](uint32_t i0) {
        if (active && i0 + slot < nprec) {
            const uint32_t row = first_row + i0 + slot, type = row + 1 >= rows ? 1 : 0;
            Rec rp;
            if (i0 == 0) {
                rp.a = recs[slot][0][lane];
                rp.b = recs[slot][1][lane];
                rp.c = recs[slot][2][lane];
                rp.e = recs[slot][3][lane];
            } else {
                rp = load_rec(F.hist + (size_t)row * n_bl * JXS_HREC, rec_line);
            }
            rate_lines(tabs[slot], rp, n_bl, g.ln_info[lane], type, g.sig, lane, lg, sub);
        }
    };
    {
        /* one Q, then one R, for the whole slice by the per-precinct binary
         * searches on the slice's summed bytes (RateControl.c:1336-1425):
         * pass 0 sums the (q, 0) columns, pass 1 the (Q, r) rows, pass 2
         * stores every precinct's headers + data bytes at (Q, R) */
        const uint32_t has_last = first_row + nprec >= rows ? 1u : 0u;
        const uint32_t hsum = (nprec - has_last) * g.rc_hdr_bytes[0] + has_last * g.rc_hdr_bytes[1];
        bool ok = sbud > hsum && (sbud + nprec - 1) / nprec <= 1048575u;
        const uint32_t avail = ok ? sbud - hsum : 0u;
        uint32_t Q = 0, R = 0;
        for (uint32_t pass = 0; pass < 3 && ok; pass++) {
            uint32_t acc = 0, dmax = 0;
            for (uint32_t i0 = 0; i0 < nprec; i0 += np) {
                const uint32_t n = min(np, nprec - i0);
                if (!(single && built)) {
                    build(i0);
                    built = true;
                    __syncthreads();
                }
                if (active && slot < n) {
                    const uint32_t row = first_row + i0 + slot, type = row + 1 >= rows ? 1 : 0;
                    const uint32_t rq = g.rq[type][lane], rw = g.rq_raw[type][lane];
                    const uint32_t lo = g.rq_seg[type][lg][sub], hi = g.rq_seg[type][lg][sub + 1];
                    const uint32_t q = pass == 0 ? (lane & (JXS_NQ - 1)) : Q;
                    const uint32_t r = pass == 0 ? 0u : pass == 1 ? (lane & (JXS_NR - 1)) : R;
                    part[wid][lane] = rate_query_seg(rq, rw, lo, hi, tabs[slot], q, r);
                }
                __syncthreads();
                if (wid == 0)
                    for (uint32_t k = 0; k < n; k++) {
                        uint32_t v = 0;
                        for (uint32_t u = 0; u < wpp; u++)
                            v += part[(k << lg) + u][lane];
                        if (pass == 2) {
                            const uint32_t row = first_row + i0 + k, type = row + 1 >= rows ? 1 : 0;
                            v = uni(v) + uni(g.rc_hdr_bytes[type]);
                            if (lane == 0)
                                gst(&F.prec[row].total, v);
                        }
                        acc += v;
                        dmax = max(dmax, v);
                    }
                __syncthreads();
            }
            if (wid == 0) {
                if (pass == 0) {
                    ok = search_q(maxq, eq, acc, avail, Q);
                } else if (pass == 1) {
                    ok = search_r(maxr, g.empty_mask[Q], acc, 0, avail, R);
                } else {
                    /* acc: every lane the slice's headers + data bytes */
                    slice_budgets(F, first_row, nprec, rc_mode, avail + hsum - uni(acc), uni(dmax), lane);
                }
                if (lane == 0)
                    slice_qr = (ok ? 1u : 0u) | Q << 8 | R << 16;
            }
            __syncthreads();
            const uint32_t sp = uni(slice_qr);
            ok = (sp & 1u) != 0;
            Q = (sp >> 8) & 0xFFu;
            R = (sp >> 16) & 0xFFu;
            __syncthreads();
        }
        if (!ok) {
            failed = true;
            if (wid == 0 && lane == 0)
                gor(F.status, 1u);
        }
    }
    const uint32_t lane_k = lane;
    for (uint32_t i0 = 0; i0 < nprec; i0 += np) {
        /* the lane index made opaque each time round: addresses derived from
         * it are formed where used, not kept live across the loop */
        const uint32_t lane = opaque_v(lane_k);
        const uint32_t n = min(np, nprec - i0);
        if (!(single && built)) {
            build(i0);
            built = true;
            __syncthreads();
        }
        const bool mine = active && slot < n;
        const uint32_t i = i0 + slot, row = first_row + i, type = row + 1 >= rows ? 1 : 0;
        const uint32_t lo = mine ? uni(g.rq_seg[type][lg][sub]) : 0u, hi = mine ? uni(g.rq_seg[type][lg][sub + 1]) : 0u;
        /* lane q (and q + 32): the run's bytes at (q, 0) */
        if (mine)
            part[wid][lane] = rate_query_seg(g.rq[type][lane], g.rq_raw[type][lane], lo, hi, tabs[slot],
                                             lane & (JXS_NQ - 1), 0);
        __syncthreads();
        const uint32_t hdr = uni(g.rc_hdr_bytes[type]);
        uint32_t cq = 0, q0 = 0, b0 = 0;
        bool ok0 = false;
        if (mine) {
            for (uint32_t u = 0; u < wpp; u++)
                cq += part[(slot << lg) + u][lane];
            b0 = uni(gld(&F.prec[row].total));
            ok0 = b0 <= 1048575u && b0 > hdr && search_q(maxq, eq, cq, b0 - hdr, q0);
            if (!ok0)
                q0 = 0;
        }
        __syncthreads();
        /* the (q, r) rows at that Q (lanes 0..31: lane r) and at Q - 1
         * (lanes 32..63) */
        if (mine) {
            const uint32_t qc = lane < JXS_NR ? q0 : (q0 > 0 ? q0 - 1 : 0);
            part[wid][lane] = rate_query_seg(g.rq[type][lane], g.rq_raw[type][lane], lo, hi, tabs[slot], qc,
                                             lane & (JXS_NR - 1));
            if (sub == 0) {
                col[slot][lane] = cq;
                if (lane == 0)
                    spec[slot] = q0;
            }
        }
        __syncthreads();
        if (mine && sub == 0) {
            uint32_t cr = 0;
            for (uint32_t u = 0; u < wpp; u++)
                cr += part[(slot << lg) + u][lane];
            rowq[slot][lane] = cr;
            /* the (Q, R) at the budget without a carry, and the smallest
             * entry above it of those the two searches look at */
            uint32_t r0 = 0, data0 = 0, thr = 0;
            if (ok0) {
                const uint32_t avail0 = b0 - hdr, er = uni(g.empty_mask[q0]);
                if (search_r(maxr, er, cr, 0, avail0, r0)) {
                    data0 = lane_of(cr, r0);
                    const bool lo32 = lane < 32;
                    const bool cand_q = lo32 && lane <= maxq && !((eq >> (lane & 31)) & 1u) && cq > avail0;
                    const bool cand_r = lo32 && lane <= maxr && !((er >> (lane & 31)) & 1u) && cr > avail0;
                    thr = wave_min(min(cand_q ? cq : ~0u, cand_r ? cr : ~0u));
                }
            }
            if (lane == 0) {
                sp_r[slot] = r0;
                sp_data[slot] = data0;
                sp_thr[slot] = thr; /* 0: search at the carried budget */
            }
        }
        __syncthreads();
        if (wid == 0) {
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t i = i0 + k, row = first_row + i, type = row + 1 >= rows ? 1 : 0;
                JxsPrec P;
                P.q = P.r = P.total = P.pad = 0;
                P.offset = off;
                P.err = 1;
                if (!failed) {
                    const uint32_t base = uni(gld(&F.prec[row].total));
                    const uint32_t budget = uni(base + carry);
                    const uint32_t hdr = uni(g.rc_hdr_bytes[type]);
                    bool ok = budget <= 1048575u && budget > hdr;
                    uint32_t q = 0, r = 0, data = 0;
                    if (ok) {
                        const uint32_t avail = budget - hdr;
                        if (avail < uni(sp_thr[k])) {
                            q = uni(spec[k]);
                            r = uni(sp_r[k]);
                            data = uni(sp_data[k]);
                        } else {
                            const uint32_t cq = col[k][lane], crs = rowq[k][lane], q0 = uni(spec[k]);
                            ok = search_q(maxq, eq, cq, avail, q);
                            if (ok) {
                                /* the bytes at (q, r) are lane base + r of cr */
                                uint32_t cr = crs, base = 0;
                                if (q0 > 0 && q + 1 == q0) {
                                    base = JXS_NR;
                                } else if (q != q0) {
                                    cr = rate_query(g.rq[type][lane], g.rq_raw[type][lane], g.rq_n[type], tabs[k], q,
                                                    lane & (JXS_NR - 1));
                                }
                                ok = search_r(maxr, uni(g.empty_mask[q]), cr, base, avail, r);
                                if (ok)
                                    data = lane_of(cr, base + r);
                            }
                        }
                        if (ok) {
                            uint32_t pad = avail - data, total = budget;
                            if (i + 1 < nprec) { /* modes 2 and 3 move the padding on */
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
                    if (!ok) {
                        failed = true;
                        if (lane == 0)
                            gor(F.status, 1u);
                    }
                }
                if (lane == 0) {
                    ps[k] = P;
                    gst_words(F.prec + row, P);
                }
            }
        }
        __syncthreads();
        if (mine && sub == 0) {
            const JxsPrec P = ps[slot];
            const uint32_t ui = g.u_info[lane];
            const bool pres = lane < g.n_units && ((ui >> (UI_NORMAL + type)) & 1u);
            const uint32_t o = (ui & 0xFFu) * RT_STRIDE +
                (pres ? truncation((ui >> 8) & 0xFFu, (ui >> 16) & 0xFFu, P.q, P.r) : RT_ZERO);
            const RateTab& T = tabs[slot];
            plan_precinct(F, g, row, type, P, T.s[o], T.g[o], T.d[o], ms_band[slot], ui, lane);
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* packing                                                             */
/* ------------------------------------------------------------------ */
#define PK_WAVES 4
#define RINGA 16       /* words of a wave's significance ring: a band line has at most 256 flags */
#define RINGB 512      /* words of a wave's bit-plane-count ring (2 KiB): a step adds <= 128 */
#define RINGC 1024     /* words of a wave's data ring (4 KiB): a step adds <= 512 */
#define RING_FLUSH 256 /* flush once this many complete words are pending */

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
    /* no segment: finish() writes nothing */
    __device__ __forceinline__ void clear() {
        lo = hi = pos = flushed = 0;
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
    /* Writes ring words [from, to) out and clears them (sh: boundary words
     * into the shadow words, sd bytes after their codestream words, else into
     * the codestream).  A word wholly inside
     * the segment is stored; the segment's first and last words, which it
     * may share with its neighbours, are ORed in with device-scope atomics
     * (the neighbours may run on other XCDs).  A device-scope atomic on a
     * line that another XCD's L2 holds dirty first waits for that line's
     * write-back, so plain stores and atomics must not share lines.  Large
     * launches OR the boundary words into the frame's shadow words, which a
     * later kernel ORs into the codestream, and keep the stores in L2; a
     * single frame's launch, where that kernel's latency would cost more,
     * writes the stores through to memory and ORs into the codestream. */
    __device__ __forceinline__ void flush(uint32_t from, uint32_t to, int64_t sd, uint32_t sh) {
        for (uint32_t w = from + lane; w < to; w += WAVE) {
            uint32_t* slot = &ring[w & mask];
            const uint32_t v = *slot;
            *slot = 0;
            const uint32_t b0 = 32 * w;
            if (b0 >= lo && b0 + 32 <= hi) {
                if (sh)
                    gst(ob + w, __builtin_bswap32(v));
                else
                    __hip_atomic_store((GAS uint32_t*)(ob + w), __builtin_bswap32(v), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            } else if (v) {
                gor(sh ? shadow_of(ob + w, sd) : ob + w, __builtin_bswap32(v));
            }
        }
    }
    __device__ __forceinline__ void maybe_flush(int64_t sd, uint32_t sh) {
        if ((pos >> 5) - flushed >= RING_FLUSH) {
            wave_sync();
            flush(flushed, pos >> 5, sd, sh);
            flushed = pos >> 5;
            wave_sync();
        }
    }
    __device__ __forceinline__ void finish(int64_t sd, uint32_t sh) {
        wave_sync();
        flush(flushed, (hi + 31) / 32, sd, sh);
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


/* Swaps the bits at positions p and p + d of x for every p in m. */
__device__ __forceinline__ uint32_t delta_swap(uint32_t x, uint32_t m, uint32_t d) {
    const uint32_t t = ((x >> d) ^ x) & m;
    return x ^ (t | (t << d));
}

/* Bit-plane nibbles of one group's four 16-bit values r0 .. r3, held as
 * rx = r0 | r1 << 16 and ry = r2 | r3 << 16: nibble j of the result holds
 * bit j of r0, r1, r2, r3 from its MSB down, for planes 0..7 (sel
 * PLANES_LO) or 8..15 (PLANES_HI).  One byte permute gathers the byte of
 * each value into byte 3, 1, 2, 0 (the value index bit-reversed), which makes
 * the wanted layout a rotation of the bit address [row:2 | plane:3] to
 * [plane:3 | row:2]: three delta swaps exchange address bits 4 and 2, 3 and
 * 1, 2 and 0. */
#define PLANES_LO 0x04000602u
#define PLANES_HI 0x05010703u
__device__ __forceinline__ uint32_t planes8(uint32_t rx, uint32_t ry, uint32_t sel) {
    uint32_t w = __builtin_amdgcn_perm(rx, ry, sel);
    w = delta_swap(w, 0x0000F0F0u, 12);
    w = delta_swap(w, 0x00CC00CCu, 6);
    return delta_swap(w, 0x0A0A0A0Au, 3);
}

#define IPL 4 /* consecutive items coded by one lane per step */

/* Significance sub-packet segment of one band line: one bit per group of
 * 8 GCLIs, 1 when its maximum is <= gt (PackPrecinct.c, significance
 * flags).  Lane L codes flags 4L .. 4L+3 from the maxima sf[] it loaded;
 * a line has at most 256 significance groups (one step). */
__device__ __forceinline__ void sig_step(WaveOut& W, const uint32_t (&sf)[IPL], uint32_t sw, uint32_t gt) {
    const uint32_t i0 = IPL * W.lane;
    uint32_t nb = 0;
    uint64_t v = 0;
#pragma unroll
    for (uint32_t k = 0; k < IPL; k++)
        if (i0 + k < sw) {
            v = (v << 1) | (sf[k] <= gt ? 1u : 0u);
            nb++;
        }
    const uint32_t incl = wave_incl_scan(nb, W.lane);
    if (nb && v)
        W.put(W.pos + incl - nb, v, nb);
    W.pos += wave_last(incl);
}

/* The inputs of one step of a band line's bit-plane-count and data
 * segments for one lane: groups i0 .. i0+N-1 (N divides 8, i0 a multiple of
 * N): GCLIs, coefficients, and the maximum of their significance group (the
 * same for all N).  Lanes past the line's end load the last group's
 * neighbourhood; the groups of a lane past the end (at most N - 1 beyond the
 * line's last group) are read but not coded: the GCLI and coefficient
 * buffers continue past every band line (later lines, then the next buffers
 * of the frame's work block). */
template <uint32_t N>
struct GdIn {
    uint32_t a[N], s8;
    uint2 cv[N];
};
template <uint32_t N>
__device__ __forceinline__ GdIn<N> gd_load(const uint8_t* gl, const uint8_t* sl, const uint16_t* cl, uint32_t gw,
                                           uint32_t ms, uint32_t i0) {
    GdIn<N> in;
    const uint32_t b = min(i0, gw - 1);
    const uint8_t* g = gl + b;
    const uint16_t* c = cl + 4 * b;
#pragma unroll
    for (uint32_t k = 0; k < N; k++) {
        in.a[k] = gld(g + k);
        in.cv[k] = gld2(c + 4 * k);
    }
    in.s8 = ms ? gld(sl + b / 8) : 0u;
    return in;
}

/* One step of the bit-plane-count and data segments, 64 * N groups: lane L
 * codes groups i0 .. i0+N-1 (i0 = base + N * L).  Bit-plane counts are unary
 * (gcli - gt ones and a zero, none for a group inside an insignificant
 * significance group) or raw 4-bit GCLIs; a lane's codes are concatenated
 * (at most 64 bits).  Data: per coded group a sign nibble, then bit-planes
 * gcli-1 .. gt of the quantised magnitudes, 4 bits each; groups no lane
 * codes are skipped.  The coefficients of a group arrive as two words of
 * two 16-bit sign-magnitude values; quantisation and signs work on both
 * halves at once.  With np = gcli - gt, r = q | sign << np (the sign of a
 * nonzero quantised value) makes planes np .. 0 of r the sign nibble and the
 * bit-planes in order, so one transpose gives the group's whole code.  A
 * lane's data codes are concatenated up to 64 bits per write.  Coefficients
 * past a band line's end are zero in the buffer, so they quantise to zero. */
template <uint32_t N, bool UNIFORM>
__device__ __forceinline__ void gd_step(WaveOut& B, WaveOut& C, const GdIn<N>& in, uint32_t i0, uint32_t gw,
                                        uint32_t gt, uint32_t ms, uint32_t raw) {
    const uint32_t lane = B.lane;
    /* bit-plane counts */
    const uint32_t rem = gw - min(i0, gw); /* groups of this lane inside the line, capped below */
    uint32_t nb = 0;
    uint64_t v = 0;
    if (raw) {
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            const uint32_t len = k < rem ? 4u : 0u;
            v = (v << len) | (k < rem ? in.a[k] : 0u);
            nb += len;
        }
    } else {
        /* a group of an insignificant significance group has no code */
        const bool sig_skip = ms && in.s8 <= gt;
#pragma unroll
        for (uint32_t k = 0; k < N; k++) {
            const bool skip = sig_skip || k >= rem;
            const uint32_t m = max(in.a[k], gt) - gt;
            const uint32_t len = skip ? 0u : m + 1;
            v = (v << len) | (skip ? 0u : (2u << m) - 2);
            nb += len;
        }
    }
    const uint32_t incl = wave_incl_scan(nb, lane);
    if (nb && v)
        B.put(B.pos + incl - nb, v, nb);
    B.pos += wave_last(incl);
    /* data lengths first: the lane's position comes from their scan */
    uint32_t dl[N], dn = 0;
#pragma unroll
    for (uint32_t k = 0; k < N; k++) {
        const bool coded = k < rem && in.a[k] > gt;
        dl[k] = coded ? 4 * (in.a[k] - gt) + 4 : 0u;
        dn += dl[k];
    }
    const uint32_t dincl = wave_incl_scan(dn, lane);
    uint32_t at = C.pos + dincl - dn;
    const uint32_t gtt = gt * 0x10001u;
    uint64_t acc = 0;
    uint32_t an = 0;
#pragma unroll
    for (uint32_t k = 0; k < N; k++) {
        if (__ballot(dl[k] != 0) == 0)
            continue;
        const uint32_t gc = in.a[k];
        /* magnitudes; zero for a group this lane does not code */
        const uint32_t mx = dl[k] ? in.cv[k].x & 0x7FFF7FFFu : 0u;
        const uint32_t my = dl[k] ? in.cv[k].y & 0x7FFF7FFFu : 0u;
        uint32_t qx, qy;
        if (UNIFORM) {
            qx = quant_mag<true>(mx & 0xFFFFu, gc, gt) | (quant_mag<true>(mx >> 16, gc, gt) << 16);
            qy = quant_mag<true>(my & 0xFFFFu, gc, gt) | (quant_mag<true>(my >> 16, gc, gt) << 16);
        } else {
            qx = as_u32(as_u16x2(mx) >> as_u16x2(gtt));
            qy = as_u32(as_u16x2(my) >> as_u16x2(gtt));
        }
        /* sign bit 15 moved down to bit np, kept where q != 0: q + 0x7FFF
         * reaches bit 15 exactly when q > 0 (q < 2^15, so no carry crosses
         * into the other half) */
        const u16x2 rs = as_u16x2((15 + gt - gc) * 0x10001u);
        const uint32_t nzx = (qx + 0x7FFF7FFFu) & 0x80008000u, nzy = (qy + 0x7FFF7FFFu) & 0x80008000u;
        const uint32_t rx = qx | as_u32(as_u16x2(in.cv[k].x & nzx) >> rs);
        const uint32_t ry = qy | as_u32(as_u16x2(in.cv[k].y & nzy) >> rs);
        uint64_t dv = planes8(rx, ry, PLANES_LO);
        if (__ballot(dl[k] > 32) != 0)
            dv |= (uint64_t)planes8(rx, ry, PLANES_HI) << 32;
        if (dl[k]) {
            if (an + dl[k] > 64) {
                if (acc)
                    C.put(at, acc, an);
                at += an;
                acc = 0;
                an = 0;
            }
            acc = an ? (acc << dl[k]) | dv : dv;
            an += dl[k];
        }
    }
    if (an && acc)
        C.put(at, acc, an);
    C.pos += wave_last(dincl);
}

/* What a packing wave loads first for a unit: the significance maxima it
 * codes (IPL per lane) and the first step's GCLIs, maxima and coefficients
 * (absent segments load nothing). */
struct PackIn {
    uint32_t sf[IPL];
    GdIn<IPL> gd;
};
/* The fields of a unit those loads need, read on their own (scalar loads)
 * so that the next unit's record is not held in registers while this one is
 * coded. */
struct PackSrc {
    const uint8_t* gl;
    const uint8_t* sl;
    const uint16_t* cl;
    uint32_t len[3], gw, sw, ms;
};
__device__ __forceinline__ PackSrc pack_src(const JxsUnit* u) {
    const CAS JxsUnit* c = (const CAS JxsUnit*)u;
    PackSrc s;
    s.gl = c->gl;
    s.sl = c->sl;
    s.cl = c->cl;
    s.len[0] = c->len[0];
    s.len[1] = c->len[1];
    s.len[2] = c->len[2];
    s.gw = c->gw;
    s.sw = c->sw;
    s.ms = c->ms;
    return s;
}
__device__ __forceinline__ PackIn pack_load(const PackSrc& U, uint32_t lane) {
    PackIn in;
#pragma unroll
    for (uint32_t k = 0; k < IPL; k++)
        in.sf[k] = 0;
    if (U.len[0]) {
#pragma unroll
        for (uint32_t k = 0; k < IPL; k++)
            in.sf[k] = gld(U.sl + min(IPL * lane + k, U.sw - 1));
    }
    if (U.len[1] | U.len[2])
        in.gd = gd_load<IPL>(U.gl, U.sl, U.cl, U.gw, U.ms, IPL * lane);
    return in;
}

/* s_waitcnt vmcnt(0) alone (expcnt and lgkmcnt fields at their no-wait
 * maxima): every vector memory access of the wave has completed.  On this
 * architecture stores and loads share the counter and complete in order, so
 * waiting for a load also waits for every store issued before it; the packer
 * places this wait where all outstanding accesses were issued one unit (or
 * one step) of coding earlier. */
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

/* The packet header (PackHeaders.c:98-127, sign size 0) of the first band
 * line of a packet, ORed into the zeroed output. */
__device__ __forceinline__ void put_header(const CAS JxsUnit& U, uint32_t lane, uint32_t shadow) {
    if (U.hlen) {
        uint32_t w0, w1, w2;
        split_bits(U.hdr, U.hlen, U.hlo, w0, w1, w2);
        const uint32_t piece = lane == 0 ? w0 : lane == 1 ? w1 : w2;
        if (lane < 3 && piece)
            gor(shadow ? shadow_of(U.hob + lane, U.sdelta) : U.hob + lane, __builtin_bswap32(piece));
    }
}

/* The significance segment of a band line into ring A (written out later).
 * Returns whether its length differs from the plan's. */
__device__ __forceinline__ bool code_sig(const CAS JxsUnit& U, const PackIn& in, uint32_t lane, WaveOut& A,
                                         uint32_t* ringa) {
    if (!U.len[0]) {
        A.clear();
        return false;
    }
    A.start(ringa, RINGA - 1, U.ob[0], U.lo[0], U.len[0], lane);
    wave_sync();
    sig_step(A, in.sf, U.sw, U.gt);
    return A.pos != A.hi;
}

/* The bit-plane-count and data segments of a band line into rings B and C,
 * IPL groups per lane and step, each step's successor loaded before the step
 * is coded.  From the second step on, a step starts by waiting for its
 * inputs and writing out the words completed so far (when many are pending),
 * then issues the next step's loads.  The segments' remaining words are
 * written out later.  Returns whether a segment's length differs from the
 * plan's. */
template <bool UNIFORM>
__device__ __forceinline__ bool code_gd(const CAS JxsUnit& U, const GdIn<IPL>& first, uint32_t lane, WaveOut& B,
                                        WaveOut& C, uint32_t* ringb, uint32_t* ringc, uint32_t shadow) {
    if (!(U.len[1] | U.len[2])) {
        B.clear();
        C.clear();
        return false;
    }
    const uint8_t* gl = U.gl;
    const uint8_t* sl = U.sl;
    const uint16_t* cl = U.cl;
    const uint32_t gw = U.gw, gt = U.gt, ms = U.ms, raw = U.raw;
    B.start(ringb, RINGB - 1, U.ob[1], U.lo[1], U.len[1], lane);
    C.start(ringc, RINGC - 1, U.ob[2], U.lo[2], U.len[2], lane);
    const int64_t sd = U.sdelta;
    wave_sync();
    GdIn<IPL> cur = first;
    for (uint32_t base = 0; base < gw; base += IPL * WAVE) {
        const bool more = base + IPL * WAVE < gw;
        if (base) {
            vm_wait_all();
            B.maybe_flush(sd, shadow);
            C.maybe_flush(sd, shadow);
        }
        GdIn<IPL> nxt;
        if (more)
            nxt = gd_load<IPL>(gl, sl, cl, gw, ms, base + IPL * WAVE + IPL * lane);
        gd_step<IPL, UNIFORM>(B, C, cur, base + IPL * lane, gw, gt, ms, raw);
        if (more)
            cur = nxt;
    }
    return B.pos != B.hi || C.pos != C.hi;
}

/* Persistent packing waves (PackPrecinct.c:235-572): each wave codes the
 * packing units (band lines of precincts) gu, gu + W, gu + 2W, ... (W =
 * waves of the launch; units ordered widest first), every one's
 * significance flags, bit-plane counts (unary or raw) and data bit-planes
 * with the quantisation (Quant.c:9-45) fused, each segment into its own
 * sub-packet, plus the packet header of a packet's first band line.  A unit
 * is coded in the wave's LDS rings and written out at the start of the next
 * unit's turn: wait for the next unit's inputs (loaded while this one was
 * coded), write out the previous unit's segments and this unit's header,
 * issue the loads for the unit after, then code from registers.  So every
 * wait covers loads and stores issued a whole unit earlier.  Waves never
 * wait on each other. */
#define PK_BLOCKS 1280 /* resident blocks: 256 CUs x 5 (5 waves per SIMD: <= 96 VGPRs, 24.3 KiB LDS per block) */
__global__ void __launch_bounds__(64 * PK_WAVES) __attribute__((amdgpu_waves_per_eu(5)))
k_pack(const JxsUnit* __restrict__ units, uint32_t total_units, uint32_t shadow) {
    __shared__ uint32_t ringa[PK_WAVES][RINGA];
    __shared__ uint32_t ringb[PK_WAVES][RINGB];
    __shared__ uint32_t ringc[PK_WAVES][RINGC];
    const uint32_t wid = wave_uniform(threadIdx.x / WAVE), lane = threadIdx.x & (WAVE - 1);
    const uint32_t stride = gridDim.x * PK_WAVES;
    uint32_t gu = blockIdx.x * PK_WAVES + wid;
    if (gu >= total_units)
        return;
    WaveOut A, B, C; /* the previous unit's segments, not yet written out */
    int64_t psd = 0; /* the previous unit's shadow offset */
    A.clear();
    B.clear();
    C.clear();
    A.lane = B.lane = C.lane = lane;
    PackIn in = pack_load(pack_src(units + gu), lane);
    for (;;) {
        const CAS JxsUnit& U = *(const CAS JxsUnit*)(units + gu);
        const uint32_t gn = gu + stride;
        const bool more = gn < total_units;
        PackSrc Un;
        if (more)
            Un = pack_src(units + gn);
        vm_wait_all();
        A.finish(psd, shadow);
        B.finish(psd, shadow);
        C.finish(psd, shadow);
        wave_sync();
        put_header(U, lane, shadow);
        PackIn inn;
        if (more)
            inn = pack_load(Un, lane);
        bool bad = code_sig(U, in, lane, A, ringa[wid]);
        bad |= U.qpih ? code_gd<true>(U, in.gd, lane, B, C, ringb[wid], ringc[wid], shadow)
                      : code_gd<false>(U, in.gd, lane, B, C, ringb[wid], ringc[wid], shadow);
        /* the plan's lengths are the segments' exact bit counts */
        if (bad && lane == 0)
            gor(U.status, 2u);
        psd = U.sdelta;
        if (!more)
            break;
        gu = gn;
        in = inn;
    }
    A.finish(psd, shadow);
    B.finish(psd, shadow);
    C.finish(psd, shadow);
}

/* The codestream dwords packing units share (a segment's first and last
 * dwords, a packet header's dwords): their bits are in the frame's shadow
 * words and are ORed into the codestream here, which already holds the
 * rate kernel's bytes (headers) and zeros elsewhere.  Units sharing a dword
 * each OR the same complete shadow word in: the result is the same whichever
 * goes first.  JOIN_SLOTS threads per unit, one per candidate dword (slots
 * 0..2: first dword of segment 0..2, 3..5: their last dwords, 6..8: the
 * packet header's dwords), so every thread reads its shadow and codestream
 * words together. */
#define JOIN_SLOTS 9 /* 16 (a shift for the unit index) left 7 of every 16 threads idle and ran 8 % longer */
__global__ void __launch_bounds__(256) k_join(const JxsUnit* __restrict__ units, uint32_t total_units) {
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t u = gid / JOIN_SLOTS, k = gid % JOIN_SLOTS;
    if (u >= total_units)
        return;
    const GAS JxsUnit& U = ((const GAS JxsUnit*)units)[u];
    uint32_t* p = nullptr;
    if (k < 6) {
        const uint32_t sgm = k < 3 ? k : k - 3, lo = U.lo[sgm], len = U.len[sgm];
        if (len) {
            const uint32_t last = (lo + len - 1) >> 5;
            if (k < 3 && (lo != 0 || len < 32))
                p = U.ob[sgm];
            else if (k >= 3 && last > 0 && ((lo + len) & 31) != 0)
                p = U.ob[sgm] + last;
        }
    } else {
        const uint32_t hlen = U.hlen;
        if (hlen && k - 6 <= ((U.hlo + hlen - 1) >> 5))
            p = U.hob + (k - 6);
    }
    if (!p)
        return;
    const uint32_t v = gld(shadow_of(p, U.sdelta)), o = gld(p);
    if (v)
        gst(p, o | v);
}

/* ------------------------------------------------------------------ */
/* launchers (C linkage for the host code)                             */
/* ------------------------------------------------------------------ */
extern "C" {

/* jobs: all of one chunk width: 8 columns per thread in 512-thread blocks,
 * 16 or 32 in 256-thread blocks */
hipError_t jxs_launch_dwt(const JxsDwtJob* jobs, const uint32_t* bases, int njobs, uint32_t blocks, uint32_t chunk,
                          uint32_t* status, uint32_t nstatus, hipStream_t st) {
    if (!njobs || !blocks)
        return hipSuccess;
    if (chunk == 8)
        hipLaunchKernelGGL((k_dwt<8, 512>), dim3(blocks), dim3(512), 0, st, jobs, bases, njobs, blocks, status, nstatus);
    else if (chunk == 16)
        hipLaunchKernelGGL((k_dwt<16, 256>), dim3(blocks), dim3(256), 0, st, jobs, bases, njobs, blocks, status, nstatus);
    else if (chunk == 32)
        hipLaunchKernelGGL((k_dwt<32, 256>), dim3(blocks), dim3(256), 0, st, jobs, bases, njobs, blocks, status, nstatus);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* rate control modes 0 and 1: one block per precinct for the rate      */
/* table and its searches (k_rate_tab), then one per precinct for the    */
/* in-order pass over its slice and its packet plan (k_rate_plan)       */
/* ------------------------------------------------------------------ */

/* The frame of precinct p of a launch (frames in precinct order). */
__device__ __forceinline__ int frame_of_prec(const JxsFrameDesc* __restrict__ frames, int nfr, uint32_t p) {
    int fl = 0, fh = nfr - 1;
    while (fl < fh) {
        const int mid = (fl + fh + 1) >> 1;
        if (frames[mid].prec_base <= p)
            fl = mid;
        else
            fh = mid - 1;
    }
    return fl;
}

/* The budget of precinct i of a slice of nprec under rate control mode 0
 * (an even split, PackStageProcess.c:600-601) or 1 (the first precinct 10 %
 * more, the rest split evenly, 621-640), before any padding carry. */
__device__ __forceinline__ uint32_t prec_budget(uint32_t sbud, uint32_t nprec, uint32_t i, uint32_t rc_mode) {
    uint32_t minb = sbud / nprec;
    uint32_t first = minb, left = sbud - minb * nprec;
    if (rc_mode == 1) {
        first = minb * 110 / 100;
        if (first > sbud)
            first = sbud;
        left = 0;
        if (nprec > 1) {
            minb = (sbud - first) / (nprec - 1);
            left = sbud - minb * (nprec - 1) - first;
        }
    }
    return (i == 0 ? first : minb) + (i < left ? 1u : 0u);
}

/* Byte k of a read-only descriptor, through a scalar load of its dword. */
__device__ __forceinline__ uint32_t cld_u8(const uint8_t* p) {
    const uintptr_t a = (uintptr_t)p;
    const uint32_t w = *(const CAS uint32_t*)(a & ~(uintptr_t)3);
    return (w >> (8 * (a & 3))) & 0xFFu;
}
__device__ __forceinline__ uint32_t cld_u32(const uint32_t* p) { return *(const CAS uint32_t*)p; }

/* The rate-table entry of band line `lane` at truncation t alone (t >= 16:
 * the line is absent from the precinct, zeros), from its histogram record:
 * the sums of rate_lines_part at one t.  The other line of a two-line band
 * (the neighbouring lane) must ask for the same t; every lane takes part. */
__device__ __forceinline__ void rate_line_at(const Rec& rc, uint32_t n_bl, uint32_t li, uint32_t type,
                                             uint32_t sig_on, uint32_t lane, uint32_t t, uint32_t& od, uint32_t& og,
                                             uint32_t& os) {
    const uint32_t w[16] = {rc.a.x, rc.a.y, rc.a.z, rc.a.w, rc.b.x, rc.b.y, rc.b.z, rc.b.w,
                            rc.c.x, rc.c.y, rc.c.z, rc.c.w, rc.e.x, rc.e.y, rc.e.z, rc.e.w};
    uint32_t tot = 0, vtot = 0, cnt = 0, vcnt = 0, sig = 0;
#pragma unroll
    for (uint32_t v = 0; v < 16; v++) {
        const uint32_t h = (w[v >> 1] >> (16 * (v & 1))) & 0xFFFFu;
        tot += h;
        vtot += v * h;
        if (v <= t) {
            cnt += h;
            vcnt += v * h;
            if (v < 15)
                sig += (w[8 + (v >> 1)] >> (16 * (v & 1))) & 0xFFFFu;
        }
    }
    const uint32_t gw = li & 0xFFFu, full = gw >> 3, part = gw & 7, sw = (li >> LI_SW) & 0xFFu;
    if (t >= 15)
        sig = full;
    const uint32_t nl = (li >> (LI_NL + 2 * type)) & 3u;
    const bool two = nl == 2;
    const int partner = (int)(two ? (((li >> LI_SECOND) & 1u) ? lane - 1 : lane + 1) : lane);
    const uint32_t tail = w[15] >> 16; /* 0xFFFF: no partial group */
    const uint32_t tt = min(t, 15u);
    const uint32_t above = tot - cnt, vabove = vtot - vcnt;
    const uint32_t data = vabove + above - tt * above; /* sum_{v>t} (v - t + 1) h(v) */
    const uint32_t ns = data + cnt, zr = 8 * sig + (tail <= tt ? part : 0u);
    const uint32_t pn = (uint32_t)__shfl((int)ns, partner), pz = (uint32_t)__shfl((int)zr, partner);
    const uint32_t nosig = ns + (two ? pn : 0u);
    const uint32_t withsig = sw * nl + (ns - zr) + (two ? pn - pz : 0u);
    const bool use = sig_on && nosig > withsig;
    const bool absent = t >= 16 || lane >= n_bl;
    od = absent ? 0u : 4 * data;
    og = absent ? 0u : (use ? ns - zr : ns);
    os = absent ? 0u : (use ? sw : 0u);
}

/* One block per precinct of NW waves (1, 2 or 4: the autotuner's choice):
 * its rate table (1/NW of the truncations a wave), its bytes at (q, 0) for
 * every q and the (q, r) rows at the Q its budget gets without a padding
 * carry and at that Q - 1 (a run of whole packets a wave, the runs summed),
 * the binary searches (RateControl.c:889-1083, BinarySearch.c) at that
 * budget, and the smallest entry above the budget of those the searches look
 * at: all into the precinct's rate record. */
}  // extern "C"

template <uint32_t NW>
__global__ void __launch_bounds__(WAVE * NW) k_rate_tab(const JxsFrameDesc* __restrict__ frames, int nfr,
                                                        uint32_t total_prec) {
    constexpr uint32_t LG = NW == 4 ? 2u : NW == 2 ? 1u : 0u;
    __shared__ RateTab T;
    __shared__ uint32_t part[2][NW][WAVE];
    const uint32_t pg = blockIdx.x;
    if (pg >= total_prec)
        return;
    const JxsFrameDesc F = cld_words(frames + frame_of_prec(frames, nfr, pg));
    if (F.ordered || F.slice_rc)
        return; /* coded by k_enc_slice, or by k_rate */
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const uint32_t row = pg - F.prec_base, rows = F.rows, pps = F.pps, n_bl = F.n_bl;
    const uint32_t s = row / pps, i = row - s * pps, first_row = s * pps;
    const uint32_t nprec = (s == F.slices - 1) ? rows - first_row : pps;
    const uint32_t type = row + 1 >= rows ? 1u : 0u;
    const JxsGeom* G = F.g;
    /* every load the block needs, issued together */
    const Rec rc = load_rec(F.hist + (size_t)row * n_bl * JXS_HREC, min(lane, n_bl - 1));
    const uint32_t li = gld(&G->ln_info[lane]);
    const uint32_t rq = gld(&G->rq[type][lane]), rw = gld(&G->rq_raw[type][lane]);
    const uint32_t em = gld(&G->empty_mask[lane & (JXS_NQ - 1)]);
    const uint32_t sig = cld_u32(&G->sig), maxq = cld_u32(&G->maxq), maxr = cld_u32(&G->maxr);
    const uint32_t rc_mode = cld_u32(&G->rc_mode), hdr = cld_u32(&G->rc_hdr_bytes[type]);
    const uint32_t lo = cld_u8(&G->rq_seg[type][LG][w]), hi = cld_u8(&G->rq_seg[type][LG][w + 1]);
    const uint32_t b0 = prec_budget(cld_u32(F.slice_bud + s), nprec, i, rc_mode);
    const uint32_t eq = (uint32_t)__ballot(em & 1u); /* bit q: (q, 0) truncates every band away */
    rate_lines(T, rc, n_bl, li, type, sig, lane, LG, w);
    lds_barrier();
    /* lane q (and q + 32): the run's bytes at (q, 0) */
    part[0][w][lane] = rate_query_seg(rq, rw, lo, hi, T, lane & (JXS_NQ - 1), 0);
    lds_barrier();
    uint32_t cq = 0;
#pragma unroll
    for (uint32_t k = 0; k < NW; k++)
        cq += part[0][k][lane];
    uint32_t q0 = 0;
    const bool ok0 = b0 <= 1048575u && b0 > hdr && search_q(maxq, eq, cq, b0 - hdr, q0);
    if (!ok0)
        q0 = 0;
    /* the (q, r) rows at that Q (lanes 0..31: lane r) and at Q - 1 (lanes 32..63) */
    const uint32_t qc = lane < JXS_NR ? q0 : (q0 > 0 ? q0 - 1 : 0);
    part[1][w][lane] = rate_query_seg(rq, rw, lo, hi, T, qc, lane & (JXS_NR - 1));
    lds_barrier();
    if (w != 0)
        return;
    uint32_t cr = 0;
#pragma unroll
    for (uint32_t k = 0; k < NW; k++)
        cr += part[1][k][lane];
    uint32_t r0 = 0, data0 = 0, thr = 0;
    if (ok0) {
        const uint32_t avail0 = b0 - hdr, er = lane_of(em, q0);
        if (search_r(maxr, er, cr, 0, avail0, r0)) {
            data0 = lane_of(cr, r0);
            const bool lo32 = lane < 32;
            const bool cand_q = lo32 && lane <= maxq && !((eq >> (lane & 31)) & 1u) && cq > avail0;
            const bool cand_r = lo32 && lane <= maxr && !((er >> (lane & 31)) & 1u) && cr > avail0;
            thr = wave_min(min(cand_q ? cq : ~0u, cand_r ? cr : ~0u));
        }
    }
    JxsRateRec* R = F.rrec + row;
    if (lane < 32)
        gst(&R->col[lane], cq);
    gst(&R->row[lane], cr);
    if (lane == 0) {
        const u32x4 sp = {q0, r0, data0, thr};
        gst((u32x4*)&R->q0, sp);
    }
}

extern "C" {

/* One wave per precinct: the in-order pass over its slice up to it, which
 * fixes its budget, (Q, R), padding and codestream offset (the budgets with
 * the padding carry, PackStageProcess.c:600-684), then its packet plan.  A
 * precinct's (Q, R) is the one its rate record notes while the carried
 * budget stays below the noted entry; otherwise the searches run again on
 * the record's column and rows, and on a row computed from a rate table
 * rebuilt here when the carry moved Q by two steps or more. */
__global__ void __launch_bounds__(WAVE) k_rate_plan(const JxsFrameDesc* __restrict__ frames, int nfr,
                                                    uint32_t total_prec) {
    __shared__ __align__(16) JxsGeom sg;
    __shared__ RateTab T;
    __shared__ uint32_t tl[WAVE], vd_l[WAVE], vg_l[WAVE], vs_l[WAVE];
    __shared__ uint32_t ms_band[JXS_NGB];
    const uint32_t pg = blockIdx.x;
    if (pg >= total_prec)
        return;
    const JxsFrameDesc F = cld_words(frames + frame_of_prec(frames, nfr, pg));
    if (F.ordered || F.slice_rc)
        return;
    const uint32_t lane = threadIdx.x;
    const uint32_t row = pg - F.prec_base, rows = F.rows, pps = F.pps, n_bl = F.n_bl;
    const uint32_t s = row / pps, i = row - s * pps, first_row = s * pps;
    const uint32_t nprec = (s == F.slices - 1) ? rows - first_row : pps;
    const uint32_t type = row + 1 >= rows ? 1u : 0u;
    const uint32_t rec_line = min(lane, n_bl - 1);
    const Rec rc = load_rec(F.hist + (size_t)row * n_bl * JXS_HREC, rec_line);
    /* with the geometry: the noted (Q, R) of the slice's first precincts (up
     * to 64) and the columns and rows of its first 4 (a carry past the noted
     * entry then needs no further load there) */
    const JxsRateRec* RR = F.rrec + first_row;
    u32x4 sp = {0u, 0u, 0u, 0u};
    if (lane <= i)
        sp = gld((const u32x4*)&RR[lane].q0);
    uint32_t pcol[4], prow[4];
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        pcol[j] = j <= i ? gld(&RR[j].col[lane & (JXS_NQ - 1)]) : 0u;
        prow[j] = j <= i ? gld(&RR[j].row[lane]) : 0u;
    }
    stage_geom<WAVE>(&sg, F.g);
    const uint32_t sbud = cld_u32(F.slice_bud + s);
    uint32_t off = cld_u32(F.slice_off + s) + 6;
    __syncthreads();
    const JxsGeom& g = sg;
    const uint32_t maxq = uni(g.maxq), maxr = uni(g.maxr), rc_mode = uni(g.rc_mode);
    const uint32_t eq = (uint32_t)__ballot(g.empty_mask[lane & (JXS_NQ - 1)] & 1u);
    const bool move_pad = rc_mode != 0;
    /* the slice budget's split (prec_budget's) */
    uint32_t minb = sbud / nprec;
    uint32_t first = minb, left = sbud - minb * nprec;
    if (rc_mode == 1) {
        first = minb * 110 / 100;
        if (first > sbud)
            first = sbud;
        left = 0;
        if (nprec > 1) {
            minb = (sbud - first) / (nprec - 1);
            left = sbud - minb * (nprec - 1) - first;
        }
    }
    uint32_t carry = 0;
    bool failed = false;
    JxsPrec P;
    P.q = P.r = P.total = P.pad = 0;
    P.offset = off;
    P.err = 1;
    for (uint32_t k0 = 0; k0 <= i; k0 += WAVE) {
        /* the noted (Q, R) of precincts k0 + lane */
        if (k0 > 0)
            sp = k0 + lane <= i ? gld((const u32x4*)&RR[k0 + lane].q0) : u32x4{0u, 0u, 0u, 0u};
        const uint32_t kn = min((uint32_t)WAVE, i + 1 - k0);
        for (uint32_t kk = 0; kk < kn; kk++) {
            const uint32_t k = k0 + kk, rk = first_row + k, tk = rk + 1 >= rows ? 1u : 0u;
            JxsPrec Q;
            Q.q = Q.r = Q.total = Q.pad = 0;
            Q.offset = off;
            Q.err = 1;
            if (!failed) {
                const uint32_t budget = (k == 0 ? first : minb) + (k < left ? 1u : 0u) + carry;
                const uint32_t hdr = uni(g.rc_hdr_bytes[tk]);
                bool ok = budget <= 1048575u && budget > hdr;
                uint32_t q = 0, r = 0, data = 0;
                if (ok) {
                    const uint32_t avail = budget - hdr;
                    const uint32_t q0 = lane_of(sp.x, kk);
                    if (avail < lane_of(sp.w, kk)) {
                        q = q0;
                        r = lane_of(sp.y, kk);
                        data = lane_of(sp.z, kk);
                    } else {
                        uint32_t cq, cr, base = 0;
                        if (k < 4) {
                            cq = k == 0 ? pcol[0] : k == 1 ? pcol[1] : k == 2 ? pcol[2] : pcol[3];
                            cr = k == 0 ? prow[0] : k == 1 ? prow[1] : k == 2 ? prow[2] : prow[3];
                        } else {
                            cq = gld(&RR[k].col[lane & (JXS_NQ - 1)]);
                            cr = gld(&RR[k].row[lane]);
                        }
                        ok = search_q(maxq, eq, cq, avail, q);
                        if (ok) {
                            /* the bytes at (q, r) are lane base + r of cr */
                            if (q0 > 0 && q + 1 == q0) {
                                base = JXS_NR;
                            } else if (q != q0) {
                                const Rec rp = load_rec(F.hist + (size_t)rk * n_bl * JXS_HREC, rec_line);
                                rate_lines(T, rp, n_bl, g.ln_info[lane], tk, g.sig, lane, 0, 0);
                                wave_sync();
                                cr = rate_query(g.rq[tk][lane], g.rq_raw[tk][lane], g.rq_n[tk], T, q,
                                                lane & (JXS_NR - 1));
                                wave_sync();
                            }
                            ok = search_r(maxr, uni(g.empty_mask[q]), cr, base, avail, r);
                            if (ok)
                                data = lane_of(cr, base + r);
                        }
                    }
                    if (ok) {
                        uint32_t pad = avail - data, total = budget;
                        if (move_pad && k + 1 < nprec) {
                            total -= pad;
                            carry = pad;
                            pad = 0;
                        }
                        Q.q = q;
                        Q.r = r;
                        Q.total = total;
                        Q.pad = pad;
                        Q.err = 0;
                        off += total;
                    }
                }
                if (!ok) {
                    failed = true;
                    if (lane == 0)
                        gor(F.status, 1u);
                }
            }
            if (k == i)
                P = Q;
        }
    }
    if (lane == 0)
        gst_words(F.prec + row, P);
    /* every unit's bits at its band's truncation, computed per band line */
    const uint32_t ui = g.u_info[lane];
    const bool pres = lane < g.n_units && ((ui >> (UI_NORMAL + type)) & 1u);
    tl[lane] = 16u;
    wave_sync();
    if (lane < g.n_units)
        tl[ui & 0xFFu] = pres ? truncation((ui >> 8) & 0xFFu, (ui >> 16) & 0xFFu, P.q, P.r) : 16u;
    wave_sync();
    uint32_t d, gb, sb;
    rate_line_at(rc, n_bl, g.ln_info[lane], type, g.sig, lane, tl[lane], d, gb, sb);
    vd_l[lane] = d;
    vg_l[lane] = gb;
    vs_l[lane] = sb;
    wave_sync();
    const uint32_t line = ui & 0xFFu & (WAVE - 1);
    plan_precinct(F, g, row, type, P, vs_l[line], vg_l[line], vd_l[line], ms_band, ui, lane);
}

/* kinds: bit 0 launches the rate control of the frames under modes 0 / 1
 * (k_rate_tab with tab_waves waves per precinct, then k_rate_plan), bit 1
 * that of the frames under modes 2 / 3 (k_rate); each skips the other's
 * frames. */
hipError_t jxs_launch_rate(const JxsFrameDesc* frames, int nfr, uint32_t slices, uint32_t precincts, uint32_t kinds,
                           uint32_t tab_waves, hipStream_t st) {
    if (!slices)
        return hipSuccess;
    if (kinds & 1u) {
        if (tab_waves == 1)
            hipLaunchKernelGGL(k_rate_tab<1>, dim3(precincts), dim3(WAVE), 0, st, frames, nfr, precincts);
        else if (tab_waves == 2)
            hipLaunchKernelGGL(k_rate_tab<2>, dim3(precincts), dim3(WAVE * 2), 0, st, frames, nfr, precincts);
        else
            hipLaunchKernelGGL(k_rate_tab<4>, dim3(precincts), dim3(WAVE * 4), 0, st, frames, nfr, precincts);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_rate_plan, dim3(precincts), dim3(WAVE), 0, st, frames, nfr, precincts);
        e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    if (kinds & 2u)
        hipLaunchKernelGGL(k_rate, dim3(slices), dim3(WAVE * RT_NW), 0, st, frames, nfr, slices);
    return hipGetLastError();
}

hipError_t jxs_launch_join(const JxsUnit* units, uint32_t n, hipStream_t st) {
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_join, dim3((n * JOIN_SLOTS + 255) / 256), dim3(256), 0, st, units, n);
    return hipGetLastError();
}

hipError_t jxs_launch_pack(const JxsUnit* units, uint32_t n, uint32_t shadow, hipStream_t st) {
    if (!n)
        return hipSuccess;
    const uint32_t blocks = min((n + PK_WAVES - 1) / PK_WAVES, (uint32_t)PK_BLOCKS);
    hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(64 * PK_WAVES), 0, st, units, n, shadow);
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
