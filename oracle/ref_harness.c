/*
 * Test-infrastructure driver for the reference SVT-JPEG-XS library built from
 * its own sources into oracle/_ref (see oracle/Makefile.ref).  It only calls
 * the reference's public API (Source/API/SvtJpegxsEnc.h:100-142,
 * Source/API/SvtJpegxsDec.h:59-148), always with use_cpu_flags = CPU_FLAGS_C.
 *
 * Encoder (HARNESS_ENCODER):
 *   ref_enc in.raw out.jxs W H depth format bpp_num bpp_den decomp_v decomp_h frames threads [slice_height [quant [significance [rc_mode]]]]
 *   in.raw holds `frames` frames back to back; each frame is the planes of the
 *   format (planar) or one interleaved plane (packed RGB), u8 for depth 8 and
 *   u16 little-endian above.  Writes the concatenated codestreams and prints one
 *   JSON line with fps and send->get latency percentiles (ms).  With
 *   JXS_REF_REPEAT=R in the environment the frames are sent R times over
 *   (timing of a longer sample; only the first pass's codestreams are written).
 *
 * Decoder (HARNESS_DECODER):
 *   ref_dec in.jxs out.raw frames threads [proxy_mode]
 *   Decodes `frames` concatenated codestreams and writes the decoded planes
 *   (at half or quarter resolution with proxy_mode 1 or 2).
 *
 * Damaged-codestream sweep (HARNESS_SWEEP):
 *   ref_sweep valid.jxs
 *   Runs the damage patterns of the reference's decoder unit tests
 *   (tests/UnitTests/TestDecoder.cc:336-618) on one valid codestream and
 *   prints one line per damaged variant:
 *     kind a b fs r1 r2 h1 h2
 *   kind: T (first a bytes, used size a), Z / O (bytes from a on set to 0 /
 *   1, full size), X (byte a inverted), B (bit 7 - b of byte a flipped),
 *   M (the cumulative walk of TestDecoder.cc:574-604: bit 7 - b of byte a
 *   flipped into a buffer that keeps every flip after which the buffer still
 *   decodes, undone otherwise).  fs: svt_jpeg_xs_decoder_get_single_frame_size
 *   of the variant; r1: a decoder initialised with the variant decodes it; r2:
 *   a decoder initialised with the valid codestream decodes it, then the
 *   variant; h1 / h2: the planes the successful decodes wrote, hashed as
 *   crc32 << 32 | adler32 (zlib) of their bytes in component order (0
 *   otherwise).  The walk keeps a flip when fs and r2 are both 0.  A final
 *   line "F hash r2 h2" gives the walk's buffer (same hash) and its decode.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>
#include "SvtJpegxsEnc.h"
#include "SvtJpegxsDec.h"

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int cmp_d(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

#ifdef HARNESS_ENCODER
typedef struct {
    svt_jpeg_xs_encoder_api_t* enc;
    svt_jpeg_xs_frame_t* frames;
    double* t_send;
    uint32_t n;      /* sends */
    uint32_t nfr;    /* distinct frames, sent in turn */
    uint8_t* bs;     /* nslots codestream buffers */
    uint32_t nslots, bpf;
    int err;
} sender_t;

static void* sender(void* p) {
    sender_t* s = (sender_t*)p;
    for (uint32_t i = 0; i < s->n; i++) {
        svt_jpeg_xs_frame_t fr = s->frames[i % s->nfr];
        fr.bitstream.buffer = s->bs + (size_t)(i % s->nslots) * s->bpf;
        fr.user_prv_ctx_ptr = (void*)(uintptr_t)(i + 1);
        s->t_send[i] = now_ms();
        SvtJxsErrorType_t e = svt_jpeg_xs_encoder_send_picture(s->enc, &fr, 1);
        if (e) {
            s->err = (int)e;
            return NULL;
        }
    }
    return NULL;
}

int main(int argc, char** argv) {
    if (argc < 13) {
        fprintf(stderr, "usage: ref_enc in out W H depth format bpp_num bpp_den V H frames threads [slice_h]\n");
        return 2;
    }
    const char* in_name = argv[1];
    const char* out_name = argv[2];
    svt_jpeg_xs_encoder_api_t enc;
    if (svt_jpeg_xs_encoder_load_default_parameters(SVT_JPEGXS_API_VER_MAJOR, SVT_JPEGXS_API_VER_MINOR, &enc))
        return 3;
    enc.source_width = (uint32_t)atoi(argv[3]);
    enc.source_height = (uint32_t)atoi(argv[4]);
    enc.input_bit_depth = (uint8_t)atoi(argv[5]);
    enc.colour_format = (ColourFormat_t)atoi(argv[6]);
    enc.bpp_numerator = (uint32_t)atoi(argv[7]);
    enc.bpp_denominator = (uint32_t)atoi(argv[8]);
    enc.ndecomp_v = (uint32_t)atoi(argv[9]);
    enc.ndecomp_h = (uint32_t)atoi(argv[10]);
    uint32_t nframes = (uint32_t)atoi(argv[11]);
    enc.threads_num = (uint32_t)atoi(argv[12]);
    if (argc > 13)
        enc.slice_height = (uint32_t)atoi(argv[13]);
    if (argc > 14)
        enc.quantization = (uint32_t)atoi(argv[14]);
    if (argc > 15)
        enc.coding_significance = (uint32_t)atoi(argv[15]);
    if (argc > 16)
        enc.rate_control_mode = (uint32_t)atoi(argv[16]);
    if (argc > 17)
        enc.coding_vertical_prediction_mode = (uint32_t)atoi(argv[17]);
    if (argc > 18)
        enc.coding_signs_handling = (uint32_t)atoi(argv[18]);
    /* JXS_REF_SIMD=1 enables the reference's AVX2 / AVX-512 intrinsic kernels
     * (the SSE2 bit stays clear: the only SSE2 entries are the unassembled NASM
     * routines, whose C twins are then selected).  Output is identical either
     * way (EncoderTest.sh:734-769); the default is the plain C path. */
    const char* simd = getenv("JXS_REF_SIMD");
    enc.use_cpu_flags = (simd && simd[0] == '1') ? (CPU_FLAGS_ALL & ~(CPU_FLAGS)CPU_FLAGS_SSE2) : CPU_FLAGS_C;
    enc.verbose = VERBOSE_ERRORS;

    svt_jpeg_xs_image_config_t cfg;
    uint32_t bytes_per_frame = 0;
    if (svt_jpeg_xs_encoder_get_image_config(SVT_JPEGXS_API_VER_MAJOR, SVT_JPEGXS_API_VER_MINOR, &enc, &cfg, &bytes_per_frame))
        return 4;
    SvtJxsErrorType_t e = svt_jpeg_xs_encoder_init(SVT_JPEGXS_API_VER_MAJOR, SVT_JPEGXS_API_VER_MINOR, &enc);
    if (e) {
        fprintf(stderr, "init error 0x%x\n", (unsigned)e);
        return 5;
    }
    uint32_t pix = enc.input_bit_depth <= 8 ? 1 : 2;
    size_t frame_bytes = 0;
    for (int c = 0; c < cfg.components_num; c++)
        frame_bytes += cfg.components[c].byte_size;
    FILE* fi = fopen(in_name, "rb");
    if (!fi)
        return 6;
    uint8_t* raw = (uint8_t*)malloc(frame_bytes * nframes);
    if (fread(raw, 1, frame_bytes * nframes, fi) != frame_bytes * nframes) {
        fprintf(stderr, "short input\n");
        return 7;
    }
    fclose(fi);
    const char* rep_env = getenv("JXS_REF_REPEAT");
    const uint32_t repeat = rep_env && atoi(rep_env) > 1 ? (uint32_t)atoi(rep_env) : 1u;
    const uint32_t total = nframes * repeat;
    /* one codestream buffer per send of the first pass, reused cyclically after it */
    const uint32_t nslots = nframes < 64 ? (total < 64 ? total : 64) : nframes;
    svt_jpeg_xs_frame_t* frames = (svt_jpeg_xs_frame_t*)calloc(nframes, sizeof(*frames));
    uint8_t* bs = (uint8_t*)malloc((size_t)bytes_per_frame * nslots);
    for (uint32_t f = 0; f < nframes; f++) {
        size_t off = f * frame_bytes;
        for (int c = 0; c < cfg.components_num; c++) {
            frames[f].image.data_yuv[c] = raw + off;
            frames[f].image.stride[c] = cfg.components[c].width * (cfg.format == COLOUR_FORMAT_PACKED_YUV444_OR_RGB ? 3 : 1);
            frames[f].image.alloc_size[c] = cfg.components[c].byte_size;
            off += cfg.components[c].byte_size;
        }
        if (cfg.format == COLOUR_FORMAT_PACKED_YUV444_OR_RGB) {
            /* send_picture validates three planes (EncHandle.c:974-990); the packed
             * frame is one interleaved plane, so all three name it */
            for (int c = 1; c < 3; c++) {
                frames[f].image.data_yuv[c] = frames[f].image.data_yuv[0];
                frames[f].image.stride[c] = frames[f].image.stride[0];
                frames[f].image.alloc_size[c] = frames[f].image.alloc_size[0];
            }
        }
        frames[f].bitstream.buffer = bs + (size_t)f * bytes_per_frame;
        frames[f].bitstream.allocation_size = bytes_per_frame;
        frames[f].bitstream.used_size = 0;
        frames[f].user_prv_ctx_ptr = (void*)(uintptr_t)(f + 1);
    }
    double* t_send = (double*)calloc(total, sizeof(double));
    double* lat = (double*)calloc(total, sizeof(double));
    sender_t s = {&enc, frames, t_send, total, nframes, bs, nslots, bytes_per_frame, 0};
    double t0 = now_ms();
    pthread_t th;
    pthread_create(&th, NULL, sender, &s);
    FILE* fo = fopen(out_name, "wb");
    int rc = 0;
    for (uint32_t f = 0; f < total; f++) {
        svt_jpeg_xs_frame_t out;
        e = svt_jpeg_xs_encoder_get_packet(&enc, &out, 1);
        double t = now_ms();
        if (e) {
            fprintf(stderr, "encode error 0x%x on frame %u\n", (unsigned)e, f);
            rc = 8;
            break;
        }
        uint32_t idx = (uint32_t)(uintptr_t)out.user_prv_ctx_ptr - 1;
        lat[f] = t - t_send[idx];
        if (idx < nframes)
            fwrite(out.bitstream.buffer, 1, out.bitstream.used_size, fo);
    }
    double t1 = now_ms();
    pthread_join(th, NULL);
    fclose(fo);
    svt_jpeg_xs_encoder_close(&enc);
    if (rc || s.err)
        return rc ? rc : 9;
    qsort(lat, total, sizeof(double), cmp_d);
    uint32_t i50 = (total - 1) * 50 / 100, i99 = (total - 1) * 99 / 100;
    printf("{\"frames\": %u, \"bytes_per_frame\": %u, \"fps\": %.3f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, \"threads\": %u}\n",
           total, bytes_per_frame, total * 1e3 / (t1 - t0), lat[i50], lat[i99], enc.threads_num);
    free(raw);
    free(bs);
    free(frames);
    free(t_send);
    free(lat);
    return 0;
}
#endif /*HARNESS_ENCODER*/

#ifdef HARNESS_DECODER
int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: ref_dec in.jxs out.raw frames threads\n");
        return 2;
    }
    FILE* fi = fopen(argv[1], "rb");
    if (!fi)
        return 3;
    fseek(fi, 0, SEEK_END);
    size_t total = (size_t)ftell(fi);
    fseek(fi, 0, SEEK_SET);
    uint8_t* data = (uint8_t*)malloc(total);
    if (fread(data, 1, total, fi) != total)
        return 4;
    fclose(fi);
    uint32_t nframes = (uint32_t)atoi(argv[3]);
    svt_jpeg_xs_decoder_api_t dec;
    memset(&dec, 0, sizeof(dec));
    dec.verbose = VERBOSE_ERRORS;
    dec.threads_num = (uint32_t)atoi(argv[4]);
    dec.use_cpu_flags = CPU_FLAGS_C;
    dec.proxy_mode = argc > 5 ? (proxy_mode_t)atoi(argv[5]) : proxy_mode_full;
    uint32_t frame_size = 0;
    SvtJxsErrorType_t e = svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(data, total, NULL, &frame_size, 1, proxy_mode_full);
    if (e || frame_size == 0)
        return 5;
    svt_jpeg_xs_image_config_t cfg;
    e = svt_jpeg_xs_decoder_init(SVT_JPEGXS_API_VER_MAJOR, SVT_JPEGXS_API_VER_MINOR, &dec, data, frame_size, &cfg);
    if (e) {
        fprintf(stderr, "decoder init error 0x%x\n", (unsigned)e);
        return 6;
    }
    uint32_t pix = cfg.bit_depth <= 8 ? 1 : 2;
    svt_jpeg_xs_image_buffer_t img;
    memset(&img, 0, sizeof(img));
    for (int c = 0; c < cfg.components_num; c++) {
        img.stride[c] = cfg.components[c].width;
        img.alloc_size[c] = cfg.components[c].width * cfg.components[c].height * pix;
        img.data_yuv[c] = malloc(img.alloc_size[c]);
    }
    FILE* fo = fopen(argv[2], "wb");
    size_t off = 0;
    double t0 = now_ms();
    for (uint32_t f = 0; f < nframes; f++) {
        uint32_t fs = 0;
        e = svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(data + off, total - off, NULL, &fs, 1, proxy_mode_full);
        if (e || fs == 0 || off + fs > total) {
            fprintf(stderr, "frame size error on frame %u\n", f);
            return 7;
        }
        svt_jpeg_xs_frame_t in;
        memset(&in, 0, sizeof(in));
        in.bitstream.buffer = data + off;
        in.bitstream.allocation_size = fs;
        in.bitstream.used_size = fs;
        in.image = img;
        e = svt_jpeg_xs_decoder_send_frame(&dec, &in, 1);
        if (e)
            return 8;
        svt_jpeg_xs_frame_t out;
        e = svt_jpeg_xs_decoder_get_frame(&dec, &out, 1);
        if (e) {
            fprintf(stderr, "decode error 0x%x on frame %u\n", (unsigned)e, f);
            return 9;
        }
        for (int c = 0; c < cfg.components_num; c++)
            fwrite(out.image.data_yuv[c], 1, cfg.components[c].byte_size, fo);
        off += fs;
    }
    double t1 = now_ms();
    fclose(fo);
    svt_jpeg_xs_decoder_close(&dec);
    printf("{\"frames\": %u, \"fps\": %.3f, \"components\": %u, \"depth\": %u}\n", nframes, nframes * 1e3 / (t1 - t0),
           cfg.components_num, cfg.bit_depth);
    return 0;
}
#endif /*HARNESS_DECODER*/

#ifdef HARNESS_SWEEP
#include <zlib.h>

/* running crc32 / adler32 of a byte sequence fed in parts */
typedef struct {
    uLong crc, adl;
} hash_t;
static void hash_init(hash_t* h) {
    h->crc = crc32(0L, Z_NULL, 0);
    h->adl = adler32(0L, Z_NULL, 0);
}
static void hash_add(hash_t* h, const uint8_t* p, size_t n) {
    h->crc = crc32(h->crc, p, (uInt)n);
    h->adl = adler32(h->adl, p, (uInt)n);
}
static uint64_t hash_get(const hash_t* h) { return ((uint64_t)(h->crc & 0xFFFFFFFFu) << 32) | (h->adl & 0xFFFFFFFFu); }

/* One decoder over `first` (init + decode), then over `second` when given;
 * returns the last call's status and the hash of what it wrote. */
static int32_t decode_pair(const uint8_t* first, size_t first_size, const uint8_t* second, size_t second_size,
                           uint64_t* hash) {
    svt_jpeg_xs_decoder_api_t dec;
    memset(&dec, 0, sizeof(dec));
    dec.verbose = VERBOSE_NONE;
    dec.threads_num = 1;
    dec.use_cpu_flags = CPU_FLAGS_C;
    svt_jpeg_xs_image_config_t cfg;
    *hash = 0;
    int32_t e = svt_jpeg_xs_decoder_init(SVT_JPEGXS_API_VER_MAJOR, SVT_JPEGXS_API_VER_MINOR, &dec, first, first_size, &cfg);
    if (e)
        return e;
    svt_jpeg_xs_image_buffer_t img;
    memset(&img, 0, sizeof(img));
    uint32_t pix = cfg.bit_depth <= 8 ? 1 : 2;
    for (int c = 0; c < cfg.components_num; c++) {
        img.stride[c] = cfg.components[c].width;
        img.alloc_size[c] = cfg.components[c].width * cfg.components[c].height * pix;
        img.data_yuv[c] = calloc(1, img.alloc_size[c]);
    }
    const uint8_t* bufs[2] = {first, second};
    size_t sizes[2] = {first_size, second_size};
    for (int k = 0; k < (second ? 2 : 1) && !e; k++) {
        for (int c = 0; c < cfg.components_num; c++)
            memset(img.data_yuv[c], 0, img.alloc_size[c]);
        svt_jpeg_xs_frame_t in, out;
        memset(&in, 0, sizeof(in));
        in.bitstream.buffer = (uint8_t*)bufs[k];
        in.bitstream.allocation_size = (uint32_t)sizes[k];
        in.bitstream.used_size = (uint32_t)sizes[k];
        in.image = img;
        e = svt_jpeg_xs_decoder_send_frame(&dec, &in, 1);
        if (!e)
            e = svt_jpeg_xs_decoder_get_frame(&dec, &out, 1);
    }
    if (!e) {
        hash_t h;
        hash_init(&h);
        for (int c = 0; c < cfg.components_num; c++)
            hash_add(&h, (const uint8_t*)img.data_yuv[c], img.alloc_size[c]);
        *hash = hash_get(&h);
    }
    for (int c = 0; c < cfg.components_num; c++)
        free(img.data_yuv[c]);
    svt_jpeg_xs_decoder_close(&dec);
    return e;
}

static int32_t variant(const uint8_t* valid, size_t n, const uint8_t* bad, size_t bad_size, char kind, uint32_t a,
                       uint32_t b) {
    /* the damaged bytes in an allocation of exactly their size */
    uint8_t* v = (uint8_t*)malloc(bad_size ? bad_size : 1);
    memcpy(v, bad, bad_size);
    uint32_t fs_size = 0;
    int32_t fs = svt_jpeg_xs_decoder_get_single_frame_size(v, bad_size, NULL, &fs_size, 0);
    uint64_t h1, h2;
    int32_t r1 = decode_pair(v, bad_size, NULL, 0, &h1);
    int32_t r2 = decode_pair(valid, n, v, bad_size, &h2);
    printf("%c %u %u %d %d %d %016llx %016llx\n", kind, a, b, fs, r1, r2, (unsigned long long)h1,
           (unsigned long long)h2);
    free(v);
    return (fs == 0 && r2 == 0) ? 0 : -1;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: ref_sweep valid.jxs\n");
        return 2;
    }
    FILE* fi = fopen(argv[1], "rb");
    if (!fi)
        return 3;
    fseek(fi, 0, SEEK_END);
    size_t n = (size_t)ftell(fi);
    fseek(fi, 0, SEEK_SET);
    uint8_t* valid = (uint8_t*)malloc(n);
    if (fread(valid, 1, n, fi) != n)
        return 4;
    fclose(fi);
    uint8_t* buf = (uint8_t*)malloc(n);
    for (uint32_t s = 1; s < n; s++)
        variant(valid, n, valid, s, 'T', s, 0);
    for (uint32_t s = 1; s < n; s++) {
        memcpy(buf, valid, n);
        memset(buf + s, 0, n - s);
        variant(valid, n, buf, n, 'Z', s, 0);
    }
    for (uint32_t s = 1; s < n; s++) {
        memcpy(buf, valid, n);
        memset(buf + s, 1, n - s);
        variant(valid, n, buf, n, 'O', s, 0);
    }
    for (uint32_t i = 0; i < n; i++) {
        memcpy(buf, valid, n);
        buf[i] = (uint8_t)~buf[i];
        variant(valid, n, buf, n, 'X', i, 0);
    }
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < 8; j++) {
            memcpy(buf, valid, n);
            buf[i] ^= (uint8_t)(1u << (7 - j));
            variant(valid, n, buf, n, 'B', i, j);
        }
    memcpy(buf, valid, n);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < 8; j++) {
            buf[i] ^= (uint8_t)(1u << (7 - j));
            if (variant(valid, n, buf, n, 'M', i, j) != 0)
                buf[i] ^= (uint8_t)(1u << (7 - j));
        }
    uint64_t h2;
    int32_t r2 = decode_pair(buf, n, NULL, 0, &h2);
    hash_t hb;
    hash_init(&hb);
    hash_add(&hb, buf, n);
    printf("F %016llx %d %016llx\n", (unsigned long long)hash_get(&hb), r2, (unsigned long long)h2);
    free(buf);
    free(valid);
    return 0;
}
#endif /*HARNESS_SWEEP*/
