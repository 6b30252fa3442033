/*
 * SvtJpegxsEnc.h -- JPEG XS encoder C API of the MI355X-native library.
 *
 * Same layout and entry points as the reference encoder API v0.10
 * (Source/API/SvtJpegxsEnc.h:18-142).  Each call is served by the GPU path:
 * send_picture copies the planes to HBM and enqueues the fused
 * transform -> GCLI -> rate control -> quantise/pack kernels on the encoder's
 * HIP stream; get_packet waits for that frame and returns the codestream in
 * the caller's bitstream buffer.
 */
#ifndef SVT_JPEGXS_AMD_ENC_H
#define SVT_JPEGXS_AMD_ENC_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "SvtJpegxs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Source/API/SvtJpegxsEnc.h:18-93 (member order and padding kept) */
typedef struct svt_jpeg_xs_encoder_api {
    uint32_t source_width;
    uint32_t source_height;
    uint8_t input_bit_depth;
    ColourFormat_t colour_format;
    uint32_t bpp_numerator;
    uint32_t bpp_denominator;
    uint32_t ndecomp_v;
    uint32_t ndecomp_h;
    uint32_t quantization;
    uint32_t slice_height;
    CPU_FLAGS use_cpu_flags;
    uint32_t threads_num;
    uint8_t cpu_profile;
    uint32_t print_bands_info;
    uint32_t coding_signs_handling;
    uint32_t coding_significance;
    uint32_t rate_control_mode;
    uint32_t coding_vertical_prediction_mode;
    uint32_t verbose;
    void (*callback_send_data_available)(struct svt_jpeg_xs_encoder_api* encoder, void* context);
    void* callback_send_data_available_context;
    void (*callback_get_data_available)(struct svt_jpeg_xs_encoder_api* encoder, void* context);
    void* callback_get_data_available_context;
    uint8_t slice_packetization_mode;
    void* private_ptr;
    uint8_t padding[64];
} svt_jpeg_xs_encoder_api_t;

/* Replaces svt_jpeg_xs_encoder_load_default_parameters
 * (Source/API/SvtJpegxsEnc.h:100-101; EncHandle.c:490-532). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_load_default_parameters(uint64_t version_api_major,
                                                                         uint64_t version_api_minor,
                                                                         svt_jpeg_xs_encoder_api_t* enc_api);

/* Replaces svt_jpeg_xs_encoder_get_image_config
 * (Source/API/SvtJpegxsEnc.h:108-111; EncHandle.c:537-612). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_get_image_config(uint64_t version_api_major, uint64_t version_api_minor,
                                                                  svt_jpeg_xs_encoder_api_t* enc_api,
                                                                  svt_jpeg_xs_image_config_t* out_image_config,
                                                                  uint32_t* out_bytes_per_frame);

/* Replaces svt_jpeg_xs_encoder_init (Source/API/SvtJpegxsEnc.h:118-119;
 * EncHandle.c:623-936): validates the parameters, builds the band layout,
 * weight tables, frame header and slice budgets, and allocates the device
 * buffers and HIP stream of this encoder. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_init(uint64_t version_api_major, uint64_t version_api_minor,
                                                      svt_jpeg_xs_encoder_api_t* enc_api);

/* Replaces svt_jpeg_xs_encoder_close (Source/API/SvtJpegxsEnc.h:122;
 * EncHandle.c:938-950). */
PREFIX_API void svt_jpeg_xs_encoder_close(svt_jpeg_xs_encoder_api_t* enc_api);

/* Replaces svt_jpeg_xs_encoder_send_picture (Source/API/SvtJpegxsEnc.h:130-131;
 * EncHandle.c:955-1018). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_send_picture(svt_jpeg_xs_encoder_api_t* enc_api,
                                                              svt_jpeg_xs_frame_t* enc_input, uint8_t blocking_flag);

/* Replaces svt_jpeg_xs_encoder_get_packet (Source/API/SvtJpegxsEnc.h:141-142;
 * EncHandle.c:1023-1059). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_encoder_get_packet(svt_jpeg_xs_encoder_api_t* enc_api,
                                                            svt_jpeg_xs_frame_t* enc_output, uint8_t blocking_flag);

#ifdef __cplusplus
}
#endif
#endif
