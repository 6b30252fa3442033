/*
 * SvtJpegxsDec.h -- JPEG XS decoder C API of the MI355X-native library.
 *
 * Same layout and entry points as the reference decoder API v0.10
 * (Source/API/SvtJpegxsDec.h:16-148).  The decode path runs on the GPU:
 * precinct unpacking, dequantisation, inverse 5/3 transform and output
 * scaling are HIP kernels; the header and precinct-length scan is host code.
 */
#ifndef SVT_JPEGXS_AMD_DEC_H
#define SVT_JPEGXS_AMD_DEC_H

#include <stddef.h>
#include <stdint.h>
#include "SvtJpegxs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Source/API/SvtJpegxsDec.h:16-57 (member order and padding kept) */
typedef struct svt_jpeg_xs_decoder_api {
    uint64_t use_cpu_flags;
    uint32_t verbose;
    uint32_t threads_num;
    uint8_t packetization_mode;
    proxy_mode_t proxy_mode;
    void (*callback_send_data_available)(struct svt_jpeg_xs_decoder_api* decoder, void* context);
    void* callback_send_data_available_context;
    void (*callback_get_data_available)(struct svt_jpeg_xs_decoder_api* decoder, void* context);
    void* callback_get_data_available_context;
    void* private_ptr;
    uint8_t padding[64];
} svt_jpeg_xs_decoder_api_t;

/* Replaces svt_jpeg_xs_decoder_init (Source/API/SvtJpegxsDec.h:59-61;
 * DecHandle.c:106-362). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_init(uint64_t version_api_major, uint64_t version_api_minor,
                                                      svt_jpeg_xs_decoder_api_t* dec_api, const uint8_t* bitstream_buf,
                                                      size_t codestream_size, svt_jpeg_xs_image_config_t* out_image_config);

/* Replaces svt_jpeg_xs_decoder_close (Source/API/SvtJpegxsDec.h:63). */
PREFIX_API void svt_jpeg_xs_decoder_close(svt_jpeg_xs_decoder_api_t* dec_api);

/* Replace svt_jpeg_xs_decoder_get_single_frame_size[_with_proxy]
 * (Source/API/SvtJpegxsDec.h:83-93). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_single_frame_size(const uint8_t* bitstream_buf,
                                                                       size_t bitstream_buf_size,
                                                                       svt_jpeg_xs_image_config_t* out_image_config,
                                                                       uint32_t* frame_size, uint32_t fast_search);
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_single_frame_size_with_proxy(
    const uint8_t* bitstream_buf, size_t bitstream_buf_size, svt_jpeg_xs_image_config_t* out_image_config,
    uint32_t* frame_size, uint32_t fast_search, proxy_mode_t proxy_mode);

/* Replaces svt_jpeg_xs_decoder_send_frame (Source/API/SvtJpegxsDec.h:106-107). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                            svt_jpeg_xs_frame_t* dec_input, uint8_t blocking_flag);

/* Replaces svt_jpeg_xs_decoder_send_packet (Source/API/SvtJpegxsDec.h:120-121). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_packet(svt_jpeg_xs_decoder_api_t* dec_api,
                                                             svt_jpeg_xs_frame_t* dec_input, uint32_t* bytes_used);

/* Replaces svt_jpeg_xs_decoder_get_frame (Source/API/SvtJpegxsDec.h:136-137). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_get_frame(svt_jpeg_xs_decoder_api_t* dec_api,
                                                           svt_jpeg_xs_frame_t* dec_output, uint8_t blocking_flag);

/* Replaces svt_jpeg_xs_decoder_send_eoc (Source/API/SvtJpegxsDec.h:148). */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_decoder_send_eoc(svt_jpeg_xs_decoder_api_t* dec_api);

#ifdef __cplusplus
}
#endif
#endif
