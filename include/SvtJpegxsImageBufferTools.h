/*
 * SvtJpegxsImageBufferTools.h -- image / bitstream buffers and frame pools.
 *
 * Same entry points and semantics as the reference
 * (Source/API/SvtJpegxsImageBufferTools.h:20-88; ImageBuffer.c:22-318).  The
 * buffers are page-locked host memory, so the library's host<->HBM copies
 * run at full PCIe rate without an extra staging copy.
 */
#ifndef SVT_JPEGXS_AMD_BUFFER_TOOLS_H
#define SVT_JPEGXS_AMD_BUFFER_TOOLS_H

#include "SvtJpegxs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces svt_jpeg_xs_image_buffer_alloc / _free (SvtJpegxsImageBufferTools.h:20-32). */
PREFIX_API svt_jpeg_xs_image_buffer_t* svt_jpeg_xs_image_buffer_alloc(svt_jpeg_xs_image_config_t* image_config);
PREFIX_API void svt_jpeg_xs_image_buffer_free(svt_jpeg_xs_image_buffer_t* image_buffer);

/* Replaces svt_jpeg_xs_bitstream_alloc / _free (SvtJpegxsImageBufferTools.h:34-46). */
PREFIX_API svt_jpeg_xs_bitstream_buffer_t* svt_jpeg_xs_bitstream_alloc(uint32_t bitstream_size);
PREFIX_API void svt_jpeg_xs_bitstream_free(svt_jpeg_xs_bitstream_buffer_t* bitstream);

typedef struct svt_jpeg_xs_frame_pool svt_jpeg_xs_frame_pool_t;

/* Replace svt_jpeg_xs_frame_pool_{alloc,free,get,release}
 * (SvtJpegxsImageBufferTools.h:51-88). */
PREFIX_API svt_jpeg_xs_frame_pool_t* svt_jpeg_xs_frame_pool_alloc(const svt_jpeg_xs_image_config_t* image_config,
                                                                  uint32_t bitstream_size, uint32_t count);
PREFIX_API void svt_jpeg_xs_frame_pool_free(svt_jpeg_xs_frame_pool_t* frame_pool);
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_frame_pool_get(svt_jpeg_xs_frame_pool_t* frame_pool, svt_jpeg_xs_frame_t* frame,
                                                        uint8_t blocking_flag);
PREFIX_API void svt_jpeg_xs_frame_pool_release(svt_jpeg_xs_frame_pool_t* frame_pool, svt_jpeg_xs_frame_t* frame);

#ifdef __cplusplus
}
#endif
#endif
