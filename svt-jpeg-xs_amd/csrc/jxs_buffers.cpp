/*
 * jxs_buffers.cpp -- image / bitstream buffers and frame pools
 * (SvtJpegxsImageBufferTools.h), backed by page-locked host memory.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <vector>

#include "SvtJpegxsImageBufferTools.h"

static void* host_alloc(size_t n) {
    void* p = nullptr;
    if (hipHostMalloc(&p, n ? n : 1, 0) != hipSuccess)
        return nullptr;
    return p;
}

static void host_free(void* p) {
    if (p)
        (void)hipHostFree(p);
}

/* A frame pool is two independent pools, one of image buffers and one of
 * bitstream buffers (ImageBuffer.c:168-246): a frame takes one of each, and
 * each part goes back on its own when the frame says it is releasable. */
struct svt_jpeg_xs_frame_pool {
    bool use_image = false, use_bitstream = false;
    svt_jpeg_xs_image_config_t config;
    uint32_t bitstream_size = 0;
    std::vector<svt_jpeg_xs_image_buffer_t*> images;
    std::vector<svt_jpeg_xs_bitstream_buffer_t*> bitstreams;
    std::deque<uint32_t> free_images, free_bitstreams;
    std::mutex mu;
    std::condition_variable cv;
};

extern "C" {

PREFIX_API svt_jpeg_xs_image_buffer_t* svt_jpeg_xs_image_buffer_alloc(svt_jpeg_xs_image_config_t* cfg) {
    if (!cfg || cfg->components_num > MAX_COMPONENTS_NUM)
        return nullptr;
    auto* b = (svt_jpeg_xs_image_buffer_t*)calloc(1, sizeof(svt_jpeg_xs_image_buffer_t));
    if (!b)
        return nullptr;
    const uint64_t pix = cfg->bit_depth <= 8 ? 1 : 2;
    if (cfg->format == COLOUR_FORMAT_PACKED_YUV444_OR_RGB) {
        const uint64_t stride = (uint64_t)cfg->components[0].width * 3;
        const uint64_t bytes = stride * cfg->components[0].height * pix;
        if (stride > UINT32_MAX || bytes > UINT32_MAX) {
            free(b);
            return nullptr;
        }
        b->stride[0] = (uint32_t)stride;
        b->alloc_size[0] = (uint32_t)bytes;
        b->data_yuv[0] = host_alloc(bytes);
        if (!b->data_yuv[0]) {
            svt_jpeg_xs_image_buffer_free(b);
            return nullptr;
        }
    } else {
        for (uint32_t c = 0; c < cfg->components_num; c++) {
            const uint64_t bytes = (uint64_t)cfg->components[c].width * cfg->components[c].height * pix;
            if (bytes > UINT32_MAX) {
                svt_jpeg_xs_image_buffer_free(b);
                return nullptr;
            }
            b->stride[c] = cfg->components[c].width;
            b->alloc_size[c] = (uint32_t)bytes;
            b->data_yuv[c] = host_alloc(bytes);
            if (!b->data_yuv[c]) {
                svt_jpeg_xs_image_buffer_free(b);
                return nullptr;
            }
        }
    }
    return b;
}

PREFIX_API void svt_jpeg_xs_image_buffer_free(svt_jpeg_xs_image_buffer_t* b) {
    if (!b)
        return;
    for (int c = 0; c < MAX_COMPONENTS_NUM; c++)
        host_free(b->data_yuv[c]);
    free(b);
}

PREFIX_API svt_jpeg_xs_bitstream_buffer_t* svt_jpeg_xs_bitstream_alloc(uint32_t size) {
    auto* b = (svt_jpeg_xs_bitstream_buffer_t*)calloc(1, sizeof(svt_jpeg_xs_bitstream_buffer_t));
    if (!b)
        return nullptr;
    b->allocation_size = size;
    b->buffer = (uint8_t*)host_alloc(size);
    if (!b->buffer) {
        free(b);
        return nullptr;
    }
    return b;
}

PREFIX_API void svt_jpeg_xs_bitstream_free(svt_jpeg_xs_bitstream_buffer_t* b) {
    if (!b)
        return;
    host_free(b->buffer);
    free(b);
}

PREFIX_API svt_jpeg_xs_frame_pool_t* svt_jpeg_xs_frame_pool_alloc(const svt_jpeg_xs_image_config_t* cfg,
                                                                  uint32_t bitstream_size, uint32_t count) {
    /* ImageBuffer.c:172-174 */
    if (count == 0 || (!cfg && bitstream_size == 0) || (cfg && cfg->components_num > MAX_COMPONENTS_NUM))
        return nullptr;
    auto* p = new (std::nothrow) svt_jpeg_xs_frame_pool;
    if (!p)
        return nullptr;
    p->use_image = cfg != nullptr;
    p->use_bitstream = bitstream_size != 0;
    if (cfg)
        p->config = *cfg;
    p->bitstream_size = bitstream_size;
    for (uint32_t i = 0; i < count; i++) {
        svt_jpeg_xs_image_buffer_t* img = nullptr;
        svt_jpeg_xs_bitstream_buffer_t* bs = nullptr;
        if (p->use_image && !(img = svt_jpeg_xs_image_buffer_alloc(&p->config))) {
            svt_jpeg_xs_frame_pool_free(p);
            return nullptr;
        }
        if (p->use_bitstream && !(bs = svt_jpeg_xs_bitstream_alloc(bitstream_size))) {
            svt_jpeg_xs_image_buffer_free(img);
            svt_jpeg_xs_frame_pool_free(p);
            return nullptr;
        }
        /* release_ctx_ptr carries the buffer's slot (index + 1) in its pool */
        if (img) {
            img->release_ctx_ptr = (void*)(uintptr_t)(i + 1);
            p->free_images.push_back(i);
        }
        if (bs) {
            bs->release_ctx_ptr = (void*)(uintptr_t)(i + 1);
            p->free_bitstreams.push_back(i);
        }
        p->images.push_back(img);
        p->bitstreams.push_back(bs);
    }
    return p;
}

PREFIX_API void svt_jpeg_xs_frame_pool_free(svt_jpeg_xs_frame_pool_t* p) {
    if (!p)
        return;
    for (auto* i : p->images)
        svt_jpeg_xs_image_buffer_free(i);
    for (auto* b : p->bitstreams)
        svt_jpeg_xs_bitstream_free(b);
    delete p;
}

/* One image and one bitstream buffer (ImageBuffer.c:248-296); blocking
 * waits for both, non-blocking takes neither unless both are free. */
PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_frame_pool_get(svt_jpeg_xs_frame_pool_t* p, svt_jpeg_xs_frame_t* frame,
                                                        uint8_t blocking_flag) {
    if (!p || !frame)
        return SvtJxsErrorBadParameter;
    std::unique_lock<std::mutex> lk(p->mu);
    auto ready = [&] {
        return (!p->use_image || !p->free_images.empty()) && (!p->use_bitstream || !p->free_bitstreams.empty());
    };
    if (!ready()) {
        if (!blocking_flag)
            return SvtJxsErrorNoErrorEmptyQueue;
        p->cv.wait(lk, ready);
    }
    if (p->use_image) {
        frame->image = *p->images[p->free_images.front()];
        p->free_images.pop_front();
    }
    if (p->use_bitstream) {
        frame->bitstream = *p->bitstreams[p->free_bitstreams.front()];
        frame->bitstream.used_size = 0;
        p->free_bitstreams.pop_front();
    }
    return SvtJxsErrorNone;
}

/* Each part goes back only when it carries its pool slot and is marked
 * releasable; the slot is then cleared from the frame, so a second release
 * of the same frame returns nothing (ImageBuffer.c:306-318). */
PREFIX_API void svt_jpeg_xs_frame_pool_release(svt_jpeg_xs_frame_pool_t* p, svt_jpeg_xs_frame_t* frame) {
    if (!p || !frame)
        return;
    bool freed = false;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        const uintptr_t it = (uintptr_t)frame->image.release_ctx_ptr;
        if (p->use_image && it && frame->image.ready_to_release) {
            if (it <= p->images.size()) {
                p->free_images.push_back((uint32_t)(it - 1));
                freed = true;
            }
            frame->image.release_ctx_ptr = nullptr;
        }
        const uintptr_t bt = (uintptr_t)frame->bitstream.release_ctx_ptr;
        if (p->use_bitstream && bt && frame->bitstream.ready_to_release) {
            if (bt <= p->bitstreams.size()) {
                p->free_bitstreams.push_back((uint32_t)(bt - 1));
                freed = true;
            }
            frame->bitstream.release_ctx_ptr = nullptr;
        }
    }
    if (freed)
        p->cv.notify_all();
}

}  // extern "C"
