/*
 * jxs_enc.h -- encoder internals shared by the launch plans and the serving
 * loop: the per-encoder geometry resident in HBM, the layout of one encode
 * batch inside a single device block (frame descriptors, transform job
 * lists, error words, packing units, per-frame work buffers), and the launch
 * sequence of a batch.
 */
#ifndef JXS_ENC_H
#define JXS_ENC_H

#include <hip/hip_runtime.h>

#include <vector>

#include "SvtJpegxsEnc.h"
#include "jxs_host.h"
#include "svt_jpegxs_amd.h"

namespace jxs {

static inline uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

/* Geometry of one encoder: the host copy, and the device copy every batch
 * that carries a frame of this encoder reads (band layout, weights, slice
 * budgets stay resident for the encoder's lifetime). */
struct DevGeom {
    HostGeom host;
    JxsGeom* d_geom = nullptr;
    uint32_t* d_slice_off = nullptr;
    uint32_t* d_slice_bud = nullptr;
    uint32_t njobs[2] = {}; /* transform jobs of one frame per stage (their tile widths may change: see autotune) */
    uint64_t work = 0;         /* work bytes of one frame (inputs and codestream excluded) */
    void release();
    SvtJxsErrorType_t upload();
};

/* The resident geometry of an initialised encoder, or null. */
const DevGeom* encoder_geom(const svt_jpeg_xs_encoder_api_t* api);

/* Device bytes of a batch, accumulated frame by frame (exact: the block
 * batch_place lays out has this size). */
struct BatchSizer {
    uint32_t n = 0, units = 0;
    uint32_t njobs[2] = {}; /* every chunk class's job list has room for all of a stage's jobs */
    uint64_t work = 0;
    void add(const DevGeom& g) {
        n++;
        /* frames with vertical prediction or sign coding are packed without units */
        units += (g.host.g.vpred || g.host.g.signs) ? 0u : g.host.g.rows * g.host.g.n_units;
        for (int s = 0; s < 2; s++)
            njobs[s] += g.njobs[s];
        work += g.work;
    }
    /* bytes of the descriptor part that is uploaded from the host */
    uint64_t upload_bytes() const;
    /* the whole block: descriptors, error words, packing units, work */
    uint64_t bytes() const;
};

/* Device addresses and launch sizes of one encode batch. */
struct BatchLaunch {
    uint32_t nfr = 0;
    uint32_t* status = nullptr;
    JxsFrameDesc* frames = nullptr;
    JxsUnit* units = nullptr;
    JxsDwtJob* dwt[2][3] = {};
    uint32_t* dbase[2][3] = {};
    int ndwt[2][3] = {};
    uint32_t dblocks[2][3] = {};
    uint32_t total_slices = 0, total_prec = 0, total_units = 0;
    bool ordered = false; /* a frame is coded by k_enc_slice */
    uint32_t rate_kinds = 0; /* bit 0: a frame under rate control mode 0 / 1, bit 1: mode 2 / 3 (k_rate) */
    uint32_t rate_waves = 4; /* waves per precinct of the rate table kernel: the choice for the batch's largest frame */
};

/* Checks one frame spec (initialised encoder, codestream buffer 4-byte
 * aligned, planes present); returns its geometry or null. */
const DevGeom* spec_geom(const svt_jpeg_xs_amd_frame_spec_t& s);

/* Lays out count frames in the device block at base (256-byte aligned,
 * BatchSizer::bytes() long).  Fills L and writes the descriptor image (the
 * block's first upload_bytes(), to be copied there before the launch) into
 * img.  Returns false when a spec is invalid or a transform job falls outside
 * what the transform kernel's layout covers. */
bool batch_place(const svt_jpeg_xs_amd_frame_spec_t* specs, uint32_t count, uint8_t* base, BatchLaunch& L,
                 std::vector<uint8_t>& img);
/* The same into a caller's buffer of exactly BatchSizer::upload_bytes()
 * bytes, with no allocation. */
bool batch_place_into(const svt_jpeg_xs_amd_frame_spec_t* specs, uint32_t count, uint8_t* base, BatchLaunch& L,
                      uint8_t* img, uint64_t img_bytes);

/* The launch sequence of a batch on one stream.  ev (optional) gets an event
 * recorded before every stage and after the last. */
hipError_t batch_launch(const BatchLaunch& L, hipStream_t st, hipEvent_t* ev);

}  // namespace jxs

#endif
