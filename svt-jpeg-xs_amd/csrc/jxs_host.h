/*
 * jxs_host.h -- host-side coding configuration: parameter validation, band
 * layout, weight tables, frame header, slice budgets, and the job lists the
 * kernels run.
 */
#ifndef JXS_HOST_H
#define JXS_HOST_H

#include <stdint.h>
#include <vector>

#include "jxs_codestream.h"
#include "jxs_geom.h"

namespace jxs {

struct EncParams {
    uint32_t width, height, depth, format;
    uint32_t bpp_num, bpp_den, decomp_v, decomp_h;
    uint32_t quantization, slice_height, significance;
    uint32_t rc_mode; /* rate_control_mode 0..3 */
    uint32_t vpred;   /* coding_vertical_prediction_mode 0..2 */
    uint32_t signs;   /* coding_signs_handling 0..2 */
};

/* Host geometry = device geometry + slice tables. */
struct HostGeom {
    JxsGeom g;
    std::vector<uint32_t> slice_size; /* bytes of each slice incl. its header (EOC in the last) */
    std::vector<uint32_t> slice_off;  /* codestream offset of each slice header */
    std::vector<uint32_t> slice_bud;  /* precinct bytes of each slice */
    uint32_t bw = 20, fq = 8;         /* decoder: coefficient precision Bw and fractional bits Fq (PIH) */
};

/* Encoder configuration: EncHandle.c:128-454 (validation), Pi.c:189-406
 * (layout), WeightTable.c:336-413, PiEnc.c:11-53, PackHeaders.c:11-140,
 * EncHandle.c:760-809.  Returns 0 or an SvtJxsErrorType_t value. */
int32_t build_encoder_geometry(const EncParams& p, HostGeom& out, int verbose);

/* Decoder configuration from a parsed codestream header: the band layout,
 * weights and per-frame coding flags; a header this decoder does not decode
 * (more than one precinct column, CWD, colour or non-linear transforms, a
 * coefficient precision other than Bw 18 or 20) is
 * SvtJxsErrorDecoderInvalidBitstream. */
int32_t geom_from_header(const CsHeader& h, uint32_t header_end, HostGeom& out, uint32_t* header_end_out);

/* cs_get_header + cs_pi_check + geom_from_header for a full-resolution
 * decode. */
int32_t parse_codestream_header(const uint8_t* cs, size_t len, HostGeom& out, uint32_t* header_end);

/* Band layout only (shared by the encoder and decoder builders). */
int build_layout(JxsGeom& g, uint32_t nc, const uint32_t sx[4], const uint32_t sy[4], uint32_t W, uint32_t H, uint32_t dh,
                 uint32_t dv, uint32_t slice_h);

/* Assigns buffer offsets and sizes (coefficients, GCLI, significance max,
 * transform scratch) for one frame of this geometry. */
void assign_buffers(JxsGeom& g);

/* Forward-transform job lists for one frame.  The base pointers are device
 * addresses of this frame's buffers (hist: band-line histogram records, may
 * be null; scratch: the int32 low-pass planes between the two vertical
 * levels); the input planes are device addresses.  dwt[0]: first vertical
 * level (or horizontal-only components), dwt[1]: second vertical level. */
struct FwdJobs {
    JxsDwtJob dwt[2][JXS_MAXC]; /* at most one job per component and stage */
    uint32_t n[2] = {0, 0};
    void push(int s, const JxsDwtJob& j) { dwt[s][n[s]++] = j; }
};
void fwd_jobs(const JxsGeom& g, const void* const planes[4], const uint32_t strides[4], uint16_t* coeff, uint8_t* gcli,
              uint8_t* smax, uint16_t* hist, int32_t* scratch, FwdJobs& jobs);

/* Transform tile widths (columns per thread: 8, 16 or 32) chosen by the
 * autotuner, per job shape; a job with no choice recorded keeps the
 * built-in rule.  The table is per process and thread-safe. */
struct TileKey {
    uint32_t kind, step, vert, lev_l, w;
};
TileKey tile_key(const JxsDwtJob& j);
/* whether the transform kernel's layout covers this width for the job */
bool tile_legal(uint32_t lev_l, uint32_t w, uint32_t chunk);
void tile_set(const TileKey& k, uint32_t chunk); /* chunk 0: back to the rule */
uint32_t tile_get(const TileKey& k);              /* 0: none recorded */
/* Waves per precinct (1, 2 or 4) of the rate table kernel for frames of this
 * geometry: the autotuner's choice, 4 when none is recorded (0 resets). */
uint32_t rate_waves_get(const JxsGeom& g);
void rate_waves_set(const JxsGeom& g, uint32_t waves);

/* Inverse-transform job lists (decoder). */
struct InvVertJob;
struct InvHorzJob;

}  // namespace jxs

#endif
