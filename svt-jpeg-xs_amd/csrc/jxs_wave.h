/*
 * jxs_wave.h -- wave-level helpers shared by the encoder and decoder
 * kernels (64-lane waves, gfx950).
 */
#ifndef JXS_WAVE_H
#define JXS_WAVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Inclusive sum over the 64 lanes of a wave with DPP moves: shifts by 1, 2,
 * 4, 8 inside each row of 16 lanes, then the row totals broadcast from lane
 * 15 into rows 1 and 3 and from lane 31 into rows 2 and 3.  Every lane of
 * the wave must be active. */
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(1), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(2), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(4), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_SHR(8), 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_BCAST15, 0xA, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, DPP_ROW_BCAST31, 0xC, 0xF, false);
    return (uint32_t)x;
}

#endif
