"""CPU tests of the C-ABI boundary: the library loads and exports every entry
point include/*.h declares, the public structs have the reference's layout,
and the host-only parts of the API (defaults, image configuration, parameter
validation, frame sizing) behave as the reference's.  No GPU calls."""
import ctypes
import glob
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from jxs_util import ROOT, load_package

GOLDEN = os.path.join(ROOT, "tests", "golden")
with open(os.path.join(GOLDEN, "golden.json")) as _f:
    G = json.load(_f)


def declared_entry_points():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        for m in re.finditer(r"PREFIX_API\s+[^;{]*?\b(svt_jpeg_xs_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_entry_point():
    P = load_package()
    L = P.lib()
    declared = declared_entry_points()
    assert len(declared) >= 30
    missing = sorted(n for n in declared if not hasattr(L, n))
    assert not missing, missing
    assert declared <= set(P.EXPORTED_SYMBOLS) | declared
    assert not P.missing_symbols()


def test_public_struct_layout_matches_reference():
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "SvtJpegxsEnc.h"
#include "SvtJpegxsDec.h"
int main(void) {
  printf("{\"svt_jpeg_xs_encoder_api_t\": %zu, \"svt_jpeg_xs_decoder_api_t\": %zu, \"svt_jpeg_xs_frame_t\": %zu,"
         " \"svt_jpeg_xs_image_config_t\": %zu, \"svt_jpeg_xs_image_buffer_t\": %zu,"
         " \"svt_jpeg_xs_bitstream_buffer_t\": %zu, \"enc.private_ptr\": %zu, \"enc.padding\": %zu,"
         " \"enc.slice_packetization_mode\": %zu, \"enc.use_cpu_flags\": %zu, \"dec.private_ptr\": %zu,"
         " \"dec.proxy_mode\": %zu, \"frame.bitstream\": %zu, \"frame.user_prv_ctx_ptr\": %zu}\n",
    sizeof(svt_jpeg_xs_encoder_api_t), sizeof(svt_jpeg_xs_decoder_api_t), sizeof(svt_jpeg_xs_frame_t),
    sizeof(svt_jpeg_xs_image_config_t), sizeof(svt_jpeg_xs_image_buffer_t), sizeof(svt_jpeg_xs_bitstream_buffer_t),
    offsetof(svt_jpeg_xs_encoder_api_t, private_ptr), offsetof(svt_jpeg_xs_encoder_api_t, padding),
    offsetof(svt_jpeg_xs_encoder_api_t, slice_packetization_mode), offsetof(svt_jpeg_xs_encoder_api_t, use_cpu_flags),
    offsetof(svt_jpeg_xs_decoder_api_t, private_ptr), offsetof(svt_jpeg_xs_decoder_api_t, proxy_mode),
    offsetof(svt_jpeg_xs_frame_t, bitstream), offsetof(svt_jpeg_xs_frame_t, user_prv_ctx_ptr));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(prog)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe])
        mine = json.loads(subprocess.check_output([exe]).decode())
    assert mine == G["api_struct_sizes"]
    P = load_package()
    sz = G["api_struct_sizes"]
    assert ctypes.sizeof(P.EncoderApi) == sz["svt_jpeg_xs_encoder_api_t"]
    assert ctypes.sizeof(P.DecoderApi) == sz["svt_jpeg_xs_decoder_api_t"]
    assert ctypes.sizeof(P.Frame) == sz["svt_jpeg_xs_frame_t"]
    assert ctypes.sizeof(P.ImageConfig) == sz["svt_jpeg_xs_image_config_t"]
    assert P.EncoderApi.private_ptr.offset == sz["enc.private_ptr"]


def test_load_default_parameters_and_version_check():
    P = load_package()
    L = P.lib()
    a = P.EncoderApi()
    assert L.svt_jpeg_xs_encoder_load_default_parameters(0, 10, ctypes.byref(a)) == 0
    assert (a.ndecomp_v, a.ndecomp_h, a.bpp_denominator, a.slice_height) == (2, 5, 1, 16)
    assert (a.rate_control_mode, a.coding_significance, a.quantization) == (1, 1, 0)
    assert L.svt_jpeg_xs_encoder_load_default_parameters(0, 11, ctypes.byref(a)) == 0x40001001
    assert L.svt_jpeg_xs_encoder_load_default_parameters(1, 0, ctypes.byref(a)) == 0x40001001


@pytest.mark.parametrize("fmt,comps,sub", [(3, 3, (2, 1)), (2, 3, (2, 2)), (4, 3, (1, 1)), (21, 1, None)])
def test_get_image_config(fmt, comps, sub):
    P = load_package()
    L = P.lib()
    a = P.EncoderApi()
    L.svt_jpeg_xs_encoder_load_default_parameters(0, 10, ctypes.byref(a))
    a.source_width, a.source_height, a.input_bit_depth, a.colour_format, a.bpp_numerator = 1920, 1080, 10, fmt, 3
    cfg = P.ImageConfig()
    nb = ctypes.c_uint32()
    assert L.svt_jpeg_xs_encoder_get_image_config(0, 10, ctypes.byref(a), ctypes.byref(cfg), ctypes.byref(nb)) == 0
    assert cfg.components_num == comps
    assert nb.value == (1920 * 1080 * 3 + 7) // 8
    if sub:
        assert (cfg.components[1].width, cfg.components[1].height) == (1920 // sub[0], 1080 // sub[1])
        assert cfg.components[1].byte_size == cfg.components[1].width * cfg.components[1].height * 2
    else:
        assert cfg.components[0].byte_size == 1920 * 1080 * 3 * 2


BAD_CONFIGS = [
    dict(input_bit_depth=7),
    dict(input_bit_depth=15),
    dict(source_width=3),
    dict(colour_format=1),
    dict(colour_format=2, source_height=1081),
    dict(colour_format=3, source_width=1921),
    dict(ndecomp_v=3),
    dict(ndecomp_h=6),
    dict(ndecomp_v=2, ndecomp_h=1),
    dict(slice_height=0),
    dict(slice_height=18),
    dict(quantization=2),
    dict(bpp_numerator=0),
    dict(bpp_denominator=0),
    dict(colour_format=2, ndecomp_v=0),
    dict(source_width=16, source_height=16, ndecomp_h=5, colour_format=3),
    dict(rate_control_mode=4),
    dict(source_width=8200, colour_format=4),  # a component wider than 8,192 samples
    dict(coding_vertical_prediction_mode=3),
    dict(coding_signs_handling=3),
    dict(coding_signs_handling=2, rate_control_mode=2),
    dict(coding_signs_handling=2, rate_control_mode=3, coding_vertical_prediction_mode=1),
]


@pytest.mark.parametrize("bad", BAD_CONFIGS, ids=lambda d: ",".join("%s=%s" % kv for kv in d.items()))
def test_encoder_init_rejects_invalid_parameters(bad):
    """Validation runs on the host before any device allocation
    (EncHandle.c:128-454), so these are CPU tests."""
    P = load_package()
    L = P.lib()
    a = P.EncoderApi()
    L.svt_jpeg_xs_encoder_load_default_parameters(0, 10, ctypes.byref(a))
    a.source_width, a.source_height, a.input_bit_depth, a.colour_format, a.bpp_numerator = 1920, 1080, 10, 3, 3
    a.verbose = 0
    for k, v in bad.items():
        setattr(a, k, v)
    assert L.svt_jpeg_xs_encoder_init(0, 10, ctypes.byref(a)) == P.ERR_BAD_PARAMETER
    assert not a.private_ptr


@pytest.mark.parametrize("name", sorted(n for n in G["small"] if not G["small"][n].get("reference_error")))
def test_get_single_frame_size(name):
    P = load_package()
    L = P.lib()
    cs = open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    buf = np.frombuffer(cs, np.uint8).copy()
    p = G["small"][name]["params"]
    for fast in (0, 1):
        cfg = P.ImageConfig()
        size = ctypes.c_uint32()
        rc = L.svt_jpeg_xs_decoder_get_single_frame_size(buf.ctypes.data, len(buf), ctypes.byref(cfg),
                                                         ctypes.byref(size), fast)
        assert rc == 0
        assert size.value == len(cs)
        assert (cfg.width, cfg.height, cfg.bit_depth) == (p["width"], p["height"], p["depth"])
    size = ctypes.c_uint32()
    rc = L.svt_jpeg_xs_decoder_get_single_frame_size(buf.ctypes.data, len(buf) // 2, None, ctypes.byref(size), 0)
    assert rc == P.ERR_DEC_TOO_SHORT


def test_server_structs_match_header_and_bad_configs_are_refused():
    """The ctypes mirrors of the serving-loop structs have the C layout, and
    invalid configurations or handles are refused before any device call."""
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "svt_jpegxs_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(svt_jpeg_xs_amd_server_config_t), sizeof(svt_jpeg_xs_amd_completion_t),
         sizeof(svt_jpeg_xs_amd_server_stats_t), offsetof(svt_jpeg_xs_amd_completion_t, done_ns),
         offsetof(svt_jpeg_xs_amd_server_stats_t, queued), sizeof(svt_jpeg_xs_amd_frame_spec_t));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(prog)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe])
        sizes = [int(x) for x in subprocess.check_output([exe]).decode().split()]
    P = load_package()
    assert sizes == [ctypes.sizeof(P.ServerConfig), ctypes.sizeof(P.Completion), ctypes.sizeof(P.ServerStats),
                     P.Completion.done_ns.offset, P.ServerStats.queued.offset, ctypes.sizeof(P.FrameSpec)]
    L = P.lib()
    assert not L.svt_jpeg_xs_amd_server_create(None)
    for bad in [(0, 4, 1), (1 << 20, 0, 1), (1 << 20, 4097, 1), (1 << 20, 4, 0), (1 << 20, 4, 5)]:
        assert not L.svt_jpeg_xs_amd_server_create(ctypes.byref(P.ServerConfig(*bad))), bad
    spec = P.FrameSpec()
    assert L.svt_jpeg_xs_amd_server_submit(None, ctypes.byref(spec), 0) == P.ERR_BAD_PARAMETER
    assert L.svt_jpeg_xs_amd_server_poll(None, (P.Completion * 1)(), 1, 1) == 0
    assert L.svt_jpeg_xs_amd_server_stats(None, ctypes.byref(P.ServerStats())) == P.ERR_BAD_PARAMETER
    L.svt_jpeg_xs_amd_server_destroy(None)
    t0 = L.svt_jpeg_xs_amd_server_now_ns()
    t1 = L.svt_jpeg_xs_amd_server_now_ns()
    assert 0 < t0 <= t1
    # an uninitialised encoder makes a footprint of nothing
    assert P.lib().svt_jpeg_xs_amd_plan_footprint(ctypes.byref(spec), 1) == 0
