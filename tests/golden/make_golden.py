"""Writes the golden fixtures of tests/golden from the reference SVT-JPEG-XS
library built from its own sources (oracle/Makefile.ref -> oracle/_ref).

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

Fixtures (data only):
  golden.json         per case: generator parameters, SHA-256 of the input
                      planes, of the reference codestream and of the reference
                      decoder's planes; struct sizes of the reference API;
                      the quantisation known-answer vectors.
  <case>.jxs          reference codestreams of the small cases.
  sample_16x16_422.jxs  the 16x16 4:2:2 codestream the reference's unit tests
                      decode (tests/UnitTests/SampleFramesData.cc:8-33).
  sweep_<case>.txt.gz the reference decoder's verdict on every damaged variant
                      of a codestream that its decoder unit tests build
                      (TestDecoder.cc:336-618: every truncation, zero and one
                      fill, every byte inverted, every bit flipped, and the
                      cumulative bit walk), one line per variant as
                      oracle/ref_harness.c (HARNESS_SWEEP) prints it.

    python tests/golden/make_golden.py --sweeps   (only the sweep fixtures)
"""
import gzip
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
REF = "/root/reference"

from jxs_util import load_oracle, load_synth  # noqa: E402

# name: (W, H, depth, format, bpp_num, bpp_den, decomp_v, decomp_h, quant, slice_h, significance, sigma, frame
#        [, rate_control_mode (default 1) [, coding_vertical_prediction_mode, coding_signs_handling (default 0)]])
SMALL = {
    "s_64x32_422_8b_bpp5_v2": (64, 32, 8, 3, 5, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "s_256x64_422_10b_bpp3_v2": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1),
    "s_200x200_444_8b_bpp3_v1": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2),
    "s_201x203_444_10b_bpp2_v2": (201, 203, 10, 4, 2, 1, 2, 5, 0, 16, 1, 4.0, 3),
    "s_256x128_420_8b_bpp2_v2": (256, 128, 8, 2, 2, 1, 2, 5, 0, 16, 1, 4.0, 4),
    "s_128x96_420_8b_bpp2_v1h3": (128, 96, 8, 2, 2, 1, 1, 3, 0, 16, 1, 4.0, 5),
    "s_160x90_444_12b_bpp4_uniform": (160, 90, 12, 4, 4, 1, 2, 5, 1, 16, 1, 4.0, 6),
    "s_160x90_422_10b_bpp5_nosig": (160, 90, 10, 3, 5, 1, 2, 5, 0, 16, 0, 4.0, 7),
    "s_96x64_rgbpacked_14b_bpp2_v1": (96, 64, 14, 21, 2, 1, 1, 5, 0, 16, 1, 16.0, 8),
    "s_128x64_444_11b_bpp3_v0h3": (128, 64, 11, 4, 3, 1, 0, 3, 0, 16, 1, 4.0, 9),
    "s_130x66_422_9b_bpp4_v1h2": (130, 66, 9, 3, 4, 1, 1, 2, 0, 16, 1, 4.0, 10),
    "s_320x180_422_10b_bpp7_2_slice32": (320, 180, 10, 3, 7, 2, 2, 5, 0, 32, 1, 4.0, 11),
    "s_256x192_422_10b_bpp20": (256, 192, 10, 3, 20, 1, 2, 5, 0, 16, 1, 4.0, 12),
    "s_30x60_444_8b_bpp4_error": (30, 60, 8, 4, 4, 1, 2, 5, 0, 16, 1, 4.0, 13),
    # rate control mode 0: budget per precinct, padding kept in each precinct
    "s_256x64_422_10b_bpp3_v2_rc0": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 0),
    "s_200x200_444_8b_bpp3_v1_rc0": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2, 0),
    "s_320x180_422_10b_bpp7_2_slice32_rc0": (320, 180, 10, 3, 7, 2, 2, 5, 0, 32, 1, 4.0, 11, 0),
    # rate control modes 2 and 3: one Q and R per slice sizes the precinct budgets
    "s_256x64_422_10b_bpp3_v2_rc2": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 2),
    "s_200x200_444_8b_bpp3_v1_rc2": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2, 2),
    "s_320x180_422_10b_bpp7_2_slice32_rc2": (320, 180, 10, 3, 7, 2, 2, 5, 0, 32, 1, 4.0, 11, 2),
    "s_256x64_422_10b_bpp3_v2_rc3": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 3),
    "s_200x200_444_8b_bpp3_v1_rc3": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2, 3),
    "s_320x180_422_10b_bpp7_2_slice32_rc3": (320, 180, 10, 3, 7, 2, 2, 5, 0, 32, 1, 4.0, 11, 3),
    "s_96x64_rgbpacked_14b_bpp2_v1_rc3": (96, 64, 14, 21, 2, 1, 1, 5, 0, 16, 1, 16.0, 8, 3),
    # uneven content (the first 8 rows of every 16 flat, synth.flatten_rows), so
    # that mode 3 raises the small precincts to a quarter of the largest
    "u_256x64_422_10b_bpp3_v2_rc2": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 8.0, 1, 2),
    "u_256x64_422_10b_bpp3_v2_rc3": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 8.0, 1, 3),
    "u_200x96_444_8b_bpp2_v1_rc3": (200, 96, 8, 4, 2, 1, 1, 5, 0, 16, 1, 8.0, 2, 3),
    "u_320x96_420_8b_bpp2_v2_slice32_rc3": (320, 96, 8, 2, 2, 1, 2, 5, 0, 32, 1, 8.0, 3, 3),
    # vertical prediction (1: zero residuals, 2: zero coefficients / Rm = 1)
    # and sign coding (1: fast, 2: full)
    "t_256x64_422_10b_bpp3_v2_vp1": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 1, 1, 0),
    "t_256x64_422_10b_bpp3_v2_vp2": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 1, 2, 0),
    "t_200x200_444_8b_bpp3_v1_vp1_rc0": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2, 0, 1, 0),
    "t_160x90_422_10b_bpp5_nosig_vp1": (160, 90, 10, 3, 5, 1, 2, 5, 0, 16, 0, 4.0, 7, 1, 1, 0),
    "t_256x128_420_8b_bpp2_v2_vp2_rc3": (256, 128, 8, 2, 2, 1, 2, 5, 0, 16, 1, 4.0, 4, 3, 2, 0),
    "t_256x64_422_10b_bpp3_v2_sg1": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 1, 0, 1),
    "t_256x64_422_10b_bpp3_v2_sg2": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 1, 0, 2),
    "t_160x90_444_12b_bpp4_uniform_sg2": (160, 90, 12, 4, 4, 1, 2, 5, 1, 16, 1, 4.0, 6, 1, 0, 2),
    "t_200x200_444_8b_bpp3_v1_sg1_rc0": (200, 200, 8, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 2, 0, 0, 1),
    "t_320x180_422_10b_bpp7_2_slice32_sg1_rc2": (320, 180, 10, 3, 7, 2, 2, 5, 0, 32, 1, 4.0, 11, 2, 0, 1),
    "t_96x64_rgbpacked_14b_bpp2_v1_vp1_sg2": (96, 64, 14, 21, 2, 1, 1, 5, 0, 16, 1, 16.0, 8, 1, 1, 2),
    "t_256x64_422_10b_bpp3_v2_vp2_sg1_rc3": (256, 64, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 1, 3, 2, 1),
    "t_130x66_422_9b_bpp4_v1h2_vp1_sg2_rc0": (130, 66, 9, 3, 4, 1, 1, 2, 0, 16, 1, 4.0, 10, 0, 1, 2),
    "t_160x90_444_12b_bpp4_uniform_vp1_sg1": (160, 90, 12, 4, 4, 1, 2, 5, 1, 16, 1, 4.0, 6, 1, 1, 1),
}

# The configurations of the benchmark workloads, first frame of each.
FULL = {
    "c1_1920x1080_422_8b_bpp5_v2": (1920, 1080, 8, 3, 5, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "c2_3840x2160_422_10b_bpp3_v2": (3840, 2160, 10, 3, 3, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "c3a_1280x720_420_8b_bpp1_v2": (1280, 720, 8, 2, 1, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "c3b_1920x1080_444_12b_bpp4_v2": (1920, 1080, 12, 4, 4, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "c3c_3840x2160_rgb_10b_bpp3_v2": (3840, 2160, 10, 4, 3, 1, 2, 5, 0, 16, 1, 4.0, 0),
    "c3d_1920x540_422_10b_bpp5_v1": (1920, 540, 10, 3, 5, 1, 1, 5, 0, 16, 1, 4.0, 0),
    "c5_7680x4320_rgbpacked_14b_bpp2_v1": (7680, 4320, 14, 21, 2, 1, 1, 5, 0, 16, 1, 16.0, 0),
    # the other decomp_v of each mixed-batch kind
    "c3a_1280x720_420_8b_bpp1_v1": (1280, 720, 8, 2, 1, 1, 1, 5, 0, 16, 1, 4.0, 0),
    "c3b_1920x1080_444_12b_bpp4_v1": (1920, 1080, 12, 4, 4, 1, 1, 5, 0, 16, 1, 4.0, 0),
    "c3c_3840x2160_rgb_10b_bpp3_v1": (3840, 2160, 10, 4, 3, 1, 1, 5, 0, 16, 1, 4.0, 0),
    "c3d_1920x540_422_10b_bpp5_v2": (1920, 540, 10, 3, 5, 1, 2, 5, 0, 16, 1, 4.0, 0),
}

# Quantisation known answers restated from tests/UnitTests/TestQuant.cc:27-136
# are read from the reference test source as data.


def sha(b):
    return hashlib.sha256(b).hexdigest()


def ref_encode(cfg, planes, td):
    W, H, D, F, bn, bd, V, Hh, q, sh, sig = cfg[:11]
    rc = cfg[13] if len(cfg) > 13 else 1
    vp, sg = (cfg[14], cfg[15]) if len(cfg) > 15 else (0, 0)
    O = load_oracle()
    raw = os.path.join(td, "in.raw")
    out = os.path.join(td, "out.jxs")
    S = load_synth()
    with open(raw, "wb") as f:
        f.write(S.frame_bytes(planes))
    cmd = [O.REF_ENC, raw, out, str(W), str(H), str(D), str(F), str(bn), str(bd), str(V), str(Hh), "1", "1", str(sh),
           str(q), str(sig), str(rc), str(vp), str(sg)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.returncode
    return open(out, "rb").read(), 0


def ref_decode(cs, td, proxy=0):
    O = load_oracle()
    p = os.path.join(td, "d.jxs")
    o = os.path.join(td, "d.raw")
    open(p, "wb").write(cs)
    r = subprocess.run([O.REF_DEC, p, o, "1", "1", str(proxy)], capture_output=True, text=True)
    if r.returncode != 0:
        return None
    return open(o, "rb").read()


def case_entry(name, cfg, td, keep):
    S = load_synth()
    W, H, D, F, bn, bd, V, Hh, q, sh, sig, sigma, frame = cfg[:13]
    planes = S.make_frame(W, H, D, F, frame=frame, sigma=sigma)
    if name.startswith("u_"):
        planes = S.flatten_rows(planes, H)
    cs, rc = ref_encode(cfg, planes, td)
    e = {"params": dict(zip(["width", "height", "depth", "format", "bpp_num", "bpp_den", "decomp_v", "decomp_h",
                             "quantization", "slice_height", "significance", "sigma", "frame"], cfg[:13])),
         "input_sha256": S.frame_sha256(planes)}
    e["params"]["rate_control_mode"] = cfg[13] if len(cfg) > 13 else 1
    if len(cfg) > 15:
        e["params"]["vertical_prediction"] = cfg[14]
        e["params"]["signs_handling"] = cfg[15]
    if name.startswith("u_"):
        e["params"]["pattern"] = "flatten_rows"
    if cs is None:
        e["reference_error"] = True
        return e
    e["codestream_sha256"] = sha(cs)
    e["codestream_bytes"] = len(cs)
    dec = ref_decode(cs, td)
    e["decoded_sha256"] = sha(dec) if dec is not None else None
    if keep:
        # proxy decoding at half and quarter resolution (None: the reference refuses it)
        for proxy, key in ((1, "decoded_half_sha256"), (2, "decoded_quarter_sha256")):
            pd = ref_decode(cs, td, proxy)
            e[key] = sha(pd) if pd is not None else None
        open(os.path.join(HERE, name + ".jxs"), "wb").write(cs)
    return e


def api_struct_sizes():
    """sizeof / offsetof of the reference API structs, from the reference
    headers (compiled here with gcc)."""
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
        subprocess.check_call(["gcc", "-I" + REF + "/Source/API", c, "-o", exe])
        return json.loads(subprocess.check_output([exe]).decode())


def quant_kats():
    src = open(os.path.join(REF, "tests/UnitTests/TestQuant.cc")).read()

    def arr(name):
        body = src[src.index(name):]
        body = body[body.index("{") + 1:body.index("};")]
        return [int(x, 16) if x.startswith("0x") else int(x) for x in re.findall(r"0x[0-9a-fA-F]+|\b\d+\b", body)]

    inp = arr("quant_input[QUANT_MAX_SIZE]")
    gclis = arr("gclis_input[GCLI_SIZE]")
    vec = src[src.index("quant_test_vector[5]"):]
    vec = vec[:vec.index("};\n\nclass")]
    tests = []
    for block in vec.split("/* Test")[1:]:
        nums = re.findall(r"0x[0-9a-fA-F]+|\b\d+\b|QUANT_TYPE_DEADZONE|QUANT_TYPE_UNIFORM", block)
        nums = nums[1:]  # drop the test number of the comment
        buf_len, gtli, qt = int(nums[0]), int(nums[1]), nums[2]
        out = [int(x, 16) if x.startswith("0x") else int(x) for x in nums[3:]]
        tests.append({"buf_len": buf_len, "gtli": gtli, "uniform": qt == "QUANT_TYPE_UNIFORM", "expected": out})
    return {"input": inp, "gclis": gclis, "tests": tests}


def sample_codestream():
    src = open(os.path.join(REF, "tests/UnitTests/SampleFramesData.cc")).read()
    body = src[src.index("Frame_Sample_1_16x16_8bit_422_bitstream[] = {"):]
    body = body[body.index("{") + 1:body.index("};")]
    return bytes(int(x, 16) for x in re.findall(r"0x[0-9A-Fa-f]+", body))


# codestreams swept: the reference's own unit-test sample (vertical
# prediction, one slice) and a two-slice codestream without it
SWEEPS = ["sample_16x16_422", "s_64x32_422_8b_bpp5_v2"]


def sweeps():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_sweep")
    for name in SWEEPS:
        r = subprocess.run([exe, os.path.join(HERE, name + ".jxs")], capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit("ref_sweep failed on %s: %s" % (name, r.stderr[-300:]))
        with gzip.open(os.path.join(HERE, "sweep_%s.txt.gz" % name), "wt", compresslevel=9) as f:
            f.write(r.stdout)
        print("sweep", name, len(r.stdout.splitlines()), "variants", flush=True)


def main():
    O = load_oracle()
    if not O.have_reference():
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    if "--sweeps" in sys.argv:
        sweeps()
        return
    out = {"small": {}, "full": {}}
    with tempfile.TemporaryDirectory() as td:
        for name, cfg in SMALL.items():
            out["small"][name] = case_entry(name, cfg, td, keep=True)
            print(name, "ok", flush=True)
        for name, cfg in FULL.items():
            out["full"][name] = case_entry(name, cfg, td, keep=False)
            print(name, "ok", flush=True)
        cs = sample_codestream()
        open(os.path.join(HERE, "sample_16x16_422.jxs"), "wb").write(cs)
        dec = ref_decode(cs, td)
        out["sample_16x16_422"] = {"codestream_sha256": sha(cs), "codestream_bytes": len(cs),
                                   "decoded_sha256": sha(dec), "decoded_bytes": len(dec),
                                   "uses_vertical_prediction": True}
        # the reference decoder's planes of the sample, kept as a file so a
        # mismatch can be located (16x16 + 2 x 8x16 bytes)
        open(os.path.join(HERE, "sample_16x16_422.yuv"), "wb").write(dec)
    out["api_struct_sizes"] = api_struct_sizes()
    out["quant_kat"] = quant_kats()
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    sweeps()


if __name__ == "__main__":
    main()
