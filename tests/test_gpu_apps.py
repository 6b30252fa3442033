"""The reference's own command-line callers, unmodified, on the MI355X library.

Source/App/EncApp and DecApp are compiled from the reference's sources against
include/ and linked with libsvtjpegxs_amd.so (oracle/Makefile.ref:
enc_app_amd, dec_app_amd).  They drive the library exactly as they drive the
CPU library: frame pools for their buffers (EncAppMain.c:113-166, 258-348;
DecAppMain.c:121-160, 378-489), a sender thread and a blocking receive loop,
slice packetization and packet input when asked.  Their output files must
equal the reference encoder's codestream and the reference decoder's planes
for the same seeded frame (SHA-256 in tests/golden/golden.json).
"""
import hashlib
import json
import os
import subprocess

import pytest

from jxs_util import ROOT, load_synth

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
ENC_APP = os.path.join(ROOT, "oracle", "_ref", "enc_app_amd")
DEC_APP = os.path.join(ROOT, "oracle", "_ref", "dec_app_amd")
FORMATS = {2: "yuv420", 3: "yuv422", 4: "yuv444"}
CASES = ["c1_1920x1080_422_8b_bpp5_v2", "c2_3840x2160_422_10b_bpp3_v2", "c3a_1280x720_420_8b_bpp1_v2"]


def sha(b):
    return hashlib.sha256(b).hexdigest()


def _run(cmd):
    print(" ".join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd)), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, "%s exited %d:\n%s" % (os.path.basename(cmd[0]), r.returncode, r.stderr[-2000:])
    return r


@pytest.mark.gpu
@pytest.mark.skipif(not (os.path.exists(ENC_APP) and os.path.exists(DEC_APP)),
                    reason="reference apps not built (make -f oracle/Makefile.ref needs the reference sources)")
@pytest.mark.parametrize("packetized", [0, 1])
@pytest.mark.parametrize("name", CASES)
def test_reference_apps_encode_and_decode_through_the_library(name, packetized, tmp_path):
    G = GOLDEN["full"][name]
    p = G["params"]
    S = load_synth()
    planes = S.make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"], sigma=p["sigma"])
    raw = S.frame_bytes(planes)
    assert sha(raw) == G["input_sha256"]
    yuv, jxs, out = (str(tmp_path / n) for n in ("in.yuv", "out.jxs", "dec.yuv"))
    open(yuv, "wb").write(raw)
    n = 3  # the app loops its one-frame input file
    _run([ENC_APP, "-i", yuv, "-b", jxs, "-w", str(p["width"]), "-h", str(p["height"]),
          "--input-depth", str(p["depth"]), "--colour-format", FORMATS[p["format"]], "--bpp", str(p["bpp_num"]),
          "--decomp_v", str(p["decomp_v"]), "--decomp_h", str(p["decomp_h"]), "-n", str(n),
          "--packetization-mode", str(packetized), "--no-progress", "1", "-v", "1"])
    cs = open(jxs, "rb").read()
    k = G["codestream_bytes"]
    assert len(cs) == n * k
    for f in range(n):
        assert sha(cs[f * k:(f + 1) * k]) == G["codestream_sha256"], "EncApp frame %d" % f
    _run([DEC_APP, "-i", jxs, "-o", out, "--packetization-mode", str(packetized), "-v", "1"])
    dec = open(out, "rb").read()
    assert len(dec) == n * len(raw)
    for f in range(n):
        assert sha(dec[f * len(raw):(f + 1) * len(raw)]) == G["decoded_sha256"], "DecApp frame %d" % f
