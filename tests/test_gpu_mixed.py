"""The mixed-batch workload at full size: every stream kind of the mix with
both of its decomp_v values, in ONE plan (one launch per stage for all eight
frames), each codestream hashed against the reference encoder's
(tests/golden/golden.json, full cases c3a..c3d).

The 1920x540 kind is one field of interlaced 1080i content.  It is fed the way
an interlaced source hands it over (EncHandle.c:974-990; README.md:127-130):
the field's rows interleaved with the other field's in a 1920x1080 buffer,
stride 2 x 1920, the last row of the field the last row of the buffer.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from jxs_util import ROOT, load_package, load_synth

G = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))["full"]
MIX = ["c3a_1280x720_420_8b_bpp1_v2", "c3a_1280x720_420_8b_bpp1_v1", "c3b_1920x1080_444_12b_bpp4_v2",
       "c3b_1920x1080_444_12b_bpp4_v1", "c3c_3840x2160_rgb_10b_bpp3_v2", "c3c_3840x2160_rgb_10b_bpp3_v1",
       "c3d_1920x540_422_10b_bpp5_v1", "c3d_1920x540_422_10b_bpp5_v2"]


def _sha(b):
    return hashlib.sha256(b).hexdigest()


def _frame(name):
    p = G[name]["params"]
    planes = load_synth().make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"],
                                     sigma=p["sigma"])
    assert load_synth().frame_sha256(planes) == G[name]["input_sha256"]
    return p, planes


def _interleave(field, second=True):
    """A frame buffer twice the field's height holding the field in its odd
    (second) rows and a foreign pattern in the others; returns the buffer and
    the field's first-row offset in elements."""
    h, w = field.shape
    buf = np.full((2 * h, w), 0x2AA if field.dtype == np.uint16 else 0xAA, dtype=field.dtype)
    buf[1::2] = field
    return buf, (w if second else 0)


@pytest.mark.gpu
def test_mixed_kinds_both_levels_in_one_plan():
    import torch
    P = load_package()
    dev = torch.device("cuda:0")
    print("mixed plan of", len(MIX), "full-size frames", flush=True)
    keep, specs, outs, encs = [], [], [], []
    for name in MIX:
        p, planes = _frame(name)
        enc = P.Encoder(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], p["bpp_den"],
                        p["decomp_v"], p["decomp_h"], p["quantization"], p["slice_height"], p["significance"],
                        verbose=0)
        encs.append(enc)
        ptrs, strides = [], []
        for pl in planes:
            if name.startswith("c3d"):
                buf, off = _interleave(pl)
                t = torch.from_numpy(buf.view(np.int16)).to(dev)
                ptrs.append(t.data_ptr() + off * 2)
                strides.append(2 * pl.shape[1])
            else:
                t = torch.from_numpy(pl.view(np.int16) if pl.dtype == np.uint16 else pl).to(dev)
                ptrs.append(t.data_ptr())
                strides.append(pl.shape[1])
            keep.append(t)
        out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
        outs.append(out)
        specs.append((enc, ptrs, strides, out.data_ptr()))
    plan = P.Plan(specs)
    plan.run(torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    for name, out in zip(MIX, outs):
        cs = out.cpu().numpy().tobytes()[:G[name]["codestream_bytes"]]
        assert _sha(cs) == G[name]["codestream_sha256"], name
    plan.destroy()
    for enc in encs:
        enc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c3d_1920x540_422_10b_bpp5_v1", "c3d_1920x540_422_10b_bpp5_v2"])
def test_interlaced_field_through_send_picture(name):
    """The same field through send_picture from host memory: planes at the
    second field's first row, stride 2 x width, allocation ending with the
    field's last row (the smallest the reference accepts)."""
    P = load_package()
    print("field through send_picture", name, flush=True)
    p, planes = _frame(name)
    enc = P.Encoder(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], decomp_v=p["decomp_v"], verbose=0)
    bufs = [_interleave(pl) for pl in planes]
    f = P.Frame()
    for c, ((buf, off), pl) in enumerate(zip(bufs, planes)):
        h, w = pl.shape
        f.image.data_yuv[c] = buf.ctypes.data + off * 2
        f.image.stride[c] = 2 * w
        f.image.alloc_size[c] = (2 * w) * 2 * (h - 1) + w * 2
    out = np.zeros(enc.bytes_per_frame, dtype=np.uint8)
    f.bitstream.buffer = out.ctypes.data
    f.bitstream.allocation_size = out.nbytes
    L = P.lib()
    assert L.svt_jpeg_xs_encoder_send_picture(ctypes.byref(enc.api), ctypes.byref(f), 1) == 0
    rc, g = enc.get()
    assert rc == 0 and g.bitstream.used_size == G[name]["codestream_bytes"]
    assert _sha(out[:g.bitstream.used_size].tobytes()) == G[name]["codestream_sha256"]
    # one byte less than the reference's minimum is refused
    f.image.alloc_size[0] -= 1
    assert L.svt_jpeg_xs_encoder_send_picture(ctypes.byref(enc.api), ctypes.byref(f), 1) == P.ERR_BAD_PARAMETER
    enc.close()
