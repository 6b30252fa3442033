"""GPU encoding with vertical prediction (coding_vertical_prediction_mode 1
and 2) and sign coding (coding_signs_handling 1 fast and 2 full), the coding
tools whose rate control runs precinct after precinct (k_enc_slice):
codestreams byte-identical to the reference encoder's (tests/golden t_*
cases) and to the oracle's over rate-control modes, quantisers, formats and
sizes; one plan mixing them with default-tool frames; and the decoders'
planes of these codestreams equal to the oracle's."""
import hashlib
import json
import os

import numpy as np
import pytest

from jxs_util import ROOT, load_oracle, load_package, load_synth

GOLDEN = os.path.join(ROOT, "tests", "golden")
G = json.load(open(os.path.join(GOLDEN, "golden.json")))
TOOLS = sorted(n for n in G["small"] if n.startswith("t_"))


def _enc(P, p, **kw):
    return P.Encoder(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], p["bpp_den"], p["decomp_v"],
                     p["decomp_h"], p["quantization"], p["slice_height"], p["significance"], verbose=0,
                     rate_control_mode=p.get("rate_control_mode", 1), vertical_prediction=p.get("vertical_prediction", 0),
                     signs_handling=p.get("signs_handling", 0), **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", TOOLS)
def test_gpu_tool_codestreams_equal_reference(name):
    print("tools encode", name, flush=True)
    P, S = load_package(), load_synth()
    p = G["small"][name]["params"]
    planes = S.make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"], sigma=p["sigma"])
    assert S.frame_sha256(planes) == G["small"][name]["input_sha256"]
    enc = _enc(P, p)
    for _ in range(2):
        rc, cs = enc.encode(planes)
        assert rc == 0
        assert cs == open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    enc.close()


# (W, H, depth, format, bpp, decomp_v, decomp_h, quantization, significance, rc mode, vpred, signs)
CASES = [
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 1, 1, 0),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 1, 2, 0),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 1, 0, 1),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 1, 0, 2),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 0, 1, 1),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 2, 1, 1),
    (640, 360, 10, 3, 3, 2, 5, 0, 1, 3, 2, 1),
    (352, 288, 12, 4, 4, 2, 5, 1, 1, 1, 1, 2),
    (352, 288, 12, 4, 4, 2, 5, 1, 1, 0, 2, 2),
    (352, 288, 12, 4, 4, 1, 5, 1, 0, 1, 1, 1),
    (320, 240, 8, 2, 2, 2, 5, 0, 1, 1, 1, 1),
    (320, 240, 8, 2, 2, 1, 3, 0, 1, 3, 2, 0),
    (250, 130, 9, 3, 1, 0, 2, 0, 1, 1, 1, 2),
    (192, 128, 14, 21, 2, 1, 5, 0, 1, 1, 2, 2),
    (1920, 1080, 8, 3, 5, 2, 5, 0, 1, 1, 1, 1),
]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", CASES, ids=lambda c: "x".join(str(v) for v in c))
def test_gpu_tools_equal_oracle(cfg):
    print("tools oracle", cfg, flush=True)
    P, S, O = load_package(), load_synth(), load_oracle()
    W, H, D, F, bpp, V, Hh, q, sig, rcm, vp, sg = cfg
    planes = S.make_frame(W, H, D, F, frame=11)
    e, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, rcm, vp, sg), planes)
    assert e == 0
    enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, verbose=0, rate_control_mode=rcm, vertical_prediction=vp,
                    signs_handling=sg)
    rc, cs = enc.encode(planes)
    enc.close()
    assert rc == 0
    if cs != ref:
        k = next(i for i in range(min(len(cs), len(ref))) if cs[i] != ref[i])
        raise AssertionError("first difference at byte %d of %d" % (k, len(ref)))
    # and back: the GPU decoder's planes equal the oracle's decode
    e, planes_ref = O.decode(ref)
    assert e == 0
    dec = P.Decoder(cs, verbose=0)
    rc, planes_gpu = dec.decode(cs)
    dec.close()
    assert rc == 0
    for a, b in zip(planes_gpu, planes_ref):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_plan_mixes_tool_and_default_frames():
    """One plan, one launch per stage: default-tool frames go through the
    packet-parallel rate control and packer, tool frames through the slice
    kernel, each codestream equal to the oracle's; replayed from a HIP graph
    three times, every replay checked."""
    import torch
    print("tools plan", flush=True)
    P, S, O = load_package(), load_synth(), load_oracle()
    dev = torch.device("cuda:0")
    mix = [(256, 64, 10, 3, 3, 2, 5, 0, 1, 1, 0, 0), (256, 64, 10, 3, 3, 2, 5, 0, 1, 1, 1, 1),
           (200, 200, 8, 4, 3, 1, 5, 0, 1, 0, 2, 0), (160, 90, 12, 4, 4, 2, 5, 1, 1, 1, 0, 2),
           (256, 128, 8, 2, 2, 2, 5, 0, 1, 3, 0, 0), (96, 64, 14, 21, 2, 1, 5, 0, 1, 1, 1, 2)]
    items, specs = [], []
    for k, (W, H, D, F, bpp, V, Hh, q, sig, rcm, vp, sg) in enumerate(mix):
        planes = S.make_frame(W, H, D, F, frame=30 + k)
        e, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, rcm, vp, sg), planes)
        assert e == 0
        enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, verbose=0, rate_control_mode=rcm,
                        vertical_prediction=vp, signs_handling=sg)
        d_in = [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]
        d_out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
        items.append((enc, d_in, d_out, ref))
        specs.append((enc, [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in], d_out.data_ptr()))
    plan = P.Plan(specs)
    stream = torch.cuda.current_stream(dev)
    for rep in range(4):
        for _, _, d_out, _ in items:
            d_out.zero_()
        torch.cuda.synchronize()
        plan.run(stream.cuda_stream, use_graph=rep > 0)
        torch.cuda.synchronize()
        assert plan.status()[0] == 0
        for k, (_, _, d_out, ref) in enumerate(items):
            assert d_out.cpu().numpy().tobytes()[:len(ref)] == ref, "run %d frame %d %s" % (rep, k, mix[k])
    plan.destroy()
    for enc, _, _, _ in items:
        enc.close()


@pytest.mark.gpu
def test_full_sign_coding_with_slice_rate_control_is_refused():
    """Full sign coding with rate-control modes 2 and 3 is the one
    combination not restated; init refuses it instead of writing another
    codestream.  Out-of-range modes are refused as by EncHandle.c:242-265."""
    P = load_package()
    for rcm, vp, sg in ((2, 0, 2), (3, 1, 2), (1, 3, 0), (1, 0, 3)):
        with pytest.raises(P.JxsError) as ei:
            P.Encoder(256, 64, 10, 3, 3, verbose=0, rate_control_mode=rcm, vertical_prediction=vp, signs_handling=sg)
        assert ei.value.code == P.ERR_BAD_PARAMETER
