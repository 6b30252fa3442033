"""GPU decoder parity: decoded planes against the oracle's decoder, through
the SvtJpegxsDec C-ABI and through device-resident decode plans."""
import numpy as np
import pytest

from jxs_util import ROOT, load_oracle, load_package, load_synth

CASES = [
    (64, 32, 8, 3, 5, 2, 5, 0, 1),
    (256, 64, 10, 3, 3, 2, 5, 0, 1),
    (200, 200, 8, 4, 3, 1, 5, 0, 1),
    (201, 203, 10, 4, 2, 2, 5, 0, 1),
    (256, 128, 8, 2, 2, 2, 5, 0, 1),
    (128, 96, 8, 2, 2, 1, 3, 0, 1),
    (160, 90, 12, 4, 4, 2, 5, 1, 1),
    (160, 90, 10, 3, 5, 2, 5, 0, 0),
    (96, 64, 14, 21, 2, 1, 5, 0, 1),
    (128, 64, 11, 4, 3, 0, 3, 0, 1),
    (1024, 768, 10, 3, 20, 2, 5, 0, 1),
    # no vertical level, rows written straight to the output plane: 8-bit
    # (4-sample stores), odd widths (per-sample tails), 16-bit
    (96, 40, 8, 3, 3, 0, 2, 0, 1),
    (131, 37, 8, 4, 4, 0, 1, 0, 1),
    (250, 30, 12, 3, 3, 0, 4, 0, 1),
    # one 7,680-sample line of three components per precinct at bpp 20: a
    # packet of about 19 KB, longer than the unpack's LDS stage (12 KiB), so
    # it is read from HBM
    (7680, 6, 8, 4, 20, 0, 1, 0, 1),
    # the widest rows taken (8,192 samples): full-width rows through the
    # row kernel sized for them, with and without vertical levels
    (8192, 16, 10, 3, 3, 1, 5, 0, 1),
    (8192, 8, 8, 4, 4, 0, 2, 0, 1),
    (8192, 18, 12, 21, 2, 2, 5, 0, 1),
]


def _codestream(cfg, frame=0, sigma=4.0):
    W, H, D, F, bpp, V, Hh, q, sig = cfg
    S, O = load_synth(), load_oracle()
    planes = S.make_frame(W, H, D, F, frame=frame, sigma=sigma)
    e, cs = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, q, 16, sig), planes)
    assert e == 0
    return planes, cs


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", CASES, ids=lambda c: "x".join(str(v) for v in c))
def test_decode_matches_oracle(cfg):
    print("case", cfg, flush=True)
    P, O = load_package(), load_oracle()
    _, cs = _codestream(cfg)
    e, ref = O.decode(cs)
    assert e == 0
    dec = P.Decoder(cs)
    rc, got = dec.decode(cs)
    dec.close()
    assert rc == 0
    assert len(got) == len(ref)
    for c, (a, b) in enumerate(zip(got, ref)):
        assert a.shape == b.shape
        assert np.array_equal(a, b), "component %d differs at %s" % (c, np.argwhere(a != b)[:3].tolist())


@pytest.mark.gpu
def test_decode_corrupt_codestream_returns_error_or_frame():
    """Bit flips in the entropy-coded data must never read outside the
    codestream: the decoder returns (an error or a frame), the process lives."""
    P = load_package()
    _, cs = _codestream((256, 64, 10, 3, 3, 2, 5, 0, 1))
    rng = np.random.default_rng(5)
    dec = P.Decoder(cs)
    for k in range(8):
        bad = bytearray(cs)
        for pos in rng.integers(200, len(cs) - 2, size=8):
            bad[int(pos)] ^= int(rng.integers(1, 256))
        rc, _ = dec.decode(bytes(bad))
        assert rc in (0, P.ERR_DEC_INVALID_BITSTREAM, P.ERR_DEC_TOO_SHORT)
    rc, _ = dec.decode(cs[: len(cs) // 2])
    assert rc != 0
    dec.close()


@pytest.mark.gpu
def test_device_roundtrip_packed_rgb14_v1():
    """Config-5 shape at reduced size: packed RGB 14-bit, bpp 2, V1 H5, sigma
    16; GPU encode plan -> GPU decode plan fed with the encoder's precinct
    table; codestream and planes bit-exact with the oracle."""
    import torch
    P, O, S = load_package(), load_oracle(), load_synth()
    W, H, D, F = 1024, 512, 14, 21
    frames = [S.make_frame(W, H, D, F, frame=i, sigma=16.0) for i in range(3)]
    enc = P.Encoder(W, H, D, F, 2, 1, 1, 5)
    dev = torch.device("cuda:0")
    d_in = [torch.from_numpy(f[0].view(np.int16)).to(dev) for f in frames]
    d_cs = torch.zeros((3, (enc.bytes_per_frame + 3) & ~3), dtype=torch.uint8, device=dev)
    plan = P.Plan([(enc, [d_in[i].data_ptr()], [W * 3], d_cs[i].data_ptr()) for i in range(3)])
    plan.run()
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    cs0 = d_cs[0].cpu().numpy().tobytes()
    dec = P.Decoder(cs0)
    d_out = [[torch.zeros((H, W), dtype=torch.int16, device=dev) for _ in range(3)] for _ in range(3)]
    specs = []
    for i in range(3):
        prec, stride = plan.precinct_offsets(i)
        specs.append((dec, d_cs[i].data_ptr(), enc.bytes_per_frame, [t.data_ptr() for t in d_out[i]], [W] * 3,
                      prec if i != 2 else None, stride))
    dplan = P.DecodePlan(specs)
    dplan.run()
    torch.cuda.synchronize()
    assert dplan.status()[0] == 0
    prm = O.make_params(W, H, D, F, 2, 1, 1, 5)
    for i in range(3):
        e, ref_cs = O.encode(prm, frames[i])
        got_cs = d_cs[i].cpu().numpy().tobytes()
        assert e == 0 and got_cs == ref_cs, "codestream %d differs" % i
        e, ref = O.decode(ref_cs)
        assert e == 0
        for c in range(3):
            got = d_out[i][c].cpu().numpy().view(np.uint16)
            assert np.array_equal(got, ref[c]), "frame %d component %d differs" % (i, c)
    dplan.destroy()
    plan.destroy()


# ---- decoded planes at full size against the reference decoder's (tests/golden) ----
import hashlib  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402

from jxs_util import ROOT  # noqa: E402

_GOLDEN = os.path.join(ROOT, "tests", "golden")
with open(os.path.join(_GOLDEN, "golden.json")) as _f:
    _G = json.load(_f)
_FULL = _G["full"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(n for n in _FULL if _FULL[n].get("decoded_sha256")))
def test_gpu_decode_full_size_equals_reference_decoder(name):
    """First frame of every benchmark workload at full size (1080p .. 8K):
    GPU encode (whose codestream hashes to the reference's), then the GPU
    decoder through the C API; the decoded planes, concatenated as the
    reference decoder writes them, hash to the reference decoder's output."""
    p = _FULL[name]["params"]
    print("full decode", name, flush=True)
    P, S = load_package(), load_synth()
    planes = S.make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"], sigma=p["sigma"])
    enc = P.Encoder(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], p["bpp_den"], p["decomp_v"],
                    p["decomp_h"], p["quantization"], p["slice_height"], p["significance"], verbose=0)
    rc, cs = enc.encode(planes)
    enc.close()
    assert rc == 0 and hashlib.sha256(cs).hexdigest() == _FULL[name]["codestream_sha256"]
    dec = P.Decoder(cs, verbose=0)
    rc, out = dec.decode(cs)
    dec.close()
    assert rc == 0
    got = hashlib.sha256(b"".join(np.ascontiguousarray(q).tobytes() for q in out)).hexdigest()
    assert got == _FULL[name]["decoded_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("proxy", [1, 2])
@pytest.mark.parametrize("name", sorted(n for n in _G["small"] if "decoded_half_sha256" in _G["small"][n]
                                        and not _G["small"][n].get("reference_error")))
def test_gpu_proxy_decode_equals_reference(name, proxy):
    """Decoding at half / quarter resolution (proxy_mode) through the C-ABI:
    the planes hash to the reference decoder's, and a mode the codestream's
    levels do not allow is refused at init as by the reference."""
    print("proxy", proxy, name, flush=True)
    P, O = load_package(), load_oracle()
    entry = _G["small"][name]
    cs = open(os.path.join(_GOLDEN, name + ".jxs"), "rb").read()
    want = entry["decoded_half_sha256" if proxy == 1 else "decoded_quarter_sha256"]
    if want is None:
        with pytest.raises(P.JxsError):
            P.Decoder(cs, verbose=0, proxy_mode=proxy)
        return
    dec = P.Decoder(cs, verbose=0, proxy_mode=proxy)
    rc, got = dec.decode(cs)
    dec.close()
    assert rc == 0
    e, ref = O.decode(cs, proxy)
    assert e == 0
    for c, (a, b) in enumerate(zip(got, ref)):
        assert a.shape == b.shape
        assert np.array_equal(a, b), "component %d differs" % c
    assert hashlib.sha256(b"".join(p.tobytes() for p in got)).hexdigest() == want


@pytest.mark.gpu
def test_decode_plan_mixes_full_and_proxy_decoders():
    """One device-resident decode batch of three codestreams decoded at full,
    half and quarter resolution (a decoder per mode): each writes its own
    plane sizes and equals the oracle's decode at that resolution."""
    import torch
    dev = torch.device("cuda:0")
    print("decode plan: full + half + quarter", flush=True)
    P, O = load_package(), load_oracle()
    cfg = (256, 64, 10, 3, 3, 2, 5, 0, 1)
    _, cs = _codestream(cfg, frame=2)
    d_cs = torch.from_numpy(np.frombuffer(cs, dtype=np.uint8).copy()).to(dev)
    specs, outs, refs, decs = [], [], [], []
    for proxy in (0, 1, 2):
        dec = P.Decoder(cs, verbose=0, proxy_mode=proxy)
        e, ref = O.decode(cs, proxy)
        assert e == 0
        planes = [torch.zeros(r.shape, dtype=torch.int16, device=dev) for r in ref]
        specs.append((dec, d_cs.data_ptr(), len(cs), [t.data_ptr() for t in planes], [r.shape[1] for r in ref], None, 1))
        outs.append(planes)
        refs.append(ref)
        decs.append(dec)
    plan = P.DecodePlan(specs)
    plan.run(torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    for proxy, (planes, ref) in enumerate(zip(outs, refs)):
        for c, (t, r) in enumerate(zip(planes, ref)):
            got = t.cpu().numpy().view(np.uint16)
            assert np.array_equal(got, r), "proxy %d component %d" % (proxy, c)
    plan.destroy()
    for dec in decs:
        dec.close()


@pytest.mark.gpu
def test_send_packet_decodes_chunked_codestreams():
    """svt_jpeg_xs_decoder_send_packet in packetization mode
    (DecThreadInit.c:347-460): three codestreams fed back to back, as
    DecAppMain.c:218-239 does, in 700-byte chunks and then in chunks larger
    than two frames.  A call takes at most the rest of the current frame
    (bytes_used), returns BitstreamTooShort until the frame is complete, then
    None; get_frame returns the frames in order, equal to the oracle's decode,
    then EndOfCodestream after send_eoc."""
    import ctypes
    print("send_packet chunks", flush=True)
    P, O = load_package(), load_oracle()
    cfg = (256, 64, 10, 3, 3, 2, 5, 0, 1)
    css = [_codestream(cfg, frame=k)[1] for k in range(3)]
    L = P.lib()
    for chunk_size in (700, 2 * len(css[0]) + 100):
        dec = P.Decoder(css[0], verbose=0, packetization_mode=1)
        c = dec.config

        def new_image():
            planes = [np.zeros((c.components[k].height, c.components[k].width), dtype=np.uint16)
                      for k in range(c.components_num)]
            f = P.Frame()
            for k, p in enumerate(planes):
                f.image.data_yuv[k] = p.ctypes.data
                f.image.stride[k] = p.shape[1]
                f.image.alloc_size[k] = p.nbytes
            return planes, f

        stream = np.frombuffer(b"".join(css), dtype=np.uint8).copy()
        pos, sent = 0, []
        while pos < len(stream):
            planes, f = new_image()
            while True:
                n = min(chunk_size, len(stream) - pos)
                f.bitstream.buffer = stream[pos:].ctypes.data
                f.bitstream.allocation_size = n
                f.bitstream.used_size = n
                used = ctypes.c_uint32(0)
                rc = L.svt_jpeg_xs_decoder_send_packet(ctypes.byref(dec.api), ctypes.byref(f), ctypes.byref(used))
                assert 0 < used.value <= n
                pos += used.value
                if rc != P.ERR_DEC_TOO_SHORT:
                    break
            assert rc == 0
            sent.append(planes)
        assert len(sent) == 3 and pos == len(stream)
        assert L.svt_jpeg_xs_decoder_send_eoc(ctypes.byref(dec.api)) == 0
        for k in range(3):
            out = P.Frame()
            assert L.svt_jpeg_xs_decoder_get_frame(ctypes.byref(dec.api), ctypes.byref(out), 1) == 0
        out = P.Frame()
        assert L.svt_jpeg_xs_decoder_get_frame(ctypes.byref(dec.api), ctypes.byref(out), 1) == P.ERR_DEC_EOC
        dec.close()
        for k, got in enumerate(sent):
            e, ref = O.decode(css[k])
            assert e == 0
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), "chunk %d frame %d" % (chunk_size, k)


@pytest.mark.gpu
def test_decoder_follows_qpih_per_frame():
    """Qpih is a per-frame header field (ParseHeader.c:238-241, used per
    precinct at Decoder.c:576): one decoder fed a deadzone codestream, then a
    uniform one of the same geometry, and so on, decodes each to the oracle's
    planes; a codestream with the reserved Qpih = 2 is refused."""
    print("qpih per frame", flush=True)
    P, O = load_package(), load_oracle()
    dz = _codestream((160, 90, 12, 4, 4, 2, 5, 0, 1))[1]
    un = _codestream((160, 90, 12, 4, 4, 2, 5, 1, 1))[1]
    assert dz != un
    dec = P.Decoder(dz, verbose=0)
    for cs in (dz, un, dz, un):
        rc, got = dec.decode(cs)
        assert rc == 0
        e, ref = O.decode(cs)
        assert e == 0
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    # PIH byte 27 (from the marker) holds Lh, Rl, Qpih (2 bits), Fs, Rm
    bad = bytearray(dz)
    off = bytes(bad).index(b"\xff\x12") + 27
    assert (bad[off] >> 4) & 3 == 0
    bad[off] = (bad[off] & 0xCF) | 0x20
    rc, _ = dec.decode(bytes(bad))
    assert rc == P.ERR_DEC_INVALID_BITSTREAM
    dec.close()


_TOOLS = sorted(n for n in _G["small"] if n.startswith("t_")) + ["sample_16x16_422"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", _TOOLS)
def test_decode_vertical_prediction_and_sign_codestreams(name):
    """Reference codestreams coded with vertical prediction (D[p,b] bit 0,
    both run modes) and separate sign sub-packets (Fs = 1, fast and full),
    and the reference's own unit-test sample (SampleFramesData.cc:8-33), which
    uses vertical prediction: the GPU decoder's planes hash to the reference
    decoder's."""
    print("tools decode", name, flush=True)
    P = load_package()
    cs = open(os.path.join(_GOLDEN, name + ".jxs"), "rb").read()
    want = (_G["small"][name] if name in _G["small"] else _G[name])["decoded_sha256"]
    dec = P.Decoder(cs, verbose=0)
    for _ in range(2):
        rc, planes = dec.decode(cs)
        assert rc == 0
        assert hashlib.sha256(b"".join(p.tobytes() for p in planes)).hexdigest() == want
    dec.close()


@pytest.mark.gpu
def test_decode_plan_of_mixed_tool_codestreams():
    """One device-resident decode plan holding codestreams with and without
    vertical prediction and sign sub-packets: slices that need in-order
    decoding and slices decoded packet-parallel side by side, each frame equal
    to the oracle's decode."""
    import torch
    print("tools decode plan", flush=True)
    P, O = load_package(), load_oracle()
    dev = torch.device("cuda:0")
    names = ["t_256x64_422_10b_bpp3_v2_vp1", "s_256x64_422_10b_bpp3_v2", "t_256x64_422_10b_bpp3_v2_sg2",
             "t_200x200_444_8b_bpp3_v1_vp1_rc0", "t_96x64_rgbpacked_14b_bpp2_v1_vp1_sg2"]
    decs, specs, outs, refs, keep = [], [], [], [], []
    for n in names:
        cs = open(os.path.join(_GOLDEN, n + ".jxs"), "rb").read()
        e, ref = O.decode(cs)
        assert e == 0
        dec = P.Decoder(cs, verbose=0)
        decs.append(dec)
        d_cs = torch.from_numpy(np.frombuffer(cs, dtype=np.uint8).copy()).to(dev)
        c = dec.config
        planes = [torch.zeros((c.components[k].height, c.components[k].width),
                              dtype=torch.int16 if c.bit_depth > 8 else torch.uint8, device=dev)
                  for k in range(c.components_num)]
        keep.append(d_cs)
        outs.append(planes)
        refs.append(ref)
        specs.append((dec, d_cs.data_ptr(), len(cs), [t.data_ptr() for t in planes], [t.shape[1] for t in planes],
                      None, 0))
    plan = P.DecodePlan(specs)
    plan.run(torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    for n, planes, ref in zip(names, outs, refs):
        for t, r in zip(planes, ref):
            got = t.cpu().numpy()
            got = got.view(np.uint16) if got.dtype == np.int16 else got
            assert np.array_equal(got, r), n
    plan.destroy()
    for dec in decs:
        dec.close()


@pytest.mark.gpu
def test_decode_plan_parts_report_each_frames_status():
    """A decode plan of six frames runs as four parts on streams of its own
    (2 + 2 + 1 + 1 frames) over one array of status words.  Frame 3's
    codestream is cut short: its precinct walk fails and only its status word
    is set; every other frame equals the oracle's decode.  A second run
    starts from cleared status words and reports the same."""
    import torch
    print("decode plan parts: statuses", flush=True)
    P, O = load_package(), load_oracle()
    dev = torch.device("cuda:0")
    cfg = (256, 64, 10, 3, 3, 2, 5, 0, 1)
    css = [_codestream(cfg, frame=i)[1] for i in range(6)]
    dec = P.Decoder(css[0], verbose=0)
    keep, specs, outs = [], [], []
    for i, cs in enumerate(css):
        d_cs = torch.from_numpy(np.frombuffer(cs, dtype=np.uint8).copy()).to(dev)
        planes = [torch.zeros((dec.config.components[c].height, dec.config.components[c].width), dtype=torch.int16,
                              device=dev) for c in range(dec.config.components_num)]
        size = len(cs) // 2 if i == 3 else len(cs)
        specs.append((dec, d_cs.data_ptr(), size, [t.data_ptr() for t in planes], [t.shape[1] for t in planes],
                      None, 1))
        keep.append(d_cs)
        outs.append(planes)
    plan = P.DecodePlan(specs)
    try:
        for _ in range(2):
            for planes in outs:
                for t in planes:
                    t.zero_()
            plan.run(torch.cuda.current_stream(dev).cuda_stream)
            torch.cuda.synchronize()
            rc, errs = plan.status()
            assert rc != 0
            assert [e != 0 for e in errs] == [False, False, False, True, False, False], errs
            for i, cs in enumerate(css):
                if i == 3:
                    continue
                e, ref = O.decode(cs)
                assert e == 0
                for t, r in zip(outs[i], ref):
                    assert np.array_equal(t.cpu().numpy().view(np.uint16), r), "frame %d differs" % i
    finally:
        plan.destroy()
        dec.close()
