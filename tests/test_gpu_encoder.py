"""GPU encoder parity: the HIP path through the C-ABI against the oracle."""
import numpy as np
import pytest

from jxs_util import load_oracle, load_package, load_synth

# (width, height, depth, format, bpp, decomp_v, decomp_h, quantization, significance)
SMALL = [
    (64, 32, 8, 3, 5, 2, 5, 0, 1),
    (256, 64, 10, 3, 3, 2, 5, 0, 1),
    (200, 200, 8, 4, 3, 1, 5, 0, 1),
    (201, 203, 10, 4, 2, 2, 5, 0, 1),
    (128, 96, 8, 2, 1, 2, 5, 0, 1),  # too few bytes: both must report a frame error
    (256, 128, 8, 2, 2, 2, 5, 0, 1),
    (128, 96, 8, 2, 2, 1, 3, 0, 1),
    (160, 90, 12, 4, 4, 2, 5, 1, 1),
    (160, 90, 10, 3, 5, 2, 5, 0, 0),
    (96, 64, 14, 21, 2, 1, 5, 0, 1),
    (96, 64, 8, 21, 3, 2, 5, 0, 1),
    (128, 64, 11, 4, 3, 0, 3, 0, 1),
    (130, 66, 9, 3, 4, 1, 2, 0, 1),
    # packed 16-bit rows read as 16-byte vectors (width a multiple of 8) at every
    # vertical depth, and a width that is not (per-sample loads), odd height
    (104, 40, 12, 21, 3, 2, 5, 0, 1),
    (64, 30, 10, 21, 3, 0, 3, 0, 1),
    (100, 33, 14, 21, 2, 1, 5, 0, 1),
]


def _encode_both(cfg, frame=0, sigma=4.0):
    W, H, D, F, bpp, V, Hh, q, sig = cfg
    S, O, P = load_synth(), load_oracle(), load_package()
    planes = S.make_frame(W, H, D, F, frame=frame, sigma=sigma)
    e_ref, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, q, 16, sig), planes)
    enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, q, 16, sig)
    rc, got = enc.encode(planes)
    enc.close()
    return e_ref, ref, rc, got


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", SMALL, ids=lambda c: "x".join(str(v) for v in c))
def test_small_configs_bit_exact(cfg):
    print("case", cfg, flush=True)
    e_ref, ref, rc, got = _encode_both(cfg)
    assert rc == e_ref
    if e_ref:
        return
    assert len(got) == len(ref)
    if got != ref:
        a = np.frombuffer(got, np.uint8)
        b = np.frombuffer(ref, np.uint8)
        first = int(np.nonzero(a != b)[0][0])
        pytest.fail("codestream differs from byte %d of %d" % (first, len(ref)))


# the widest rows taken (8,192 samples, DCI 8K): the transform's 256 threads
# of 32 columns cover one row; 4:2:2 with a vertical level, 4:4:4 without,
# packed RGB with two
WIDEST = [
    (8192, 16, 10, 3, 3, 1, 5, 0, 1),
    (8192, 8, 8, 4, 4, 0, 2, 0, 1),
    (8192, 18, 12, 21, 2, 2, 5, 0, 1),
]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", WIDEST, ids=lambda c: "x".join(str(v) for v in c))
def test_widest_rows_bit_exact(cfg):
    print("case", cfg, flush=True)
    e_ref, ref, rc, got = _encode_both(cfg)
    assert e_ref == 0 and rc == 0
    assert got == ref


@pytest.mark.gpu
def test_1080p_422_8bit_bpp5_bit_exact():
    print("case 1920x1080 422 8-bit bpp5", flush=True)
    e_ref, ref, rc, got = _encode_both((1920, 1080, 8, 3, 5, 2, 5, 0, 1))
    assert e_ref == 0 and rc == 0
    assert got == ref


# ---- against the reference's own codestreams (tests/golden) ----
import hashlib  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402

from jxs_util import ROOT  # noqa: E402

_GOLDEN = os.path.join(ROOT, "tests", "golden")
with open(os.path.join(_GOLDEN, "golden.json")) as _f:
    _G = json.load(_f)


def _gpu_encode(p):
    P, S = load_package(), load_synth()
    planes = S.make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"], sigma=p["sigma"])
    if p.get("pattern") == "flatten_rows":
        planes = S.flatten_rows(planes, p["height"])
    enc = P.Encoder(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], p["bpp_den"], p["decomp_v"],
                    p["decomp_h"], p["quantization"], p["slice_height"], p["significance"],
                    rate_control_mode=p.get("rate_control_mode", 1), vertical_prediction=p.get("vertical_prediction", 0),
                    signs_handling=p.get("signs_handling", 0))
    rc, cs = enc.encode(planes)
    enc.close()
    return rc, cs


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(n for n in _G["small"] if not _G["small"][n].get("reference_error")))
def test_gpu_codestream_equals_reference_golden(name):
    print("golden", name, flush=True)
    rc, cs = _gpu_encode(_G["small"][name]["params"])
    assert rc == 0
    assert cs == open(os.path.join(_GOLDEN, name + ".jxs"), "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(_G["full"]))
def test_gpu_full_size_workloads_equal_reference(name):
    """First frame of every benchmark workload at full size (1080p .. 8K):
    the GPU codestream hashes to the reference encoder's."""
    print("full", name, flush=True)
    rc, cs = _gpu_encode(_G["full"][name]["params"])
    assert rc == 0
    assert hashlib.sha256(cs).hexdigest() == _G["full"][name]["codestream_sha256"]


@pytest.mark.gpu
def test_slice_packetization_packets():
    """slice_packetization_mode 1 (FinalStageProcess.c:126-213): get_packet
    hands out the frame header, then one packet per slice, back to back in
    the caller's buffer; only the last is marked last / releasable.  The
    packets put together are the oracle's codestream, every slice packet
    opens with its SLH (FF20, Lslh = 4, slice index) and the last one ends
    with EOC (FF11).  Two frames in a row."""
    W, H, D, F, bpp, V, Hh = 256, 128, 10, 3, 3, 2, 5
    S, O, P = load_synth(), load_oracle(), load_package()
    enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, 0, 16, 1, slice_packetization=1)
    slices = H // 16
    for frame in range(2):
        print("case slice packetization frame", frame, flush=True)
        planes = S.make_frame(W, H, D, F, frame=frame)
        e_ref, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, 0, 16, 1), planes)
        assert e_ref == 0
        out = np.zeros(enc.bytes_per_frame, dtype=np.uint8)
        assert enc.send(planes, out) == 0
        pkts = []
        for _ in range(slices + 1):
            rc, f = enc.get()
            assert rc == 0
            b = f.bitstream
            pkts.append((b.buffer - out.ctypes.data, b.used_size, b.last_packet_in_frame, b.ready_to_release,
                         f.image.ready_to_release))
            if b.last_packet_in_frame:
                break
        assert len(pkts) == slices + 1
        pos = 0
        for k, (off, size, last, rdy, img) in enumerate(pkts):
            assert off == pos
            pos += size
            final = k == slices
            assert (last, rdy, img) == ((1, 1, 1) if final else (0, 0, 0))
            if k > 0:
                assert bytes(out[off:off + 6]) == bytes([0xFF, 0x20, 0x00, 0x04, (k - 1) >> 8, (k - 1) & 0xFF])
        assert pos == len(ref)
        assert bytes(out[pos - 2:pos]) == b"\xff\x11"
        assert out[:pos].tobytes() == ref
    enc.close()


# (width, height, depth, format, bpp, decomp_v, decomp_h, slice_height)
RC0 = [
    (1920, 1080, 8, 3, 5, 2, 5, 16),
    (640, 360, 10, 4, 3, 1, 5, 32),
    (352, 288, 12, 2, 2, 2, 4, 16),
    (96, 64, 14, 21, 2, 1, 5, 16),
    (250, 130, 9, 3, 1, 0, 2, 8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", RC0, ids=lambda c: "x".join(str(v) for v in c))
def test_rate_control_mode0_bit_exact(cfg):
    """Rate control mode 0 (budget per precinct, padding kept in each
    precinct) against the oracle's restatement, which the reference's mode-0
    codestreams pin (tests/golden *_rc0)."""
    print("rc0", cfg, flush=True)
    W, H, D, F, bpp, V, Hh, sh = cfg
    S, O, P = load_synth(), load_oracle(), load_package()
    planes = S.make_frame(W, H, D, F, frame=3, sigma=6.0)
    e_ref, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, 0, sh, 1, 0), planes)
    enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, 0, sh, 1, rate_control_mode=0)
    rc, got = enc.encode(planes)
    enc.close()
    assert (rc == 0) == (e_ref == 0)
    if e_ref == 0:
        assert got == ref


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("cfg", RC0, ids=lambda c: "x".join(str(v) for v in c))
def test_rate_control_slice_modes_bit_exact(cfg, mode):
    """Rate control modes 2 and 3 (one Q and R per slice sizing the precinct
    budgets, 3 with a minimum precinct size) against the oracle, on even and
    on uneven content (the first 8 rows of every 16 flat, where mode 3 raises
    the small precincts), which the reference's codestreams pin
    (tests/golden *_rc2, *_rc3)."""
    print("rc%d" % mode, cfg, flush=True)
    W, H, D, F, bpp, V, Hh, sh = cfg
    S, O, P = load_synth(), load_oracle(), load_package()
    for planes in (S.make_frame(W, H, D, F, frame=4, sigma=6.0),
                   S.flatten_rows(S.make_frame(W, H, D, F, frame=5, sigma=8.0), H)):
        e_ref, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, 0, sh, 1, mode), planes)
        enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, 0, sh, 1, rate_control_mode=mode)
        rc, got = enc.encode(planes)
        enc.close()
        assert (rc == 0) == (e_ref == 0)
        if e_ref == 0:
            assert got == ref


@pytest.mark.gpu
def test_rate_control_mode_out_of_range_is_rejected():
    P = load_package()
    for mode in (4, 7):
        with pytest.raises(P.JxsError):
            P.Encoder(64, 32, 8, 3, 3, rate_control_mode=mode, verbose=0)
