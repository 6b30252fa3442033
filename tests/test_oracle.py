"""CPU tests of the oracle (oracle/jxs_oracle.c): pinned against the reference
library's golden codestreams and decoded planes (tests/golden, written by
tests/golden/make_golden.py from the reference built from source), the
reference's quantisation known-answer vectors, the binary-search property of
its rate control, and closed forms of the 5/3 lifting."""
import hashlib
import json
import os

import numpy as np
import pytest

from jxs_util import ROOT, load_oracle, load_synth

GOLDEN = os.path.join(ROOT, "tests", "golden")
with open(os.path.join(GOLDEN, "golden.json")) as _f:
    G = json.load(_f)


def sha(b):
    return hashlib.sha256(b).hexdigest()


def _params(p):
    O = load_oracle()
    return O.make_params(p["width"], p["height"], p["depth"], p["format"], p["bpp_num"], p["bpp_den"], p["decomp_v"],
                         p["decomp_h"], p["quantization"], p["slice_height"], p["significance"],
                         p.get("rate_control_mode", 1), p.get("vertical_prediction", 0), p.get("signs_handling", 0))


def _planes(p):
    S = load_synth()
    planes = S.make_frame(p["width"], p["height"], p["depth"], p["format"], frame=p["frame"], sigma=p["sigma"])
    if p.get("pattern") == "flatten_rows":
        planes = S.flatten_rows(planes, p["height"])
    return planes


# ---------------------------------------------------------------- generator
@pytest.mark.parametrize("name", sorted(G["small"]))
def test_generator_is_reproducible(name):
    p = G["small"][name]["params"]
    S = load_synth()
    assert S.frame_sha256(_planes(p)) == G["small"][name]["input_sha256"]


# ---------------------------------------------------------------- known answers
@pytest.mark.parametrize("k", range(5))
def test_quantization_known_answers(k):
    """TestQuant.cc:27-136: five vectors (deadzone gtli 0/3, uniform 0/4/6)."""
    O = load_oracle()
    kat = G["quant_kat"]
    t = kat["tests"][k]
    n = t["buf_len"]
    got = O.quantize(kat["input"][:n], kat["gclis"], t["gtli"], t["uniform"])
    assert got.tolist() == t["expected"][:n]


def test_quantization_planted_mistake_is_caught():
    """A dropped rounding term in the uniform quantiser fails the vector."""
    kat = G["quant_kat"]
    t = kat["tests"][3]
    n, gt = t["buf_len"], t["gtli"]
    wrong = []
    for i, v in enumerate(kat["input"][:n]):
        gc = kat["gclis"][i // 4]
        if gc <= gt:
            wrong.append(0)
            continue
        m = v & 0x7FFF
        d = ((m << (gc - gt + 1)) - m) >> (gc + 1)  # rounding term (1 << gcli) dropped
        o = (d << gt) & 0xFFFF
        wrong.append(o | (v & 0x8000) if o else 0)
    assert wrong != t["expected"][:n]


@pytest.mark.parametrize("n", [3, 7, 16, 29, 32])
@pytest.mark.parametrize("below", [0, 1])
def test_binary_search_equals_linear_search(n, below):
    """BinarySearch.c against a linear scan for every threshold of monotone
    tables, with and without trailing out-of-range entries (the property the
    reference checks exhaustively in TestReateControl.cc and asserts in
    RateControl.c:1301-1318)."""
    O = load_oracle()
    for empty_from in (n + 1, n, n - 1):
        for thr in range(-1, n + 2):
            if below == 0:
                # smallest i with f(i) <= thr, f decreasing: f(i) = n - i
                def verdict(i):
                    if i >= empty_from:
                        return 3
                    return 1 if (n - i) > thr else 2
                cands = [i for i in range(0, n + 1) if i < empty_from and n - i <= thr]
                expect = min(cands) if cands else None
            else:
                # largest i with f(i) <= thr, f increasing: f(i) = i
                def verdict(i):
                    if i >= empty_from:
                        return 3
                    return 1 if i <= thr else 2
                cands = [i for i in range(0, n + 1) if i < empty_from and i <= thr]
                expect = max(cands) if cands else None
            step = 6 if (6 <= n and (below == 0 or 6 < n)) else 0
            got = O.binary_search(0, n, below, step, verdict)
            assert got == expect, (n, below, empty_from, thr, got, expect)


# ---------------------------------------------------------------- 5/3 lifting
@pytest.mark.parametrize("n", list(range(2, 41)))
def test_lifting_perfect_reconstruction(n):
    O = load_oracle()
    rng = np.random.default_rng(n)
    x = rng.integers(-(1 << 19), 1 << 19, size=n).astype(np.int32)
    lf, hf = O.fwd53(x)
    assert np.array_equal(O.inv53(lf, hf), x)


@pytest.mark.parametrize("n", [2, 3, 8, 9, 31])
def test_lifting_closed_forms(n):
    """A constant signal has no high band and a low band equal to it; a ramp
    x[i] = i has zero high band inside and the low band x[2i] + 0 except at
    the even-length end where the one-sided difference is 1."""
    O = load_oracle()
    lf, hf = O.fwd53(np.full(n, 37, dtype=np.int32))
    assert (hf == 0).all() and (lf == 37).all()
    lf, hf = O.fwd53(np.arange(n, dtype=np.int32))
    inner = hf[:-1] if n % 2 == 0 else hf
    assert (inner == 0).all()
    if n % 2 == 0:
        assert hf[-1] == 1


# ---------------------------------------------------------------- codestreams
@pytest.mark.parametrize("name", sorted(G["small"]))
def test_oracle_codestream_matches_reference(name):
    O = load_oracle()
    entry = G["small"][name]
    p = entry["params"]
    e, cs = O.encode(_params(p), _planes(p))
    if entry.get("reference_error"):
        assert e != 0
        return
    assert e == 0
    ref = open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    assert len(cs) == len(ref)
    if cs != ref:
        a, b = np.frombuffer(cs, np.uint8), np.frombuffer(ref, np.uint8)
        pytest.fail("first differing byte %d of %d" % (int(np.nonzero(a != b)[0][0]), len(ref)))


@pytest.mark.parametrize("name", sorted(G["small"]))
def test_oracle_decode_matches_reference(name):
    O = load_oracle()
    entry = G["small"][name]
    if entry.get("reference_error"):
        return
    ref = open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    e, planes = O.decode(ref)
    assert e == 0
    assert sha(b"".join(p.tobytes() for p in planes)) == entry["decoded_sha256"]


@pytest.mark.parametrize("name", sorted(G["full"]))
def test_oracle_full_size_workloads_match_reference(name):
    """First frame of every benchmark workload at full size: input planes,
    codestream and decoded planes against the reference's hashes."""
    O, S = load_oracle(), load_synth()
    entry = G["full"][name]
    p = entry["params"]
    planes = _planes(p)
    assert S.frame_sha256(planes) == entry["input_sha256"]
    e, cs = O.encode(_params(p), planes)
    assert e == 0
    assert sha(cs) == entry["codestream_sha256"]
    e, dec = O.decode(cs)
    assert e == 0
    assert sha(b"".join(x.tobytes() for x in dec)) == entry["decoded_sha256"]


def test_reference_sample_codestream_decodes_to_reference_planes():
    """The 16x16 4:2:2 codestream of the reference's unit tests
    (SampleFramesData.cc:8-33, decoded by TestDecoder.cc) uses vertical
    prediction: the oracle decodes it to the reference decoder's planes."""
    O = load_oracle()
    cs = open(os.path.join(GOLDEN, "sample_16x16_422.jxs"), "rb").read()
    assert sha(cs) == G["sample_16x16_422"]["codestream_sha256"]
    info = O.ImageInfo()
    import ctypes
    buf = np.frombuffer(cs, np.uint8).copy()
    assert O.lib().jxso_probe(buf.ctypes.data, len(buf), ctypes.byref(info)) == 0
    assert (info.width, info.height, info.depth, info.comps) == (16, 16, 8, 3)
    assert (info.comp_w[1], info.comp_h[1]) == (8, 16)
    e, planes = O.decode(cs)
    assert e == 0
    got = b"".join(p.tobytes() for p in planes)
    assert got == open(os.path.join(GOLDEN, "sample_16x16_422.yuv"), "rb").read()
    assert sha(got) == G["sample_16x16_422"]["decoded_sha256"]


def _clear_vpred_bits(cs):
    """The codestream with bit 0 (vertical prediction) of every band's
    D[p,b] cleared in every precinct header."""
    cs = bytearray(cs)
    wgt = cs.index(b"\xff\x14")
    nb = ((cs[wgt + 2] << 8 | cs[wgt + 3]) - 2) // 2
    pos = cs.index(b"\xff\x20")
    hdr = (40 + 2 * nb + 7) // 8
    while cs[pos:pos + 2] != b"\xff\x11":
        if cs[pos:pos + 2] == b"\xff\x20":
            pos += 6
            continue
        lprc = cs[pos] << 16 | cs[pos + 1] << 8 | cs[pos + 2]
        for k in range(nb):
            bit = 40 + 2 * k + 1  # low bit of the k-th 2-bit field
            cs[pos + bit // 8] &= ~(0x80 >> (bit % 8)) & 0xFF
        pos += hdr + lprc
    return bytes(cs)


def test_vertical_prediction_planted_mistake_is_caught():
    """Planted mistake: read as zero-predicted (the D[p,b] vertical bits
    cleared), the sample's bit-plane counts no longer give the reference
    decoder's planes -- the vertical-prediction decode is what pins them."""
    O = load_oracle()
    cs = open(os.path.join(GOLDEN, "sample_16x16_422.jxs"), "rb").read()
    bad = _clear_vpred_bits(cs)
    assert bad != cs
    e, planes = O.decode(bad)
    ref = open(os.path.join(GOLDEN, "sample_16x16_422.yuv"), "rb").read()
    assert e != 0 or b"".join(p.tobytes() for p in planes) != ref


@pytest.mark.parametrize("proxy", [1, 2])
@pytest.mark.parametrize("name", sorted(n for n in G["small"] if "decoded_half_sha256" in G["small"][n]))
def test_oracle_proxy_decode_matches_reference(name, proxy):
    """Decoding at half / quarter resolution (the reference decoder's
    proxy_mode, Pi.c:567-626) against the reference decoder's planes; where
    the reference refuses the mode (fewer levels than it drops) the oracle
    refuses it too."""
    O = load_oracle()
    entry = G["small"][name]
    if entry.get("reference_error"):
        return
    ref = open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    want = entry["decoded_half_sha256" if proxy == 1 else "decoded_quarter_sha256"]
    e, planes = O.decode(ref, proxy)
    if want is None:
        assert e != 0
        return
    assert e == 0
    p = entry["params"]
    assert planes[0].shape[1] == (p["width"] + (1 << proxy) - 1) >> proxy
    assert sha(b"".join(pl.tobytes() for pl in planes)) == want
