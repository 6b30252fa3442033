"""Damaged and truncated codestreams, host side (no GPU).

The fixtures tests/golden/sweep_*.txt.gz hold the reference decoder's verdict
on every damaged variant its own unit tests build (TestDecoder.cc:336-618) for
two codestreams.  Here: the fixtures satisfy the reference test's own
assertions (so the harness that wrote them followed those tests), and the
library's frame-size walk (svt_jpeg_xs_decoder_get_single_frame_size, a host
call) gives the reference's result on every variant.  The decodes themselves
are checked on the GPU (test_gpu_robustness.py).
"""
import ctypes

import numpy as np
import pytest

import sweep_util as SW
from jxs_util import load_package

P = load_package()
TOO_SHORT, INVALID = P.ERR_DEC_TOO_SHORT, P.ERR_DEC_INVALID_BITSTREAM


@pytest.mark.parametrize("name", SW.SWEEPS)
def test_fixture_holds_the_reference_tests_assertions(name):
    valid, rows, final = SW.load(name)
    n = len(valid)
    kinds = {}
    for kind, a, b, fs, r1, r2, h1, h2 in rows:
        kinds[kind] = kinds.get(kind, 0) + 1
        if kind == "T":  # TruncationSize / TruncationBufferCut: every size is too short
            assert fs == TOO_SHORT and r1 == TOO_SHORT and r2 == TOO_SHORT, (a, fs, r1, r2)
        if kind == "Z":  # TruncationBufferFillZero: the last decode's verdict is Invalid
            assert (r2 or r1) == INVALID, (a, r1, r2)
        if kind == "O":  # TruncationBufferFillOne
            assert (r2 or r1) in (TOO_SHORT, INVALID), (a, r1, r2)
        # a variant whose size walk fails never decodes (INVALID_ERROR_CODE_B)
        if fs != 0:
            assert r1 != 0 and r2 != 0, (kind, a, b)
        assert (h1 != 0) == (r1 == 0) and (h2 != 0) == (r2 == 0)
    assert kinds["T"] == kinds["Z"] == kinds["O"] == n - 1 and kinds["X"] == n
    assert kinds["B"] == kinds["M"] == 8 * n
    assert final[1] == 0  # the walk's buffer decodes (Corruprion_max_bits)


@pytest.mark.parametrize("name", SW.SWEEPS)
def test_frame_size_walk_equals_reference_on_every_variant(name):
    """svt_jpeg_xs_decoder_get_single_frame_size (ParseHeader.c:652-889
    restated) on every damaged variant: the reference's error code."""
    L = P.lib()
    valid, rows, _ = SW.load(name)
    bad = []
    for kind, a, b, fs, *_ in rows:
        if kind == "M":
            continue
        buf = np.frombuffer(SW.variant(valid, kind, a, b), dtype=np.uint8).copy()
        size = ctypes.c_uint32(0)
        e = ctypes.c_int32(L.svt_jpeg_xs_decoder_get_single_frame_size(buf.ctypes.data, len(buf), None,
                                                                       ctypes.byref(size), 0)).value
        if e != fs:
            bad.append((kind, a, b, fs, e))
    assert not bad, bad[:10]


def test_frame_size_of_the_valid_codestreams():
    L = P.lib()
    for name in SW.SWEEPS:
        valid, _, _ = SW.load(name)
        buf = np.frombuffer(valid + b"\0" * 7, dtype=np.uint8).copy()  # trailing bytes are not part of the frame
        size = ctypes.c_uint32(0)
        assert L.svt_jpeg_xs_decoder_get_single_frame_size(buf.ctypes.data, len(buf), None, ctypes.byref(size), 0) == 0
        assert size.value == len(valid)
