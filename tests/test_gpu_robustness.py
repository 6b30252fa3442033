"""Damaged and truncated codestreams through the C-API decoder on the GPU,
against the reference decoder's verdicts (tests/golden/sweep_*.txt.gz; the
damage patterns of TestDecoder.cc:336-618).

For every variant: r1 = a decoder initialised with the variant decodes it,
r2 = a decoder initialised with the valid codestream decodes the variant; the
error codes equal the reference's, and where a decode succeeds its planes hash
to the reference decoder's planes.  The cumulative bit walk is replayed with
this decoder's verdicts and ends on the reference's buffer.

A decoder's initialisation reads only the header (the marker segments up to
the first slice marker), so r1 of a variant runs on a decoder initialised with
any buffer whose header bytes are the same: the valid codestream's decoder,
or the last one initialised from a damaged header.
"""
import ctypes
import sys

import numpy as np
import pytest

import sweep_util as SW
from jxs_util import load_package

P = load_package()


class _Api:
    """A C-API decoder handle over raw bytes (ctypes, no numpy copies kept)."""

    def __init__(self, buf):
        L = P.lib()
        self.api = P.DecoderApi()
        self.api.verbose = 0
        self.cfg = P.ImageConfig()
        arr = np.frombuffer(buf, dtype=np.uint8).copy()
        self.rc = ctypes.c_int32(L.svt_jpeg_xs_decoder_init(P.API_MAJOR, P.API_MINOR, ctypes.byref(self.api),
                                                            arr.ctypes.data, len(arr), ctypes.byref(self.cfg))).value
        if self.rc == 0:
            dt = np.uint8 if self.cfg.bit_depth <= 8 else np.uint16
            self.planes = [np.zeros((self.cfg.components[c].height, self.cfg.components[c].width), dtype=dt)
                           for c in range(self.cfg.components_num)]

    def decode(self, buf):
        L = P.lib()
        for p in self.planes:
            p.fill(0)
        arr = np.frombuffer(buf, dtype=np.uint8).copy()
        f = P.Frame()
        for c, p in enumerate(self.planes):
            f.image.data_yuv[c] = p.ctypes.data
            f.image.stride[c] = p.shape[1]
            f.image.alloc_size[c] = p.nbytes
        f.bitstream.buffer = arr.ctypes.data
        f.bitstream.allocation_size = len(arr)
        f.bitstream.used_size = len(arr)
        rc = ctypes.c_int32(L.svt_jpeg_xs_decoder_send_frame(ctypes.byref(self.api), ctypes.byref(f), 1)).value
        if rc == 0:
            out = P.Frame()
            rc = ctypes.c_int32(L.svt_jpeg_xs_decoder_get_frame(ctypes.byref(self.api), ctypes.byref(out), 1)).value
        return rc, (SW.plane_hash(p.tobytes() for p in self.planes) if rc == 0 else 0)

    def close(self):
        if self.rc == 0 and self.api.private_ptr:
            P.lib().svt_jpeg_xs_decoder_close(ctypes.byref(self.api))


def _header_bytes(buf):
    """The bytes a decoder's initialisation can read: the marker segments from
    SOC through the first slice marker, following each segment's length
    field; None when that walk leaves the buffer or meets an NLT segment (the
    only one whose parse does not follow its length field, ParseHeader.c:393)."""
    pos = 2
    while pos + 4 <= len(buf):
        if buf[pos:pos + 2] == b"\xff\x20":
            return buf[:pos + 2]
        if buf[pos:pos + 2] == b"\xff\x16":
            return None
        pos += 2 + ((buf[pos + 2] << 8) | buf[pos + 3])
    return None


class _Verdicts:
    """fs, r1, r2, h1, h2 of a buffer, as the reference harness measures them."""

    def __init__(self, valid):
        self.valid = valid
        self.main = _Api(valid)
        assert self.main.rc == 0
        assert self.main.decode(valid)[0] == 0
        self.key = _header_bytes(valid)
        self.other, self.other_key = None, None  # the last decoder initialised from another header

    def _r1(self, buf):
        key = _header_bytes(buf)
        if key is not None and key == self.key:
            return self.main.decode(buf)
        if key is not None and key == self.other_key:
            return self.other.decode(buf) if self.other.rc == 0 else (self.other.rc, 0)
        d = _Api(buf)
        if key is None:  # init may read the whole buffer: a decoder of its own
            r = d.decode(buf) if d.rc == 0 else (d.rc, 0)
            d.close()
            return r
        if self.other is not None:
            self.other.close()
        self.other, self.other_key = d, key
        return d.decode(buf) if d.rc == 0 else (d.rc, 0)

    def __call__(self, buf):
        L = P.lib()
        arr = np.frombuffer(buf, dtype=np.uint8).copy()
        size = ctypes.c_uint32(0)
        fs = ctypes.c_int32(L.svt_jpeg_xs_decoder_get_single_frame_size(arr.ctypes.data, len(arr), None,
                                                                        ctypes.byref(size), 0)).value
        r1, h1 = self._r1(buf)
        r2, h2 = self.main.decode(buf)
        return fs, r1, r2, h1, h2

    def close(self):
        self.main.close()
        if self.other is not None:
            self.other.close()


GROUPS = {"truncate_and_fill": "TZO", "bytes": "X", "bits": "B"}


@pytest.mark.gpu
@pytest.mark.parametrize("name", SW.SWEEPS)
@pytest.mark.parametrize("group", list(GROUPS))
def test_gpu_damaged_codestreams_match_reference(name, group):
    valid, rows, _ = SW.load(name)
    print("sweep", name, group, flush=True)
    V = _Verdicts(valid)
    bad, n = [], 0
    try:
        for kind, a, b, fs, r1, r2, h1, h2 in rows:
            if kind not in GROUPS[group]:
                continue
            n += 1
            got = V(SW.variant(valid, kind, a, b))
            if got != (fs, r1, r2, h1, h2):
                bad.append(((kind, a, b), (fs, r1, r2, h1, h2), got))
    finally:
        V.close()
    for x in bad[:20]:
        print(x, file=sys.stderr)
    assert n > 0 and not bad, "%d of %d variants differ" % (len(bad), n)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SW.SWEEPS)
def test_gpu_cumulative_bit_walk_matches_reference(name):
    """Corruprion_max_bits (TestDecoder.cc:574-604): every step's verdicts and
    the final buffer and its decode equal the reference's."""
    valid, rows, final = SW.load(name)
    print("walk", name, flush=True)
    V = _Verdicts(valid)
    try:
        steps, buf = SW.walk(valid, rows, V)
        ref = [r for r in rows if r[0] == "M"]
        bad = [(s[0], r[3:], s[1:]) for s, r in zip(steps, ref) if tuple(s[1:]) != tuple(r[3:])]
        for x in bad[:20]:
            print(x, file=sys.stderr)
        assert len(steps) == len(ref) and not bad, "%d of %d steps differ" % (len(bad), len(ref))
        assert SW.plane_hash([buf]) == final[0]
        assert V.main.decode(buf) == (final[1], final[2])
    finally:
        V.close()
