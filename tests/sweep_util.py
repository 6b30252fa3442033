"""Damaged-codestream sweeps: the damage patterns of the reference's decoder
unit tests (tests/UnitTests/TestDecoder.cc:336-618) and the reference
decoder's verdict on each variant, read from tests/golden/sweep_<case>.txt.gz
(written by tests/golden/make_golden.py through oracle/ref_harness.c).

A fixture line is "kind a b fs r1 r2 h1 h2":
  kind T  the first a bytes (the frame cut short, TruncationSize / Cut)
       Z  bytes from a on set to 0, full size (TruncationBufferFillZero)
       O  bytes from a on set to 1, full size (TruncationBufferFillOne)
       X  byte a inverted (Corruprion_xor_bytes)
       B  bit 7 - b of byte a flipped (Corruprion_xor_bits)
       M  the cumulative walk (Corruprion_max_bits): bit 7 - b of byte a
          flipped into a buffer that keeps every earlier flip after which it
          still decoded (fs and r2 both 0), the others undone
  fs  svt_jpeg_xs_decoder_get_single_frame_size of the variant
  r1  a decoder initialised with the variant decoding it
  r2  a decoder initialised with the valid codestream decoding it, then the variant
  h1 / h2  the decoded planes' hash (plane_hash) for r1 / r2 == 0, else 0
The last line "F hash r2 h2" is the walk's final buffer and its decode.
"""
import gzip
import os
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
SWEEPS = ["sample_16x16_422", "s_64x32_422_8b_bpp5_v2"]


def plane_hash(chunks):
    """crc32 << 32 | adler32 of the byte strings in order."""
    c, a = 0, 1
    for b in chunks:
        c = zlib.crc32(b, c)
        a = zlib.adler32(b, a)
    return (c << 32) | a


def load(name):
    valid = open(os.path.join(GOLDEN, name + ".jxs"), "rb").read()
    rows, final = [], None
    with gzip.open(os.path.join(GOLDEN, "sweep_%s.txt.gz" % name), "rt") as f:
        for ln in f:
            p = ln.split()
            if not p:
                continue
            if p[0] == "F":
                final = (int(p[1], 16), int(p[2]), int(p[3], 16))
                continue
            rows.append((p[0], int(p[1]), int(p[2]), int(p[3]), int(p[4]), int(p[5]), int(p[6], 16), int(p[7], 16)))
    return valid, rows, final


def variant(valid, kind, a, b):
    """The damaged bytes of a T / Z / O / X / B line."""
    n = len(valid)
    if kind == "T":
        return valid[:a]
    buf = bytearray(valid)
    if kind == "Z":
        buf[a:] = bytes(n - a)
    elif kind == "O":
        buf[a:] = b"\x01" * (n - a)
    elif kind == "X":
        buf[a] ^= 0xFF
    elif kind == "B":
        buf[a] ^= 1 << (7 - b)
    else:
        raise ValueError(kind)
    return bytes(buf)


def walk(valid, rows, verdict):
    """Replays the cumulative walk: verdict(buf) -> (fs, r1, r2, h1, h2) for
    each step; returns the per-step results and the final buffer."""
    buf = bytearray(valid)
    out = []
    for (kind, a, b, *_rest) in rows:
        if kind != "M":
            continue
        buf[a] ^= 1 << (7 - b)
        v = verdict(bytes(buf))
        out.append(((kind, a, b),) + tuple(v))
        if not (v[0] == 0 and v[2] == 0):
            buf[a] ^= 1 << (7 - b)
    return out, bytes(buf)
