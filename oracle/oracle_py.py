"""ctypes front end of the C oracle (oracle/jxs_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libjxs_oracle.so")
REF_ENC = os.path.join(HERE, "_ref", "ref_enc")
REF_DEC = os.path.join(HERE, "_ref", "ref_dec")


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "depth", "format", "bpp_num", "bpp_den", "decomp_v", "decomp_h",
        "quantization", "slice_height", "significance", "rc_mode", "vpred", "signs")]


class ImageInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("comps", ctypes.c_uint32), ("comp_w", ctypes.c_uint32 * 4), ("comp_h", ctypes.c_uint32 * 4)]


class BSearch(ctypes.Structure):
    _fields_ = [("beg", ctypes.c_int32), ("end", ctypes.c_int32), ("step", ctypes.c_int32), ("best", ctypes.c_int32),
                ("last", ctypes.c_int32), ("below", ctypes.c_uint8)]


_lib = None


def quantize(values, gclis, gtli, uniform):
    buf = np.array(values, dtype=np.uint16)
    g = np.array(gclis, dtype=np.uint8)
    lib().jxso_quantize(buf.ctypes.data, len(buf), g.ctypes.data, gtli, 1 if uniform else 0)
    return buf


def fwd53(x):
    x = np.ascontiguousarray(x, dtype=np.int32)
    n = len(x)
    lf = np.zeros((n + 1) // 2, dtype=np.int32)
    hf = np.zeros(n // 2, dtype=np.int32)
    lib().jxso_fwd53(x.ctypes.data, lf.ctypes.data, hf.ctypes.data, n)
    return lf, hf


def inv53(lf, hf):
    lf = np.ascontiguousarray(lf, dtype=np.int32)
    hf = np.ascontiguousarray(hf, dtype=np.int32)
    x = np.zeros(len(lf) + len(hf), dtype=np.int32)
    lib().jxso_inv53(lf.ctypes.data, hf.ctypes.data, x.ctypes.data, len(x))
    return x


def binary_search(begin, end, below, step, verdict):
    """Runs the rate control's binary search; verdict(i) -> 1 too small,
    2 too big, 3 out of range.  Returns the found index or None."""
    b = BSearch()
    lib().jxso_bsearch_init(ctypes.byref(b), begin, end, below, step)
    nxt = ctypes.c_uint32(0)
    res = 0
    while True:
        r = lib().jxso_bsearch_next(ctypes.byref(b), res, ctypes.byref(nxt))
        if r == 1:
            return nxt.value
        if r == 2:
            return None
        res = verdict(nxt.value)


def build():
    """Compile the oracle with gcc into oracle/_build (always rebuilds)."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    src = os.path.join(HERE, "jxs_oracle.c")
    tmp = LIB_PATH + ".tmp"
    cmd = "gcc -O2 -fPIC -shared -o '%s' '%s'" % (tmp, src)
    if os.system(cmd) != 0:
        raise RuntimeError("oracle build failed: " + cmd)
    os.replace(tmp, LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        src = os.path.join(HERE, "jxs_oracle.c")
        if not os.path.exists(LIB_PATH) or (os.path.exists(src) and os.path.getmtime(src) > os.path.getmtime(LIB_PATH)):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.jxso_default_params.argtypes = [ctypes.POINTER(Params)]
        L.jxso_codestream_bytes.argtypes = [ctypes.POINTER(Params)]
        L.jxso_codestream_bytes.restype = ctypes.c_uint32
        L.jxso_encode.argtypes = [ctypes.POINTER(Params), ctypes.c_void_p * 4, ctypes.c_uint32 * 4, ctypes.c_void_p,
                                  ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.jxso_encode.restype = ctypes.c_int32
        L.jxso_encode_dump.argtypes = [ctypes.POINTER(Params), ctypes.c_void_p * 4, ctypes.c_uint32 * 4, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32 * 3]
        L.jxso_encode_dump.restype = ctypes.c_int32
        L.jxso_probe.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ImageInfo)]
        L.jxso_probe.restype = ctypes.c_int32
        L.jxso_decode.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p * 4, ctypes.POINTER(ImageInfo)]
        L.jxso_decode.restype = ctypes.c_int32
        L.jxso_quantize.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint8, ctypes.c_uint32]
        L.jxso_quantize.restype = None
        L.jxso_fwd53.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        L.jxso_fwd53.restype = None
        L.jxso_inv53.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        L.jxso_inv53.restype = None
        L.jxso_bsearch_init.argtypes = [ctypes.POINTER(BSearch), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint8,
                                        ctypes.c_uint32]
        L.jxso_bsearch_init.restype = None
        L.jxso_bsearch_next.argtypes = [ctypes.POINTER(BSearch), ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32)]
        L.jxso_bsearch_next.restype = ctypes.c_int32
        _lib = L
    return _lib


def make_params(width, height, depth, fmt, bpp_num, bpp_den=1, decomp_v=2, decomp_h=5, quantization=0,
                slice_height=16, significance=1, rc_mode=1, vpred=0, signs=0):
    p = Params()
    lib().jxso_default_params(ctypes.byref(p))
    p.width, p.height, p.depth, p.format = width, height, depth, fmt
    p.bpp_num, p.bpp_den, p.decomp_v, p.decomp_h = bpp_num, bpp_den, decomp_v, decomp_h
    p.quantization, p.slice_height, p.significance = quantization, slice_height, significance
    p.rc_mode = rc_mode
    p.vpred, p.signs = vpred, signs
    return p


def _plane_args(planes, fmt):
    ptrs = (ctypes.c_void_p * 4)()
    strides = (ctypes.c_uint32 * 4)()
    for i, pl in enumerate(planes):
        assert pl.flags["C_CONTIGUOUS"]
        ptrs[i] = pl.ctypes.data
        strides[i] = pl.shape[1]
    return ptrs, strides


def codestream_bytes(p):
    return lib().jxso_codestream_bytes(ctypes.byref(p))


def encode(p, planes):
    """Returns (error_code, bytes)."""
    n = codestream_bytes(p)
    if n == 0:
        return (int(np.int32(np.uint32(0x80001005))), b"")
    out = np.zeros(n, dtype=np.uint8)
    ptrs, strides = _plane_args(planes, p.format)
    used = ctypes.c_uint32(0)
    e = lib().jxso_encode(ctypes.byref(p), ptrs, strides, out.ctypes.data, n, ctypes.byref(used))
    return e, out.tobytes()


def encode_dump(p, planes, total_coeffs, total_gcli, rows):
    coeff = np.zeros(total_coeffs, dtype=np.uint16)
    gcli = np.zeros(total_gcli, dtype=np.uint8)
    qr = np.zeros(2 * rows, dtype=np.uint8)
    ptrs, strides = _plane_args(planes, p.format)
    sizes = (ctypes.c_uint32 * 3)()
    e = lib().jxso_encode_dump(ctypes.byref(p), ptrs, strides, coeff.ctypes.data, gcli.ctypes.data, qr.ctypes.data, sizes)
    return e, coeff[:sizes[0]], gcli[:sizes[1]], qr[:sizes[2]]


def decode(cs, proxy=0):
    """Returns (error_code, [planes]); proxy 1 / 2 decodes at half / quarter resolution."""
    buf = np.frombuffer(cs, dtype=np.uint8).copy()
    info = ImageInfo()
    e = lib().jxso_probe_proxy(ctypes.c_void_p(buf.ctypes.data), ctypes.c_uint32(len(buf)), ctypes.c_uint32(proxy),
                               ctypes.byref(info))
    if e:
        return e, []
    dt = np.uint8 if info.depth <= 8 else np.uint16
    planes = [np.zeros((info.comp_h[c], info.comp_w[c]), dtype=dt) for c in range(info.comps)]
    ptrs = (ctypes.c_void_p * 4)()
    for c, pl in enumerate(planes):
        ptrs[c] = pl.ctypes.data
    e = lib().jxso_decode_proxy(ctypes.c_void_p(buf.ctypes.data), ctypes.c_uint32(len(buf)), ctypes.c_uint32(proxy),
                                ptrs, ctypes.byref(info))
    return e, planes


def have_reference():
    return os.path.exists(REF_ENC) and os.path.exists(REF_DEC)
